"""Throughput of the transport-and-scoring path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--histories N] [--cpu-sample S]

Workload (one step): the heterogeneous slab phantom - 64x64x96 voxels at 2 mm, water
with a lung slab (0.26 g/cc, depth 3-7 cm) and a bone slab (1.85 g/cc, 9-11 cm), a
10x10 cm parallel field of 6 MV-like spectral photons, 1e7 histories in 20 batches,
electron transport with Goudsmit-Saunderson multiple scattering, analytic water
cross-sections - one complete ``HipEngine.run``: generation, transport, per-batch
float64 fold and finalize, with the phantom, tables and GS grid already resident in
HBM. Data are synthetic. Prints one JSON line.

Before the warmup, ``pyradmc_amd.autotune`` picks the chunk size and the rounds' launch
shape for this workload (untimed; no result depends on them; ``--no-autotune`` keeps
the defaults). The timed steps run with per-kernel HIP events, so their rounds launch
kernel by kernel; ``unprofiled`` times the same steps without events, where each group
of four rounds replays a captured HIP graph when the tuned shape has graphs on.

``roofline`` is for the dominant kernel (the electron round): algorithmic bytes per
launch - 48 B per particle popped, 48 B per queue push, 5 B per voxel fetch (float32
density + uint8 material), 16 B per int64 tally atomic; tables and the GS grid are
cache-resident and not counted - divided by that kernel's mean launch time from HIP
events recorded on the engine's own stream during the timed steps. ``traffic`` is read
from ``profiles/traffic_electron_round.json``: HBM bytes per launch from two PMC runs of
this workload (FETCH_SIZE, WRITE_SIZE), each scaled by what the counter reads for a
1 GiB device copy (``--copy-probe-mib 1024`` runs that copy first); null if absent.

``cpu_baseline`` times the CPU restatement of the reference algorithm (``oracle/``,
double precision, one stream per history, pthreads) on this host's cores for a
bounded sample of the same workload (and the quick-start pencil, ``pencil``).

Further legs, reported beside the throughput line: ``pencil`` (the reference's
quick-start pencil in both voxel readings), ``dij`` (the batched uneven 100-beamlet
Dij), ``denoised_dij`` (denoised against plain Monte Carlo at the same error, with the
denoiser's MFMA roofline), ``full_plan`` (five gantry beams, 1125 columns, and the
1000-vector planning product) and ``full_plan_denoised`` (the same comparison as
``denoised_dij`` on the full plan's gantry-0 beam).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0  # dense bf16 MFMA, MI355X


def _algorithmic_bytes(c: dict, prefix: str) -> float:
    return (48.0 * c[f"{prefix}_pops"] + 48.0 * c[f"{prefix}_pushes"]
            + 5.0 * c[f"{prefix}_voxels"] + 16.0 * c[f"{prefix}_deposits"])


EVENT_KEYS = ("photon_steps", "electron_steps", "photon_voxels", "electron_voxels", "photon_deposits",
              "electron_deposits", "photon_pushes", "electron_pushes")


def _events(counters: dict, histories: float) -> dict:
    """Per-history event counts from a run's ledger counters: the denominators of the
    algorithmic bytes (5 B per voxel fetch, 16 B per tally atomic, 48 B per push and pop)."""
    return {k: round(counters.get(k, 0) / histories, 3) for k in EVENT_KEYS}


def _per_beamlet_sigma(d) -> np.ndarray:
    """Median over each column's voxels with D >= 0.5 max of sigma/D."""
    out = np.zeros(d.n_beamlets)
    for j in range(d.n_beamlets):
        lo, hi = int(d.indptr[j]), int(d.indptr[j + 1])
        dose = np.asarray(d.dose[lo:hi], dtype=np.float64)
        sig = np.asarray(d.sigma[lo:hi], dtype=np.float64)
        keep = dose >= 0.5 * dose.max() if hi > lo else np.zeros(0, bool)
        out[j] = float(np.median(sig[keep] / dose[keep])) if keep.any() else np.nan
    return out


def _shaped_engine(grid, xs, launch: dict | None):
    """An engine with the slab-phantom step's launch shape (the chunk size stays
    automatic: it follows each workload's size)."""
    from pyradmc_amd import HipEngine

    engine = HipEngine(grid, xs)
    if launch:
        engine.set_launch(launch["round_lanes_per_cu"], launch["refill_block"], launch["push_chunk"],
                          launch["use_graphs"])
    return engine


def _pencil_leg(repeats: int = 7) -> dict:
    """The reference's quick start: a 6 MeV pencil into 16^3 water, 80,000 histories in 10
    batches, in both readings (1 cm voxels entering at (8, 8, -1) cm; 1 mm voxels at
    (0.8, 0.8, -0.1) cm). Median wall of whole runs, dose and sigma on the host; the
    workload is too small to fill the GPU, so this measures launch and host overhead."""
    import statistics

    from pyradmc_amd import HipEngine, workloads

    out = {}
    for units in ("cm", "mm"):
        grid, xs, src, kw = workloads.small_pencil(units)
        engine = HipEngine(grid, xs)
        engine.run(src, kw["n_histories"], kw["n_batches"], seed=1)  # warm
        walls = []
        for k in range(repeats):
            started = time.perf_counter()
            engine.run(src, kw["n_histories"], kw["n_batches"], seed=kw["seed"] + k)
            walls.append(time.perf_counter() - started)
        engine.close()
        wall = statistics.median(walls)
        out[units] = {"wall_s": round(wall, 5), "histories_per_s": round(kw["n_histories"] / wall, 1)}
    out["workload"] = "6 MeV pencil, 16^3 water (1 cm and 1 mm voxels), 80,000 histories in 10 batches"
    return out


def _head_leg(n_presolve: int = 8_000_000, histories: int = 8_000_000, repeats: int = 3) -> dict:
    """The treatment head on one field: a 6 MV-like Gaussian-spot source through two jaw
    pairs and a 40-pair rounded-tip MLC onto 64^3 water at 5 mm. Times (median of
    ``repeats``, whole calls on the host clock): the device pre-solve (forced first
    Compton, air column) with its primaries drawn in the kernel, and with them sampled on
    the host; transport of its exit population left in HBM; the same field as a
    collimated in-kernel source and open; and, for contrast, the copied-back population
    on the host-sampled route."""
    import statistics
    import warnings

    from pyradmc_amd import (MLC, AirColumn, AnalyticCrossSections, BeamFrame, BeamLimitingStack,
                             CollimatedSource, GaussianSpotBeamSource, HipEngine, JawPair, VoxelGrid,
                             bremsstrahlung_6mv_like, presolve_head)

    grid = VoxelGrid.uniform_water((64, 64, 64), (0.5, 0.5, 0.5))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    focal = (16.0, 16.0, -100.0)
    leaves = np.linspace(-2.0, 2.0, 41)
    mid = 0.5 * (leaves[:-1] + leaves[1:])
    # the analytic source is water only: the devices are water (material 0) at tungsten's density
    stack = BeamLimitingStack(BeamFrame(focal), (
        JawPair("v", 20.0, 27.8, -2.2, 2.2, 0, 19.3, focused=True),
        JawPair("u", 28.0, 35.8, -2.2, 2.2, 0, 19.3, focused=True),
        MLC(40.0, 46.5, tuple(leaves), tuple(-1.2 - 0.6 * np.cos(mid)), tuple(0.9 + 0.7 * np.sin(mid) ** 2), 8.0,
            0, 18.0, focused_sides=True),
    ))
    src = GaussianSpotBeamSource(bremsstrahlung_6mv_like(), focal, (16.0, 16.0, -85.0), 3.0, 3.0, 0.15, 0.15)
    air = AirColumn(50.0, 95.0, material=0)  # water at air's density (the analytic source is water only)
    engine = HipEngine(grid, xs)

    def median_wall(fn):
        walls = []
        for _ in range(repeats):
            started = time.perf_counter()
            fn()
            walls.append(time.perf_counter() - started)
        return statistics.median(walls)

    class HostSampledSpot(GaussianSpotBeamSource):
        """A subclass: the pre-solve samples its primaries on the host."""

    host_src = HostSampledSpot(bremsstrahlung_6mv_like(), focal, (16.0, 16.0, -85.0), 3.0, 3.0, 0.15, 0.15)
    kw = dict(cross_sections=xs, exit_z=97.0, stack=stack, air=air, engine=engine, resident=True)
    presolve_head(src, n_histories=100_000, seed=1, **kw)  # warm
    presolve_wall = median_wall(lambda: presolve_head(src, n_histories=n_presolve, seed=2, **kw))
    host_presolve_wall = median_wall(lambda: presolve_head(host_src, n_histories=n_presolve, seed=2, **kw))
    population = presolve_head(src, n_histories=n_presolve, seed=3, **kw)
    engine.run(population, 1_000_000, 10, seed=4)  # warm
    resident = median_wall(lambda: engine.run(population, histories, 10, seed=5))
    collimated_src = CollimatedSource(src, stack, xs)
    engine.run(collimated_src, 1_000_000, 10, seed=6)
    collimated = median_wall(lambda: engine.run(collimated_src, histories, 10, seed=7))
    engine.run(src, 1_000_000, 10, seed=8)
    open_field = median_wall(lambda: engine.run(src, histories, 10, seed=9))
    host_n = 2_000_000
    in_memory = population.to_phase_space()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = median_wall(lambda: engine.run(in_memory, host_n, 10, seed=10))
    engine.close()
    return {
        "workload": "6 MV-like Gaussian spot (0.15 cm) through two focused jaw pairs and a 40-pair "
                    "rounded-tip MLC (water at tungsten density), air column; 64^3 water at 5 mm",
        "presolve": {"primaries": n_presolve, "exit_particles": len(population), "wall_s": round(presolve_wall, 4),
                     "primaries_per_s": round(n_presolve / presolve_wall, 1),
                     "host_sampled_primaries_wall_s": round(host_presolve_wall, 4)},
        "histories_per_s": {"resident_phase_space": round(histories / resident, 1),
                            "collimated_in_kernel": round(histories / collimated, 1),
                            "open_spot_field": round(histories / open_field, 1),
                            "host_sampled_phase_space": round(host_n / host, 1)},
        "histories": histories,
    }


def _dij_leg(median: float, launch: dict | None = None) -> dict:
    """The batched uneven Dij: 128x128x96 CT-like phantom at 2.5 mm, 100 spectral
    beamlets of 5x5 mm from 100 cm SAD, lognormal per-beamlet counts, truncation 1e-4,
    float32 CSC values - one complete run_dij, timed wall to wall (CSC on the host)."""
    from pyradmc_amd import workloads

    grid, xs, src, kw = workloads.ct_like_dij(median=median)
    counts = kw["n_histories_per_beamlet"]
    engine = _shaped_engine(grid, xs, launch)
    if launch is None:  # launch shape for this workload (untimed; results do not depend on it)
        engine.autotune_dij(src, counts, kw["n_batches"], seed=999, repeats=2, use_cache=False,
                            truncation=kw["truncation"], value_dtype=np.float32)
    small = np.maximum(kw["n_batches"], counts // 100 // kw["n_batches"] * kw["n_batches"])
    engine.run_dij(src, small, kw["n_batches"], seed=1, truncation=kw["truncation"])  # warm
    started = time.perf_counter()
    d = engine.run_dij(src, counts, kw["n_batches"], seed=kw["seed"], truncation=kw["truncation"],
                       value_dtype=np.float32)
    wall = time.perf_counter() - started
    sig = _per_beamlet_sigma(d)
    med = float(np.nanmedian(sig))
    # Sigma scales as N^-1/2 at a fixed layout: every count scaled so that the worst
    # (90th percentile) column reaches 2 percent, and that Dij run and timed.
    scale = (float(np.nanpercentile(sig, 90)) / 0.02) ** 2
    counts2 = np.maximum(kw["n_batches"], np.ceil(counts * scale / kw["n_batches"]).astype(np.int64)
                         * kw["n_batches"])
    started = time.perf_counter()
    d2 = engine.run_dij(src, counts2, kw["n_batches"], seed=kw["seed"] + 1, truncation=kw["truncation"],
                        value_dtype=np.float32)
    wall2 = time.perf_counter() - started
    sig2 = _per_beamlet_sigma(d2)
    at_2pct = {"count_scale": round(scale, 3), "histories": int(counts2.sum()), "wall_s": round(wall2, 4),
               "histories_per_s": round(float(counts2.sum()) / wall2, 1),
               "per_beamlet_sigma_median": round(float(np.nanmedian(sig2)), 5),
               "per_beamlet_sigma_p90": round(float(np.nanpercentile(sig2, 90)), 5)}
    launch_used = dict(engine.launch, chunk_size=engine.chunk_size or "auto")
    engine.close()
    del d2
    return {
        "workload": "CT-like 128x128x96 @ 2.5 mm, 100 beamlets 5x5 mm, SAD 100, lognormal counts "
                    f"(median {median:g}, sigma 0.8), truncation 1e-4, f32 CSC",
        "histories": int(counts.sum()),
        "wall_s": round(wall, 4),
        "histories_per_s": round(float(counts.sum()) / wall, 1),
        "nnz": int(d.indptr[-1]),
        "per_beamlet_sigma_median": round(med, 5),
        "per_beamlet_sigma_p90": round(float(np.nanpercentile(sig, 90)), 5),
        "at_2pct": at_2pct,
        "events_per_history": _events(d.counters or {}, float(counts.sum())),
        "seconds": _phase_seconds(d.counters or {}),
        "launch": launch_used,
    }


def _window_errors(d, ref, source, grid) -> np.ndarray:
    """Per-column dose-weighted RMS error of Dij d against Dij ref inside each
    beamlet's beam-aligned window."""
    from pyradmc_amd import denoise

    spacing = float(grid.spacing[0])
    shape = tuple(int(n) for n in grid.shape)
    out = np.zeros(source.n_beamlets)

    def window(dij, j, win):
        x, y, z, dose, _ = denoise._window_entries(dij, j, win, shape)
        a = np.zeros((denoise.WINDOW_XY, denoise.WINDOW_XY, shape[2]))
        a[x, y, z] = dose
        return a

    for j in range(source.n_beamlets):
        cx, cy, _ = source.centers[j]
        win = denoise.beam_window((int(cx / spacing), int(cy / spacing)), shape)
        out[j] = denoise.dose_weighted_rms(window(d, j, win), window(ref, j, win))
    return out


def _denoiser_roofline(model, shape: tuple[int, int, int, int], repeats: int = 5) -> dict:
    """The denoiser's forward pass on a batch of windows (the builder's inference
    batch) against the bf16 MFMA peak: algorithmic FLOPs over the mean time between HIP
    events on the stream its kernels run on."""
    import torch

    from pyradmc_amd import denoise

    rng = np.random.default_rng(5)
    noisy = rng.random(shape, dtype=np.float32)
    dens = rng.uniform(0.1, 1.0, shape).astype(np.float32)
    x, _ = denoise.pack_input(noisy, dens)
    model.predict_tensor(x)  # warm
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        model.predict_tensor(x)
    stop.record()
    stop.synchronize()
    seconds = start.elapsed_time(stop) / 1e3 / repeats
    flops = denoise.forward_flops(shape, model.widths)
    achieved = flops / seconds / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_BF16_TFLOPS, 5), "batch": list(shape), "flops_per_batch": flops,
            "ms_per_batch": round(seconds * 1e3, 3)}


def _denoise_leg(model_path: Path, low: int, audit: float, target: float, launch: dict | None = None) -> dict:
    """The 100-beamlet CT-like Dij at a dose-weighted RMS error target against a
    100x-history Monte Carlo reference: plain Monte Carlo (histories per beamlet set
    from a measured error at 1e5 and sigma ~ N^-1/2, then run and timed) versus
    low-history Monte Carlo plus the trained denoiser with plain Monte Carlo fallback.
    Held-out phantom (a seed never used in training)."""
    from pyradmc_amd import denoise, workloads
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    grid = workloads.ct_like_grid(seed=20261111)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    src = workloads.ct_like_beamlets()
    engine = _shaped_engine(grid, xs, launch)
    model = denoise.load(model_path)
    nb = 10
    engine.run_dij(src, 10 * nb, nb, seed=1, truncation=1e-4)  # warm
    ref = engine.run_dij(src, 100 * low, nb, seed=7, truncation=1e-4)
    probe_n = 100_000
    probe = engine.run_dij(src, probe_n, nb, seed=8, truncation=1e-4)
    eps = float(np.percentile(_window_errors(probe, ref, src, grid), 90))
    n_plain = int(math.ceil(probe_n * (eps / target) ** 2 / nb)) * nb
    started = time.perf_counter()
    plain = engine.run_dij(src, n_plain, nb, seed=9, truncation=1e-4, value_dtype=np.float32)
    wall_plain = time.perf_counter() - started
    plain_err = _window_errors(plain, ref, src, grid)
    builder = denoise.DenoisedDijBuilder(engine, model, low_histories=low, audit_fraction=audit,
                                         tolerance=target, fallback_histories=n_plain, n_batches=nb)
    mfma = _denoiser_roofline(model, (32, grid.shape[2], denoise.WINDOW_XY, denoise.WINDOW_XY))
    builder.build(src, seed=11, value_dtype=np.float32)  # warm (kernels, allocator)
    started = time.perf_counter()
    den, report = builder.build(src, seed=12, truncation=1e-4, value_dtype=np.float32)
    wall_den = time.perf_counter() - started
    den_err = _window_errors(den, ref, src, grid)
    engine.close()
    return {
        "workload": "CT-like 128x128x96 @ 2.5 mm (held-out seed), 100 beamlets 5x5 mm, SAD 100, "
                    f"reference {100 * low} histories per beamlet",
        "target_error": target,
        "plain_histories_per_beamlet": n_plain,
        "plain_wall_s": round(wall_plain, 4),
        "plain_error_median": round(float(np.median(plain_err)), 5),
        "plain_error_p90": round(float(np.percentile(plain_err, 90)), 5),
        "denoised_low_histories_per_beamlet": low,
        "denoised_audit_fraction": audit,
        "denoised_wall_s": round(wall_den, 4),
        "denoised_seconds": {k: round(v, 4) for k, v in report.seconds.items()},
        "denoised_fallback_columns": int(report.fallback.sum()),
        "denoised_error_median": round(float(np.median(den_err)), 5),
        "denoised_error_p90": round(float(np.percentile(den_err, 90)), 5),
        # the plain-MC audit's estimate of the denoised error, and the measured error of
        # the columns that stayed denoised (against the 100x reference)
        "audit_columns": int(report.audited.sum()),
        "audit_error": round(float(report.audit_error), 5),
        "denoised_only_error_median": (round(float(np.median(den_err[~report.fallback])), 5)
                                       if (~report.fallback).any() else None),
        "inference_roofline": mfma,
        "speedup": round(wall_plain / wall_den, 3) if wall_den > 0 else None,
    }


def _plan_denoise_leg(model_path: Path, low: int, audit: float, target: float,
                      launch: dict | None = None) -> dict:
    """The full plan's gantry-0 beam (225 beamlets of 5x5 mm along +z on the 256x256x128
    phantom at 2 mm) at a dose-weighted RMS error target against a 100x-history
    reference, plain Monte Carlo versus low-history Monte Carlo plus the trained
    denoiser, as the 100-beamlet leg does it. It is the plan's one beam whose beamlets
    run along a grid axis: the denoiser's windows are grid-aligned (64x64 voxels across,
    the grid's whole depth along the beam), so the four rotated beams have none and
    stay plain Monte Carlo. The network was trained on 2.5 mm phantoms 96 voxels deep;
    here it sees 2 mm voxels 128 deep, and the plain-MC audit judges it on this plan."""
    from pyradmc_amd import denoise, workloads

    grid, xs, beams, kw = workloads.full_plan()
    src = beams[0]
    nb, trunc = kw["n_batches"], kw["truncation"]
    engine = _shaped_engine(grid, xs, launch)
    model = denoise.load(model_path)
    engine.run_dij(src, 10 * nb, nb, seed=1, truncation=trunc)  # warm
    ref = engine.run_dij(src, 100 * low, nb, seed=7, truncation=trunc)
    probe_n = 100_000
    probe = engine.run_dij(src, probe_n, nb, seed=8, truncation=trunc)
    eps = float(np.percentile(_window_errors(probe, ref, src, grid), 90))
    del probe
    n_plain = int(math.ceil(probe_n * (eps / target) ** 2 / nb)) * nb
    started = time.perf_counter()
    plain = engine.run_dij(src, n_plain, nb, seed=9, truncation=trunc, value_dtype=np.float32)
    wall_plain = time.perf_counter() - started
    plain_err = _window_errors(plain, ref, src, grid)
    del plain
    builder = denoise.DenoisedDijBuilder(engine, model, low_histories=low, audit_fraction=audit,
                                         tolerance=target, fallback_histories=n_plain, n_batches=nb)
    builder.build(src, seed=11, value_dtype=np.float32)  # warm (kernels, allocator)
    started = time.perf_counter()
    den, report = builder.build(src, seed=12, truncation=trunc, value_dtype=np.float32)
    wall_den = time.perf_counter() - started
    den_err = _window_errors(den, ref, src, grid)
    engine.close()
    kept = ~report.fallback
    return {
        "workload": "full plan's gantry-0 beam: CT-like 256x256x128 @ 2 mm, 15x15 beamlets of 5x5 mm along +z "
                    f"from 100 cm SAD (225 columns); reference {100 * low} histories per beamlet; the four "
                    "rotated beams have no grid-aligned denoiser window",
        "target_error": target,
        "plain_histories_per_beamlet": n_plain,
        "plain_wall_s": round(wall_plain, 4),
        "plain_error_median": round(float(np.median(plain_err)), 5),
        "plain_error_p90": round(float(np.percentile(plain_err, 90)), 5),
        "denoised_low_histories_per_beamlet": low,
        "denoised_audit_fraction": audit,
        "denoised_wall_s": round(wall_den, 4),
        "denoised_seconds": {k: round(v, 4) for k, v in report.seconds.items()},
        "denoised_fallback_columns": int(report.fallback.sum()),
        "denoised_error_median": round(float(np.median(den_err)), 5),
        "denoised_error_p90": round(float(np.percentile(den_err, 90)), 5),
        "audit_columns": int(report.audited.sum()),
        "audit_error": round(float(report.audit_error), 5),
        "denoised_only_error_median": round(float(np.median(den_err[kept])), 5) if kept.any() else None,
        "speedup": round(wall_plain / wall_den, 3) if wall_den > 0 else None,
    }


def _phase_seconds(counters: dict) -> dict:
    """run_dij's host wall split: transport (with fold and finalize) and CSC assembly
    (truncation, device-to-host copy, join), summed over calls; and the column groups."""
    return {"transport": round(float(counters.get("transport_s", 0.0)), 4),
            "assemble": round(float(counters.get("assemble_s", 0.0)), 4),
            "groups": int(counters.get("groups", 0))}


def _plan_leg(target: float, probe_n: int, n_vectors: int, cap: int,
              launch: dict | None = None) -> tuple[dict, object]:
    """The full plan: 256x256x128 CT-like phantom at 2 mm, five gantry beams of 15x15
    spectral beamlets (1125 columns). A probe Dij at ``probe_n`` histories per beamlet
    gives each beamlet's sigma; sigma ~ N^-1/2 then sets the histories that bring the
    90th-percentile beamlet to ``target``, and that plain Dij is run and timed wall to
    wall (five run_dij calls, float32 CSC on the host). Then the planning product for
    ``n_vectors`` seeded U[0,1) weight vectors on the device, float64 (bit-exact with
    dose_for_weights) and float32, timed with HIP events around the kernel."""
    from pyradmc_amd import spmm, workloads

    grid, xs, beams, kw = workloads.full_plan()
    nb, trunc = kw["n_batches"], kw["truncation"]
    engine = _shaped_engine(grid, xs, launch)
    if launch is None:  # tuned on one beam at 1e5 histories per beamlet, one sweep (untimed)
        engine.autotune_dij(beams[0], 100_000, nb, seed=999, repeats=2, passes=1, use_cache=False,
                            truncation=trunc, value_dtype=np.float32)
    engine.run_dij(beams[0], 10 * nb, nb, seed=1, truncation=trunc)  # warm

    events: dict = {}

    def plan(n: int, seed: int):
        started = time.perf_counter()
        parts = [engine.run_dij(b, n, nb, seed=seed + i, truncation=trunc, value_dtype=np.float32)
                 for i, b in enumerate(beams)]
        elapsed = time.perf_counter() - started
        events.clear()
        for part in parts:
            for k, v in (part.counters or {}).items():
                events[k] = events.get(k, 0) + v
        return workloads.hstack_dij(parts), elapsed

    probe, probe_wall = plan(probe_n, kw["seed"])
    p90 = float(np.nanpercentile(_per_beamlet_sigma(probe), 90))
    del probe
    n_2pct = int(math.ceil(probe_n * (p90 / target) ** 2 / nb)) * nb
    n_run = min(n_2pct, cap)
    d, wall = plan(n_run, kw["seed"] + 100)
    engine.close()
    sig = _per_beamlet_sigma(d)
    out = {
        "workload": "CT-like 256x256x128 @ 2 mm; 5 beams at gantry 0/72/144/216/288 deg about y through "
                    "the isocentre, 15x15 beamlets of 5x5 mm each from 100 cm SAD (1125 columns), 6 MV-like "
                    "spectrum, truncation 1e-4, f32 CSC",
        "probe_histories_per_beamlet": probe_n,
        "probe_wall_s": round(probe_wall, 3),
        "probe_sigma_p90": round(p90, 5),
        "histories_per_beamlet": n_run,
        "histories_per_beamlet_for_target": n_2pct,
        "target_sigma": target,
        "wall_s": round(wall, 3),
        "histories_per_s": round(n_run * d.n_beamlets / wall, 1),
        "nnz": int(d.indptr[-1]),
        "per_beamlet_sigma_median": round(float(np.nanmedian(sig)), 5),
        "per_beamlet_sigma_p90": round(float(np.nanpercentile(sig, 90)), 5),
        "events_per_history": _events(events, float(n_run) * d.n_beamlets),
        "seconds": _phase_seconds(events),
        "launch": dict(engine.launch, chunk_size=engine.chunk_size or "auto"),
    }
    dev = spmm.DeviceDij(d)
    rng = np.random.default_rng(kw["seed"])
    weights = rng.random((d.n_beamlets, n_vectors))
    products = {}
    for precision, elem in (("f64", 8), ("f32", 4)):
        dev.apply(weights, precision=precision, fetch=False)  # warm (buffers, code objects)
        started = time.perf_counter()
        dev.apply(weights, precision=precision, fetch=False)
        call_s = time.perf_counter() - started
        kernel_s = dev.last_kernel_ms / 1e3
        algorithmic = 8.0 * dev.nnz + elem * dev.n_rows * n_vectors + elem * dev.n_cols * n_vectors
        products[precision] = {
            "kernel_ms": round(kernel_s * 1e3, 3),
            "call_s": round(call_s, 4),
            "vectors_per_s": round(n_vectors / kernel_s, 1),
            "roofline": {"bound": "hbm", "achieved": round(algorithmic / kernel_s / 1e9, 1), "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": round(algorithmic / kernel_s / 1e9 / PEAK_HBM_GBS, 4),
                         "algorithmic_bytes": int(algorithmic)},
        }
    out["product"] = {"weight_vectors": n_vectors, "block_entries": dev.block_entries,
                      "output": "[voxels][vectors] on the device", **products}
    dev.close()
    out["optimise"] = _guarded("plan optimisation", _optimise_leg, d, grid)
    return out, d


def _optimise_leg(d, grid) -> dict:
    """The study's fluence optimisation on the full plan's Dij: a 2.5 cm-radius
    spherical PTV at the isocentre and a 2x6x6 cm OAR box beside it, prescription the
    median PTV dose of uniform weights, OAR limit 0.3, penalty 10. L-BFGS-B on the host
    with the region's rows on the device (wall time, iterations, per-evaluation time),
    beside the same two products per evaluation in scipy on the host CPU."""
    from scipy.sparse import csc_array

    from pyradmc_amd.study import ToyPlanProblem, objective_and_gradient, optimize_weights

    shape = tuple(int(n) for n in grid.shape)
    iso = np.array([shape[i] * float(grid.spacing[i]) / 2.0 for i in range(3)])
    centres = [(np.arange(shape[i]) + 0.5) * float(grid.spacing[i]) - iso[i] for i in range(3)]
    x, y, z = np.meshgrid(*centres, indexing="ij")
    ptv = (x ** 2 + y ** 2 + z ** 2 < 2.5 ** 2).ravel()
    oar = ((x > 3.0) & (x < 5.0) & (np.abs(y) < 3.0) & (np.abs(z) < 3.0)).ravel()
    dij = d.dose_csc()
    uniform_ptv = csc_array(dij)[ptv, :] @ np.ones(d.n_beamlets)
    rx = float(np.median(uniform_ptv))
    started = time.perf_counter()
    problem = ToyPlanProblem(d, ptv, oar, prescription=rx, oar_max=0.3, oar_penalty=10.0)
    problem._roi  # region rows to the device
    setup_s = time.perf_counter() - started
    info: dict = {}
    started = time.perf_counter()
    w = optimize_weights(problem, info=info)
    solve_s = time.perf_counter() - started
    evals = []
    for _ in range(5):
        t0 = time.perf_counter()
        objective_and_gradient(problem, w)
        evals.append(time.perf_counter() - t0)
    roi = csc_array(dij)[ptv | oar, :].tocsc()
    a = np.ones(roi.shape[0])
    cpu = []
    for _ in range(3):
        t0 = time.perf_counter()
        roi @ w
        roi.T @ a
        cpu.append(time.perf_counter() - t0)
    dose = csc_array(dij)[ptv, :] @ w
    return {"ptv_voxels": int(ptv.sum()), "oar_voxels": int(oar.sum()), "region_nnz": int(roi.nnz),
            "setup_s": round(setup_s, 3), "solve_s": round(solve_s, 3), **info,
            "evaluation_ms": round(1e3 * float(np.median(evals)), 3),
            "cpu_scipy_products_per_evaluation_ms": round(1e3 * float(np.median(cpu)), 2),
            "ptv_d98_d2_over_rx": [round(float(np.percentile(dose, 2.0)) / rx, 4),
                                   round(float(np.percentile(dose, 98.0)) / rx, 4)]}


def _cpu_dose_for_weights(d, vectors: int = 2) -> dict:
    """The reference's host product (one weight vector at a time) on the full-plan Dij,
    timed with the oracle's restatement of DijResult.dose_for_weights."""
    from oracle import mc

    rng = np.random.default_rng(1)
    started = time.perf_counter()
    for _ in range(vectors):
        mc.dose_for_weights(d.indptr, d.indices, d.dose, rng.random(d.n_beamlets), d.n_voxels)
    per = (time.perf_counter() - started) / vectors
    return {"s_per_vector": round(per, 3), "cores": 1, "kind": "port",
            "sample": f"full-plan Dij, {vectors} weight vectors, one thread (NumPy)"}


def _cpu_pencil(threads: int) -> dict:
    """The CPU restatement on the quick-start pencil (both readings), for the pencil leg."""
    from oracle import mc
    from pyradmc_amd import workloads

    out = {}
    for units in ("cm", "mm"):
        grid, _, src, kw = workloads.small_pencil(units)
        ora = mc.for_grid(grid, src)
        ora.run(1000, 1, seed=1, threads=threads)  # warm
        started = time.perf_counter()
        ora.run(kw["n_histories"], kw["n_batches"], seed=kw["seed"], threads=threads)
        wall = time.perf_counter() - started
        out[units] = {"wall_s": round(wall, 4), "histories_per_s": round(kw["n_histories"] / wall, 1)}
    out.update(cores=threads, kind="port")
    return out


def _cpu_baseline(sample_histories: int, threads: int) -> dict:
    from oracle import mc
    from pyradmc_amd import workloads

    grid, _, src, _ = workloads.slab_phantom()
    ora = mc.for_grid(grid, src)
    ora.run(max(threads * 100, 1000), 1, seed=1, threads=threads)  # warm (GS grid load, threads)
    started = time.perf_counter()
    ora.run(sample_histories, 10, seed=2, threads=threads)
    wall = time.perf_counter() - started
    return {
        "value": sample_histories / wall,
        "unit": "histories/s",
        "cores": threads,
        "kind": "port",
        "sample": f"heterogeneous slab phantom, {sample_histories} histories in 10 batches, "
                  f"{threads} threads, {wall:.1f} s",
        "host_cpu": _cpu_model(),
        "reference_engine": _reference_engine_rate(),
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _reference_engine_rate() -> dict | None:
    """The reference's own CPU path on the same workload, as recorded with its slab
    fixture: pyRadMC's single-threaded NumPy ReferenceEngine, timed when the fixture was
    written in the build container (it cannot run where the bench runs)."""
    fixture = ROOT / "tests" / "golden" / "ref_hi_slab.npz"
    if not fixture.is_file():
        return None
    with np.load(fixture) as g:
        n, cpu_s = int(g["n_histories"]), float(g["cpu_seconds"])
        shards, batches = int(g["shards"]), int(g["n_batches"])
    return {"value": round(n / cpu_s, 1), "unit": "histories/s", "cores": 1, "kind": "reference",
            "sample": f"pyRadMC ReferenceEngine on the heterogeneous slab phantom, {n} histories in "
                      f"{shards} single-threaded processes x {batches // shards} batches, {cpu_s:.0f} "
                      "CPU-seconds; measured in the build container (Intel Xeon @ 2.70 GHz, not the "
                      "GPU host)"}


def _guarded(name: str, leg, *args):
    """Run a secondary leg; a Python error is recorded in its place, so that the
    throughput line is still printed."""
    print(f"bench: {name} leg", file=sys.stderr, flush=True)
    try:
        return leg(*args)
    except Exception as exc:  # noqa: BLE001 - reported in the JSON line
        print(f"bench: {name} leg failed: {exc!r}", file=sys.stderr, flush=True)
        return {"error": repr(exc)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=int, default=10_000_000)
    ap.add_argument("--batches", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="histories for the CPU baseline (default: about 10 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pencil", action="store_true", help="skip the quick-start pencil leg")
    ap.add_argument("--no-dij", action="store_true", help="skip the batched Dij leg")
    ap.add_argument("--no-head", action="store_true", help="skip the treatment-head leg")
    ap.add_argument("--head-only", action="store_true",
                    help="run only the treatment-head leg and print its JSON (for profiling it)")
    ap.add_argument("--dij-median", type=float, default=2e5,
                    help="median per-beamlet history count of the Dij leg")
    ap.add_argument("--denoise-model", default=str(ROOT / "pyradmc_amd" / "models" / "denoiser_ct.npz"),
                    help="trained denoiser weights for the denoised-Dij leg (skipped if absent)")
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--denoise-low", type=int, default=20_000,
                    help="histories per beamlet of the denoised build's input run")
    ap.add_argument("--denoise-audit", type=float, default=0.1,
                    help="fraction of the denoised build's columns re-simulated with plain Monte Carlo as its judge")
    ap.add_argument("--no-plan", action="store_true", help="skip the full-plan Dij and product leg")
    ap.add_argument("--no-plan-denoise", action="store_true",
                    help="skip the denoised-versus-plain comparison on the full plan's gantry-0 beam")
    ap.add_argument("--plan-denoise-only", action="store_true",
                    help="run only that comparison and print its JSON")
    ap.add_argument("--plan-probe", type=int, default=20_000,
                    help="probe histories per beamlet of the full-plan leg")
    ap.add_argument("--plan-vectors", type=int, default=1000, help="weight vectors of the planning product")
    ap.add_argument("--plan-cap", type=int, default=4_000_000,
                    help="largest histories per beamlet the full-plan leg runs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape and chunk size")
    ap.add_argument("--launch", default=None,
                    help="LANES,REFILL[,PUSH]: fix the rounds' lanes per CU, refill block and push chunk "
                         "instead of tuning "
                         "(e.g. the shape a tuned run reported, to profile that run's kernels)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run three device-to-device copies of this many MiB (a known-traffic "
                         "kernel that calibrates the HBM byte counters of a PMC profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(histories shard across GPUs with no exchange for Dij columns; see the design notes)",
              file=sys.stderr)
        return 2
    if args.head_only:
        print(json.dumps({"head": _head_leg(repeats=2)}), flush=True)
        from pyradmc_amd import _native

        _native.shutdown()
        return 0

    if args.copy_probe_mib > 0:
        import torch

        n = args.copy_probe_mib << 18  # float32 elements
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    denoise_leg = not args.no_denoise and Path(args.denoise_model).is_file()
    if args.plan_denoise_only:  # the engine's default launch shape
        print(json.dumps({"full_plan_denoised": _plan_denoise_leg(Path(args.denoise_model), args.denoise_low,
                                                                  args.denoise_audit, 0.02)}), flush=True)
        from pyradmc_amd import _native

        _native.shutdown()
        return 0
    if denoise_leg:
        # The denoiser runs on torch's HIP runtime: load it before any engine exists.
        print("bench: loading the denoiser (PyTorch)", file=sys.stderr, flush=True)
        import pyradmc_amd.denoise  # noqa: F401

    from pyradmc_amd import HipEngine, workloads

    grid, xs, src, kw = workloads.slab_phantom(n_histories=args.histories, n_batches=args.batches)
    engine = HipEngine(grid, xs)
    tuned = None
    if args.launch:
        shape = [int(v) for v in args.launch.split(",")]
        engine.set_launch(shape[0], shape[1], shape[2] if len(shape) > 2 else 64, True)
    elif not args.no_autotune:
        # Launch shape and chunk size for this workload (untimed; results do not depend on them).
        print("bench: autotune", file=sys.stderr, flush=True)
        tuned = engine.autotune(src, args.histories, args.batches, seed=999, repeats=3, use_cache=False)
    for w in range(args.warmup):
        engine.run(src, args.histories, args.batches, seed=1000 + w)
    engine.set_profiling(True)
    counters: dict = {}
    started = time.perf_counter()
    for k in range(args.steps):
        r = engine.run(src, args.histories, args.batches, seed=kw["seed"] + k)
        for key, v in r.counters.items():
            counters[key] = counters.get(key, 0) + v
    elapsed = time.perf_counter() - started
    engine.set_profiling(False)
    total = args.histories * args.steps
    value = total / elapsed
    # The same steps unprofiled: with graphs on, groups of rounds replay the captured
    # HIP graph (the profiled steps launch kernel by kernel so that each one is timed).
    graph_counts = 0
    started = time.perf_counter()
    for k in range(args.steps):
        graph_counts += engine.run(src, args.histories, args.batches, seed=kw["seed"] + k).counters["graph_launches"]
    replay_elapsed = time.perf_counter() - started

    launches = max(int(counters["electron_launches"]), 1)
    mean_launch_s = counters["electron_kernel_ms"] / 1e3 / launches
    bytes_per_launch = _algorithmic_bytes(counters, "electron") / launches
    achieved = bytes_per_launch / mean_launch_s / 1e9 if mean_launch_s > 0 else 0.0
    traffic, traffic_at = None, None
    tfile = ROOT / "profiles" / "traffic_electron_round.json"
    if tfile.is_file():
        try:
            tinfo = json.loads(tfile.read_text())
            traffic, traffic_at = tinfo.get("hbm_bytes_per_launch"), tinfo.get("workload")
        except (OSError, ValueError):
            traffic = None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 3),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 6),
        "traffic": traffic,
        # HBM bytes come from a separate rocprofv3 --pmc run (FETCH_SIZE, WRITE_SIZE), not
        # from this one: the launch shape it was profiled at is named here
        "traffic_profiled": {"source": "profiles/traffic_electron_round.json", "workload": traffic_at},
        "kernel": "k_electron_round",
        "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
        "mean_launch_ms": round(mean_launch_s * 1e3, 4),
        "launches": launches,
        "electron_kernel_share": round(
            counters["electron_kernel_ms"] / max(counters["electron_kernel_ms"] + counters["photon_kernel_ms"], 1e-9), 4),
    }

    # The Dij legs tune their own launch shape (on their own workloads, untimed) unless
    # a shape is given; the denoised leg runs at the batched Dij's.
    shape = dict(engine.launch) if (args.launch or args.no_autotune) else None
    pencil = None if args.no_pencil else _guarded("pencil", _pencil_leg)
    dij = None if args.no_dij else _guarded("dij", _dij_leg, args.dij_median, shape)
    head = None if args.no_head else _guarded("head", _head_leg)
    dij_shape = shape if shape is not None else (dij or {}).get("launch")
    denoised = (_guarded("denoised_dij", _denoise_leg, Path(args.denoise_model), args.denoise_low,
                         args.denoise_audit, 0.02, dij_shape)
                if denoise_leg else None)
    plan_denoised = (_guarded("full_plan_denoised", _plan_denoise_leg, Path(args.denoise_model), args.denoise_low,
                              args.denoise_audit, 0.02, dij_shape)
                     if denoise_leg and not args.no_plan_denoise else None)
    plan, plan_dij = None, None
    if not args.no_plan:
        got = _guarded("full_plan", _plan_leg, 0.02, args.plan_probe, args.plan_vectors, args.plan_cap, shape)
        plan, plan_dij = got if isinstance(got, tuple) else (got, None)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sample = args.cpu_sample or 2_500_000 * threads  # ~10 s at ~0.25M histories/s/thread
        cpu = _cpu_baseline(sample, threads)
        if not args.no_pencil:
            cpu["pencil"] = _guarded("cpu pencil", _cpu_pencil, threads)
        if plan_dij is not None:
            cpu["dose_for_weights"] = _guarded("cpu dose_for_weights", _cpu_dose_for_weights, plan_dij)

    line = {
        "metric": "histories/s",
        "value": round(value, 1),
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "heterogeneous slab phantom 64x64x96 @ 2 mm (lung 0.26 at 3-7 cm, bone 1.85 "
                        "at 9-11 cm), 10x10 cm parallel 6 MV-like spectral field, GS MSC, analytic water",
            "histories_per_step": args.histories,
            "n_batches": args.batches,
            "device": engine.device_name,
            "tally": "int64 fixed point (1e-9 MeV), f64 fold",
            "launch": {**engine.launch, "chunk_size": engine.chunk_size or "auto",
                       "autotuned": tuned is not None, "profiled_steps": "kernel by kernel"},
        },
        "unprofiled": {"ms_per_step": round(replay_elapsed / args.steps * 1e3, 3),
                         "histories_per_s": round(total / replay_elapsed, 1),
                         "graph_launches": int(graph_counts)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pencil": pencil,
        "dij": dij,
        "denoised_dij": denoised,
        "full_plan": plan,
        "full_plan_denoised": plan_denoised,
        "head": head,
        "events_per_history": _events(counters, total),
    }
    engine.close()
    print(json.dumps(line), flush=True)
    # Release the large page-locked results and every native handle here, while the
    # HIP runtime is certainly alive, rather than during interpreter teardown.
    del plan_dij, engine
    import gc

    from pyradmc_amd import _native

    gc.collect()
    _native.shutdown()
    return 0


if __name__ == "__main__":
    sys.exit(main())
