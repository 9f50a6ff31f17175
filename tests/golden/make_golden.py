"""Write the golden vectors the oracle and the HIP engine are pinned against.

Run only in the build container, where the reference checkout is present at
/root/reference. It imports the reference package (pure Python + NumPy) and records
input/output vectors for the functions that closed forms and invariants alone cannot
pin to rounding (the flattened cross-section tables, individual Goudsmit-Saunderson
nodes, end-to-end ReferenceEngine doses, and the summation order of
DijResult.dose_for_weights). The fixtures are data: inputs and the
reference's outputs, nothing of its source.

Usage::

    HOME=<scratch> python tests/golden/make_golden.py tables gs_nodes
    HOME=<scratch> python tests/golden/make_golden.py run_c1_kerma run_c1_cm ...
    HOME=<scratch> python tests/golden/make_golden.py dose_for_weights
    HOME=<scratch> python tests/golden/make_golden.py hi slab <part> <histories> <shard dir>
    HOME=<scratch> python tests/golden/make_golden.py merge_hi slab <shard dir> <part> ...

The ``hi`` shards (one process each, so several run side by side) are written to the
shard directory given on the command line; ``merge_hi`` pools them into one fixture.

HOME is redirected by the caller because the reference persists its GS grid under
``~/.cache``.
"""

from __future__ import annotations

import math
import sys
import time
from pathlib import Path

import numpy as np

REFERENCE = Path("/root/reference")
OUT = Path(__file__).resolve().parent

sys.path.insert(0, str(REFERENCE))


def _tables() -> None:
    from pyradmc.data.analytic import AnalyticCrossSections

    xs = AnalyticCrossSections()
    e_max = 6.0 * (1.0 + 1.0e-6)
    t = xs.build_tables(ecut=0.2, pcut=0.05, e_max=e_max)
    arrays = {
        name: np.asarray(getattr(t, name))
        for name in (
            "mu_compton",
            "mu_photo",
            "mu_pair",
            "mu_rayleigh",
            "majorant",
            "stopping_restricted",
            "stopping_radiative",
            "moller",
            "csda_range",
            "scattering_power",
            "log_eta",
            "restricted_range",
            "energy_of_restricted_range",
            "coherent_x",
            "coherent_cumulative",
        )
    }
    scalars = {
        name: np.float64(getattr(t, name))
        for name in (
            "photon_log_e_min",
            "photon_inv_dlog",
            "electron_log_e_min",
            "electron_inv_dlog",
            "range_log_r_min",
            "range_inv_dlog",
            "ecut",
            "pcut",
            "e_max",
        )
    }
    # Host-API point queries at off-grid energies: pins the formulas themselves.
    energies = np.array([0.06, 0.1, 0.2, 0.35, 0.5, 1.0, 1.022, 1.5, 2.0, 3.3, 6.0])
    queries = {
        "q_energy": energies,
        "q_mu_compton": np.array([xs.mu_over_rho(float(e), 0, 0) for e in energies]),
        "q_mu_photo": np.array([xs.mu_over_rho(float(e), 0, 1) for e in energies]),
        "q_mu_pair": np.array([xs.mu_over_rho(float(e), 0, 2) for e in energies]),
        "q_restricted_stopping": np.array(
            [xs.restricted_stopping_power(float(e), 0, 0.2) for e in energies]
        ),
        "q_radiative": np.array([xs.radiative_stopping_power(float(e), 0) for e in energies]),
        "q_moller": np.array([xs.moller_cross_section(float(e), 0, 0.2) for e in energies]),
        "q_csda_range": np.array([xs.csda_range(float(e), 0) for e in energies]),
        "q_scattering_power": np.array(
            [xs.scattering_power(float(e), 0, 0.2) for e in energies]
        ),
        "q_screening": np.array([xs.elastic_screening(float(e), 0) for e in energies]),
        "q_restricted_range": np.array(
            [xs.restricted_range(float(e), 0, 0.2) for e in energies]
        ),
    }
    np.savez_compressed(OUT / "tables_water_analytic.npz", **arrays, **scalars, **queries)
    print("tables written")


def _gs_nodes() -> None:
    from pyradmc.data import goudsmit_saunderson as gs
    from pyradmc.data.analytic import AnalyticCrossSections

    xs = AnalyticCrossSections()
    t = xs.build_tables(ecut=0.2, pcut=0.05, e_max=6.0 * (1.0 + 1.0e-6))
    log_eta_min, log_eta_max, theta2_max = gs.gs_window_from_tables(t, [0])
    bins = 4.0
    ix0 = math.floor(log_eta_min * bins)
    ix_top = math.floor(log_eta_max * bins) + 1
    iy0 = math.floor(math.log(1.0e-12) * bins)
    iy_top = math.floor(math.log(theta2_max) * bins) + 1
    # Representative nodes: tiny-Lambda composition, doubling composition, and series.
    picks = []
    for i in (ix0, (ix0 + ix_top) // 2, ix_top):
        eta = math.exp(i / bins)
        g1 = gs.first_transport_moment(eta)
        for target_lam in (0.05, 0.7, 3.0, 40.0, 250.0, 400.0, 3000.0):
            j = round(math.log(2.0 * g1 * target_lam) * bins)
            j = min(max(j, iy0), iy_top)
            picks.append((i, j))
    picks = sorted(set(picks))
    values = np.stack(
        [
            gs.gs_scaled_deflection_table(math.exp(i / bins), math.exp(j / bins), 512, 4096)
            for i, j in picks
        ]
    )
    lam = np.array(
        [
            0.5 * math.exp(j / bins) / gs.first_transport_moment(math.exp(i / bins))
            for i, j in picks
        ]
    )
    # Moments of one eta for the series branch.
    eta_mid = math.exp(((ix0 + ix_top) // 2) / bins)
    moments = gs.screened_rutherford_moments(eta_mid, 1024)
    np.savez_compressed(
        OUT / "gs_nodes_water.npz",
        nodes=np.array(picks, dtype=np.int64),
        lam=lam,
        values=values,
        window=np.array([ix0, ix_top, iy0, iy_top], dtype=np.int64),
        window_float=np.array([log_eta_min, log_eta_max, theta2_max]),
        moments_eta=np.float64(eta_mid),
        moments=moments,
    )
    print("gs nodes written:", len(picks), "window", ix0, ix_top, iy0, iy_top)


def _engine():
    from pyradmc.backends.ref.engine import ReferenceEngine
    from pyradmc.rng.host import HostRNG

    return ReferenceEngine, HostRNG


def _run_c1(spacing: float, position: tuple[float, float, float], electrons: bool, tag: str,
            n: int = 80_000, batches: int = 10) -> None:
    from pyradmc.data.analytic import AnalyticCrossSections
    from pyradmc.geometry.grid import VoxelGrid
    from pyradmc.geometry.source import PencilBeamSource

    ReferenceEngine, HostRNG = _engine()
    grid = VoxelGrid.uniform_water((16, 16, 16), (spacing, spacing, spacing))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    engine = ReferenceEngine(grid, xs, HostRNG())
    source = PencilBeamSource(6.0, position, (0.0, 0.0, 1.0))
    started = time.perf_counter()
    r = engine.run(source, n, batches, seed=20260726, transport_electrons=electrons)
    wall = time.perf_counter() - started
    np.savez_compressed(
        OUT / f"ref_run_{tag}.npz",
        dose=r.dose,
        dose_sigma=r.dose_sigma,
        energy_emitted=r.energy_emitted,
        energy_deposited=r.energy_deposited,
        energy_escaped=r.energy_escaped,
        energy_unscored=r.energy_unscored,
        n_histories=n,
        n_batches=batches,
        seed=20260726,
        spacing=spacing,
        position=np.array(position),
        transport_electrons=electrons,
        wall_seconds=wall,
    )
    print(f"{tag}: {n} histories in {wall:.1f} s ({n / wall:.1f} hist/s)")


def _dij_small() -> None:
    from pyradmc.data.analytic import AnalyticCrossSections
    from pyradmc.geometry.grid import VoxelGrid
    from pyradmc.geometry.source import BeamletGridSource

    ReferenceEngine, HostRNG = _engine()
    grid = VoxelGrid.uniform_water((8, 8, 16), (0.5, 0.5, 0.5))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    engine = ReferenceEngine(grid, xs, HostRNG())
    source = BeamletGridSource(2.0, -0.1, (1.0, 3.0), (1.0, 3.0), 2, 2)
    out = {}
    for correlated in (True, False):
        started = time.perf_counter()
        d = engine.run_dij(source, 3000, 6, seed=20260711, correlated=correlated)
        tag = "corr" if correlated else "indep"
        for name in ("indptr", "indices", "dose", "sigma"):
            out[f"{tag}_{name}"] = np.asarray(getattr(d, name))
        for name in ("energy_emitted", "energy_deposited", "energy_escaped", "energy_unscored"):
            out[f"{tag}_{name}"] = np.float64(getattr(d, name))
        print(f"dij {tag}: {time.perf_counter() - started:.1f} s")
    np.savez_compressed(OUT / "ref_dij_small.npz", **out)


def _dij_fan() -> None:
    """Divergent spectral beamlets (SpectralBeamletSource) through a parallelogram
    aperture into a water block with a lung-density slab: pins the fan generator
    (focal point, axes, spectrum draw order) and density scaling in a Dij."""
    from pyradmc.data.analytic import AnalyticCrossSections
    from pyradmc.geometry.grid import VoxelGrid
    from pyradmc.geometry.source import SpectralBeamletSource
    from pyradmc.geometry.spectrum import Spectrum

    ReferenceEngine, HostRNG = _engine()
    shape = (16, 16, 24)
    density = np.ones(shape)
    density[:, :, 6:12] = 0.3
    grid = VoxelGrid(shape=shape, spacing=(0.5, 0.5, 0.5), density=density,
                     material=np.zeros(shape, dtype=np.int32))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    engine = ReferenceEngine(grid, xs, HostRNG())
    edges = np.linspace(0.1, 6.0, 61)
    mid = 0.5 * (edges[:-1] + edges[1:])
    spectrum = Spectrum(edges, mid * np.exp(-mid / 1.8), convention="number")
    centers = [(3.5, 3.5, 4.0), (4.5, 3.5, 4.0), (3.5, 4.5, 4.0), (4.5, 4.8, 4.0)]
    source = SpectralBeamletSource(spectrum, (4.0, 4.0, -96.0), centers, 1.0, 0.8,
                                   u_axis=(1.0, 0.0, 0.0), v_axis=(0.3, 1.0, 0.0))
    started = time.perf_counter()
    d = engine.run_dij(source, 16000, 8, seed=20260712, correlated=True, truncation=1e-3)
    out = {name: np.asarray(getattr(d, name)) for name in ("indptr", "indices", "dose", "sigma")}
    for name in ("energy_emitted", "energy_deposited", "energy_escaped", "energy_unscored"):
        out[name] = np.float64(getattr(d, name))
    out["density"] = density
    print(f"dij fan: {time.perf_counter() - started:.1f} s")
    np.savez_compressed(OUT / "ref_dij_fan.npz", **out)


def _dose_for_weights() -> None:
    """DijResult.dose_for_weights on (a) the fan-beamlet Dij the reference engine wrote
    (ref_dij_fan.npz) and (b) a synthetic CSC whose float32 entries and weights span
    sixteen decades, with repeated (voxel, beamlet) entries, so that a sum in any order
    other than the reference's (ascending beamlet, CSC order inside a column, product
    rounded before the add) shows in the low bits."""
    from pyradmc.scoring.dij import DijResult

    def product(shape, indptr, indices, dose, weights):
        d = DijResult(grid_shape=shape, n_beamlets=len(indptr) - 1, n_histories_per_beamlet=1,
                      n_batches=1, truncation=0.0, indptr=indptr, indices=indices, dose=dose,
                      sigma=np.zeros_like(dose), energy_emitted=0.0, energy_deposited=0.0,
                      energy_escaped=0.0)
        return np.stack([d.dose_for_weights(w).ravel() for w in weights], axis=1)

    out = {}
    fan = np.load(OUT / "ref_dij_fan.npz")
    rng = np.random.default_rng(20261201)
    w_fan = rng.random((5, len(fan["indptr"]) - 1))
    out["fan_weights"] = w_fan
    out["fan_dose"] = product((16, 16, 24), fan["indptr"], fan["indices"], fan["dose"], w_fan)

    n_rows, n_cols = 203, 37
    indptr, indices = [0], []
    for _ in range(n_cols):
        k = int(rng.integers(0, 60))
        rows = np.sort(rng.choice(n_rows, size=k, replace=False))
        rows = np.concatenate([rows, rng.choice(rows, size=min(3, k), replace=False)]) if k else rows
        indices.append(rng.permutation(rows))
        indptr.append(indptr[-1] + len(rows))
    indices = np.concatenate(indices).astype(np.int64)
    indptr = np.asarray(indptr, dtype=np.int64)
    dose = (rng.random(len(indices)) * 10.0 ** rng.integers(-8, 8, len(indices))).astype(np.float32)
    w_syn = rng.random((4, n_cols)) * 10.0 ** rng.integers(-8, 8, (4, n_cols))
    w_syn[1] *= np.where(rng.random(n_cols) < 0.3, -1.0, 1.0)
    out.update(syn_indptr=indptr, syn_indices=indices, syn_dose=dose, syn_weights=w_syn,
               syn_n_rows=np.int64(n_rows),
               syn_out=product((n_rows, 1, 1), indptr, indices, dose, w_syn))
    np.savez_compressed(OUT / "ref_dose_for_weights.npz", **out)
    print("dose_for_weights: fan", out["fan_dose"].shape, "synthetic", out["syn_out"].shape)


def _spectral_parallel_source():
    """Parallel 10x10 cm field of 6 MV-like spectral photons: the reference's sanctioned
    extension point (a Source subclass); 4 uniforms: energy (2), then x, then y."""
    from pyradmc.geometry.source import Primary, Source
    from pyradmc.geometry.spectrum import Spectrum
    from pyradmc.rng import uniform

    edges = np.linspace(0.1, 6.0, 61)
    mid = 0.5 * (edges[:-1] + edges[1:])
    spectrum = Spectrum(edges, mid * np.exp(-mid / 1.8), convention="number")

    class SpectralParallel(Source):
        @property
        def max_energy(self) -> float:
            return spectrum.max_energy

        def emit(self, rng_state):
            energy = spectrum.sample_energy(rng_state)
            x = 1.4 + 10.0 * uniform(rng_state)
            y = 1.4 + 10.0 * uniform(rng_state)
            return Primary(energy, x, y, -0.1, 0.0, 0.0, 1.0)

    return SpectralParallel()


def _slab_grid():
    from pyradmc.geometry.grid import VoxelGrid

    shape = (64, 64, 96)
    density = np.ones(shape)
    density[:, :, 15:35] = 0.26
    density[:, :, 45:55] = 1.85
    return VoxelGrid(shape=shape, spacing=(0.2, 0.2, 0.2), density=density,
                     material=np.zeros(shape, dtype=np.int32))


def _run_hi(config: str, part: int, n: int, shard_dir: Path) -> None:
    """One shard of a high-statistics reference run (distinct seed per shard)."""
    from pyradmc.data.analytic import AnalyticCrossSections
    from pyradmc.geometry.grid import VoxelGrid
    from pyradmc.geometry.source import PencilBeamSource
    from pyradmc.scoring.grid import ScoringGrid

    ReferenceEngine, HostRNG = _engine()
    scoring = None
    if config == "c1_cm":
        grid = VoxelGrid.uniform_water((16, 16, 16), (1.0, 1.0, 1.0))
        source = PencilBeamSource(6.0, (8.0, 8.0, -1.0), (0.0, 0.0, 1.0))
    elif config == "c1_mm":
        grid = VoxelGrid.uniform_water((16, 16, 16), (0.1, 0.1, 0.1))
        source = PencilBeamSource(6.0, (0.8, 0.8, -0.1), (0.0, 0.0, 1.0))
    elif config == "slab":
        grid = _slab_grid()
        source = _spectral_parallel_source()
        scoring = ScoringGrid.rebin(grid, (16, 16, 48), (0.8, 0.8, 0.4), (0.0, 0.0, 0.0))
    else:
        raise ValueError(config)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    engine = ReferenceEngine(grid, xs, HostRNG())
    started = time.perf_counter()
    r = engine.run(source, n, 10, seed=1000 + part, scoring_grid=scoring)
    wall = time.perf_counter() - started
    np.savez_compressed(
        shard_dir / f"hi_{config}_{part}.npz",
        dose=r.dose, dose_sigma=r.dose_sigma, energy_emitted=r.energy_emitted,
        energy_deposited=r.energy_deposited, energy_escaped=r.energy_escaped,
        energy_unscored=r.energy_unscored, n=n, wall=wall,
    )
    print(f"hi {config} part {part}: {n} in {wall:.1f} s ({n / wall:.1f} hist/s)")


def _merge_hi(config: str, shard_dir: Path, parts: list[int]) -> None:
    """Pool equal-size shards: mean of shard means; standard error from the shard
    standard errors (each from 10 batches), i.e. 10 * shards batches in total."""
    runs = [np.load(shard_dir / f"hi_{config}_{p}.npz") for p in parts]
    if len({int(r["n"]) for r in runs}) != 1:
        raise ValueError("shards must hold equal history counts")
    k = len(runs)
    dose = sum(r["dose"] for r in runs) / k
    sigma = np.sqrt(sum(r["dose_sigma"] ** 2 for r in runs)) / k
    books = {name: sum(float(r[name]) for r in runs)
             for name in ("energy_emitted", "energy_deposited", "energy_escaped", "energy_unscored")}
    n = int(sum(int(r["n"]) for r in runs))
    wall = float(sum(float(r["wall"]) for r in runs))
    np.savez_compressed(OUT / f"ref_hi_{config}.npz", dose=dose, dose_sigma=sigma,
                        n_batches=10 * k, n_histories=n, cpu_seconds=wall, shards=k, **books)
    print(f"merged {k} shards of {config}: {n} histories, {n / wall:.1f} hist/s per core")


TASKS = {
    "tables": _tables,
    "gs_nodes": _gs_nodes,
    "run_c1_kerma": lambda: _run_c1(1.0, (8.0, 8.0, -1.0), False, "c1_cm_kerma"),
    "run_c1_cm": lambda: _run_c1(1.0, (8.0, 8.0, -1.0), True, "c1_cm"),
    "run_c1_mm": lambda: _run_c1(0.1, (0.8, 0.8, -0.1), True, "c1_mm"),
    "dij_small": _dij_small,
    "dij_fan": _dij_fan,
    "dose_for_weights": _dose_for_weights,
}


def _two_material() -> None:
    """The reference's synthetic two-material slab (water, and 'air' at half the water
    values) under a 2 MeV parallel beam: KERMA, and electron transport scored as dose to
    medium and as dose to water. The inputs are tests/problems.py's two_material_*."""
    from pyradmc.data.tabulated.model import TabulatedData
    from pyradmc.data.tabulated.source import TabulatedCrossSections
    from pyradmc.geometry.grid import VoxelGrid
    from pyradmc.geometry.source import ParallelBeamSource

    sys.path.insert(0, str(OUT.parent.parent))
    from tests import problems as P

    ReferenceEngine, HostRNG = _engine()
    d = P.two_material_data()
    data = TabulatedData(
        photon_energies=d["photon_energies"], mu_over_rho=dict(d["mu_over_rho"]),
        electron_energies=d["electron_energies"], restricted_stopping=d["restricted_stopping"],
        radiative_stopping=d["radiative_stopping"], moller=d["moller"], csda_range=d["csda_range"],
        scattering_power=d["scattering_power"], delta_cut=d["delta_cut"], materials=d["materials"],
        provenance=d["provenance"])
    material = np.zeros(P.TWO_MATERIAL_SHAPE, dtype=np.int32)
    material[:, :, P.TWO_MATERIAL_SLAB] = 1
    grid = VoxelGrid(shape=P.TWO_MATERIAL_SHAPE, spacing=P.TWO_MATERIAL_SPACING,
                     density=np.ones(P.TWO_MATERIAL_SHAPE), material=material)
    xs = TabulatedCrossSections(data, geometry_densities=grid.max_density_by_material())
    source = ParallelBeamSource(energy=2.0, z=-1.0, x_range=(0.0, 16.0), y_range=(0.0, 16.0))
    out = {}
    for tag, electrons, mode, n, batches, seed in (
        ("kerma", False, "dose_to_medium", 120_000, 12, 20261201),
        ("medium", True, "dose_to_medium", 60_000, 12, 20261202),
        ("water", True, "dose_to_water", 60_000, 12, 20261203),
    ):
        engine = ReferenceEngine(grid, xs, HostRNG())
        started = time.perf_counter()
        r = engine.run(source, n, batches, seed=seed, transport_electrons=electrons, scoring_mode=mode)
        wall = time.perf_counter() - started
        out.update({f"{tag}_dose": r.dose, f"{tag}_sigma": r.dose_sigma,
                    f"{tag}_books": np.array([r.energy_emitted, r.energy_deposited, r.energy_escaped,
                                              r.energy_unscored]),
                    f"{tag}_run": np.array([n, batches, seed]), f"{tag}_seconds": wall})
        print(f"two_material {tag}: {n} histories in {wall:.1f} s", flush=True)
    np.savez_compressed(OUT / "ref_two_material_slab.npz", **out)


TASKS["two_material"] = _two_material


if __name__ == "__main__":
    if sys.argv[1] == "hi":
        shards = Path(sys.argv[5])
        shards.mkdir(parents=True, exist_ok=True)
        _run_hi(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), shards)
    elif sys.argv[1] == "merge_hi":
        _merge_hi(sys.argv[2], Path(sys.argv[3]), [int(v) for v in sys.argv[4:]])
    else:
        for task in sys.argv[1:]:
            TASKS[task]()
