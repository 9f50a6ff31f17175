"""Treatment-head pre-solve on the GPU: photons from a planar head source are traced
through the beam-limiting devices (and an air column) once, and the exit-plane phase
space they leave is what the engine then transports (pyradmc/geometry/head.py,
pyradmc/backends/warp/presolve.py).

Fidelity ``mode``:

* ``"attenuation"``: Beer-Lambert weights through the stack and the air, photons only
  (the physics of ``CollimatedSource``, as a phase space);
* ``"first_compton"``: additionally one forced Compton interaction per primary: in the
  stack, at the mid-plane of the device chosen by its share of the optical depth, the
  scattered photon kept with its exit attenuation (collimator scatter); in the air
  column, at an exponentially sampled depth, both the photon and the recoil electron
  (contaminant electrons);
* ``"auto"``: ``"first_compton"``.

Stated approximations (those of the reference, head.py:14-55): forced first
interaction with weight splitting (later head scatters neglected); non-Compton
interactions absorb; electrons born in the devices are neglected; the in-device site is
the mid-plane crossing; air interactions are forced on primaries only; kept particles
fly straight and lossless to the exit plane. Forced secondaries below ``weight_floor``
play fair roulette. The population is normalised per primary (every weight times
``kept / n_histories``), so ``run(phase_space, M)`` estimates dose per head primary.

The device kernel gives each primary four fixed record slots (stack photon, air photon,
air electron, transmitted photon) and a scan compacts the kept ones in that order, so
the population, and every run drawn from it, is reproducible. Built-in open-field
sources are generated in the same kernel (the reference samples every head source on
the host, pyradmc/backends/warp/presolve.py:615-620), and the head's draws follow the
source's on each primary's own counter-based stream; the reference's host pre-solve
draws from numpy's PCG64, so the two agree in distribution, not bit for bit.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pyradmc_amd import _native as nat
from pyradmc_amd.constants import ECUT_MEV, PCUT_MEV
from pyradmc_amd.data.materials import AIR
from pyradmc_amd.phasespace import DevicePhaseSpace, InMemoryPhaseSpaceSource

__all__ = ["AirColumn", "presolve_head", "resolve_presolve"]

MODES = ("auto", "first_compton", "attenuation")
# Open-field photon sources whose primaries the pre-solve kernel draws itself.
_GENERATED = frozenset({"PencilBeamSource", "ParallelBeamSource", "SpectralParallelBeamSource",
                        "SpectralBeamSource", "GaussianSpotBeamSource"})
_KIND_TO_IAEA = np.array([1, 2, 3], dtype=np.int32)
_FETCH_BLOCK = 1 << 22


@dataclass(frozen=True)
class AirColumn:
    """The air between the last device and the exit plane, in local ``w`` (cm)
    (pyradmc/geometry/head.py:92-110)."""

    z_start: float
    z_end: float
    material: int = AIR
    density: float = 1.205e-3

    def __post_init__(self) -> None:
        if not 0.0 < self.z_start < self.z_end:
            raise ValueError(f"need 0 < z_start < z_end, got {self.z_start}, {self.z_end}")
        if self.density <= 0.0:
            raise ValueError(f"density must be positive, got {self.density}")


def resolve_presolve(stack, frame, air, exit_z: float, mode: str):
    """(resolved mode, beam frame), validating the geometry (pyradmc/geometry/head.py:
    348-374): a known mode, a frame (the stack's if there is one), ``exit_z`` at or below
    the stack exit and the end of the air column, the air below the stack."""
    if mode not in MODES:
        raise ValueError(f"mode must be one of {MODES}, got {mode!r}")
    resolved = "first_compton" if mode == "auto" else mode
    if stack is not None:
        frame = stack.frame
    if frame is None:
        raise ValueError("without a stack, pass the beam frame explicitly")
    if stack is not None and exit_z < stack.exit_z:
        raise ValueError(f"exit_z {exit_z} is above the stack exit {stack.exit_z}")
    if air is not None:
        if stack is not None and air.z_start < stack.exit_z:
            raise ValueError(f"the air column starts at {air.z_start}, inside the stack (exit {stack.exit_z})")
        if exit_z < air.z_end:
            raise ValueError(f"exit_z {exit_z} is above the air column end {air.z_end}")
    return resolved, frame


def _tables(cross_sections, rows, e_min: float, e_max: float, n_table: int):
    """Log of the macroscopic total and Compton attenuation per (material, density) row
    on a geometric energy grid (pyradmc/geometry/head.py:113-180)."""
    if rows and max(m for m, _ in rows) >= int(cross_sections.n_materials):
        raise ValueError("a device or air material index is beyond this cross-section source; "
                         "tungsten devices need the tabulated backend")
    energies = np.geomspace(e_min, max(e_max, e_min * (1.0 + 1e-9)), int(n_table))
    tot = np.stack([np.log(rho * np.asarray(cross_sections.mu_total_bulk(energies, m), dtype=np.float64))
                    for m, rho in rows])
    com = np.stack([np.log(np.maximum(rho * np.asarray(cross_sections.mu_channels(energies, m)[0], dtype=np.float64),
                                      1e-300)) for m, rho in rows])
    return np.log(energies), tot, com


class _ResidentPopulation:
    """The pre-solve population left in an engine's device memory."""

    def __init__(self, engine, n: int, max_energy: float) -> None:
        self.engine = engine
        self.n = int(n)
        self.max_energy = float(max_energy)
        self.token = None  # the DevicePhaseSpace's token, set by presolve_head

    def __len__(self) -> int:
        return self.n

    def population_blocks(self, block: int = _FETCH_BLOCK):
        """Copy the population back in blocks (what another engine uploads)."""
        if self.engine._population_token != self.token or self.engine._ctx is None:
            raise RuntimeError("the pre-solve population was replaced on its engine")
        for start in range(0, self.n, int(block)):
            yield _fetch(self.engine._ctx, start, min(int(block), self.n - start))


def _fetch(ctx, start: int, count: int) -> dict:
    kind = np.empty(count, np.int32)
    cols = [np.empty(count, np.float32) for _ in range(8)]
    ctx.call("hmc_get_population", int(start), int(count), nat.ptr(kind, nat.c_int32),
             *[nat.ptr(c, nat.ctypes.c_float) for c in cols])
    out = {"particle_type": _KIND_TO_IAEA[kind]}
    out.update(zip(("energy", "x", "y", "z", "ux", "uy", "uz", "weight"), cols))
    return out


def presolve_head(source, *, cross_sections, n_histories: int, seed: int, exit_z: float, stack=None, frame=None,
                  air: AirColumn | None = None, mode: str = "auto", pcut: float = PCUT_MEV, ecut: float = ECUT_MEV,
                  weight_floor: float = 1.0e-3, e_min: float = 0.010, n_table: int = 512, engine=None,
                  device="hip:0", resident: bool = False):
    """Pre-solve the head once on the GPU and return the exit-plane phase space
    (pyradmc/geometry/head.py:377-560, pyradmc/backends/warp/presolve.py:572-687).

    ``source`` emits photons only (a ``GaussianSpotBeamSource`` is the intended input).
    The built-in open-field sources (pencil and parallel photon beams, spectral parallel
    field, spectral and spot fans) are drawn in-kernel on history ``h``'s stream; any
    other source is sampled on the host through ``sample_batch``. ``exit_z`` and the air column are
    local ``w`` in the beam frame (the stack's; pass ``frame`` without a stack). By
    default the population is copied back as an ``InMemoryPhaseSpaceSource``. With
    ``resident=True`` and an ``engine``, it stays in that engine's device memory and a
    ``DevicePhaseSpace`` is returned, which ``engine.run`` transports with no copy
    (another engine copies it over on first use)."""
    from pyradmc_amd.engine import _device_index, _in_kernel, generator_desc

    resolved, frame = resolve_presolve(stack, frame, air, exit_z, mode)
    if resident and engine is None:
        raise ValueError("a resident pre-solve needs the engine that will transport it")
    if n_histories < 1:
        raise ValueError(f"need at least one history, got {n_histories}")
    keep = []
    generator = None
    if _in_kernel(source) and type(source).__name__ in _GENERATED:
        # drawn in-kernel: no host sampling, no upload
        generator, gen_keep, _, _ = generator_desc(source, "photon")
        keep.extend(gen_keep)
        cols = None
    else:
        cols = source.sample_batch(seed, 0, int(n_histories))
        if not np.all(np.asarray(cols["particle_type"]) == 1):
            raise ValueError("the pre-solve transports photon sources only")

    def f32(name):
        if cols is None:
            return None
        a = np.ascontiguousarray(np.asarray(cols[name], dtype=np.float32).reshape(-1))
        keep.append(a)
        return nat.ptr(a, nat.ctypes.c_float)

    def f64(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        keep.append(a)
        return nat.ptr(a, nat.c_double)

    rows = [] if stack is None else list(zip(stack.materials, stack.densities))
    if air is not None:
        rows.append((air.material, air.density))
    q = nat.PresolveDesc()
    q.first_compton = 1 if resolved == "first_compton" else 0
    q.has_air = 1 if air is not None else 0
    q.exit_z = float(exit_z)
    if air is not None:
        q.air_z_start, q.air_z_end = float(air.z_start), float(air.z_end)
    q.pcut, q.ecut, q.weight_floor = float(pcut), float(ecut), float(weight_floor)
    q.origin[:] = [float(c) for c in frame.origin]
    q.basis[:] = [float(c) for c in frame.basis.reshape(-1)]
    q.n_rows = len(rows)
    if rows:
        le, tot, com = _tables(cross_sections, rows, e_min, float(source.max_energy), n_table)
        q.n_table = int(le.size)
        q.log_energies, q.log_mu_total, q.log_mu_compton = f64(le), f64(tot), f64(com)
    stack_desc = None
    if stack is not None:
        if len(stack.devices) > 16:
            raise ValueError("the device pre-solve takes at most 16 devices")
        cs = stack.compile()
        stack_desc = nat.RayWeightsDesc()
        stack_desc.kind = 1
        stack_desc.n_devices = int(cs.kind.size)
        stack_desc.origin[:] = [float(c) for c in cs.origin]
        stack_desc.basis[:] = [float(c) for c in cs.basis.reshape(-1)]
        for name, values in (("device_kind", cs.kind), ("flags", cs.flags), ("n_pairs", cs.n_pairs),
                             ("edge_offset", cs.offset), ("pair_offset", cs.pair_offset)):
            a = np.ascontiguousarray(values, dtype=np.int32)
            keep.append(a)
            setattr(stack_desc, name, nat.ptr(a, nat.c_int32))
        for name, values in (("z_top", cs.z_top), ("z_bottom", cs.z_bottom), ("edge_neg", cs.edge_neg),
                             ("edge_pos", cs.edge_pos), ("density", cs.density)):
            setattr(stack_desc, name, f64(values))
        stack_desc.n_leaf_edges, stack_desc.n_tips = int(cs.leaf_edges.size), int(cs.tips_pos.size)
        stack_desc.leaf_edges = f64(cs.leaf_edges)
        stack_desc.tips_neg = f64(cs.tips_neg)
        stack_desc.tips_pos = f64(cs.tips_pos)

    ctx = engine._ctx if engine is not None else nat.Context(_device_index(device))
    try:
        n_out = nat.c_int64(0)
        ctx.call("hmc_presolve_head", nat.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), int(n_histories),
                 None if generator is None else nat.ctypes.byref(generator),
                 *[f32(c) for c in ("energy", "x", "y", "z", "ux", "uy", "uz", "weight")],
                 None if stack_desc is None else nat.ctypes.byref(stack_desc), nat.ctypes.byref(q),
                 nat.ctypes.byref(n_out))
        n_pop = int(n_out.value)
        if engine is not None:
            engine._population_token = None  # the context's population was replaced
        if resident:
            pop = _ResidentPopulation(engine, n_pop, float(source.max_energy))
            dev = DevicePhaseSpace(pop)
            pop.token = dev.token
            engine._population_token = dev.token
            return dev
        blocks = [_fetch(ctx, s, min(_FETCH_BLOCK, n_pop - s)) for s in range(0, n_pop, _FETCH_BLOCK)]
    finally:
        if engine is None:
            ctx.close()
    merged = {k: np.concatenate([b[k] for b in blocks]) for k in blocks[0]}
    d = np.column_stack([merged["ux"], merged["uy"], merged["uz"]]).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)  # float32 round-off
    return InMemoryPhaseSpaceSource(merged["particle_type"], merged["energy"], merged["x"], merged["y"], merged["z"],
                                    d[:, 0], d[:, 1], d[:, 2], merged["weight"])
