"""HipEngine: drop-in for the reference's ``WarpEngine`` on one MI355X.

Same constructor (``grid, cross_sections, device, chunk_size, queue_factor``) and the
same ``run`` / ``run_dij`` signatures and semantics as
``pyradmc/backends/warp/engine.py:369-1230`` (and ``ReferenceEngine``,
``pyradmc/backends/ref/engine.py:110-463``). The host side validates arguments,
flattens the cross-sections (``build_tables``), derives the Goudsmit-Saunderson
window, and drives the native library through its C ABI (``include/hipmc.h``):
generation, transport, per-batch float64 fold, finalize, and the device truncation and
CSC compaction all run on the GPU. There is no CPU fallback: the engine refuses to
construct without the native library and a HIP device.

History-to-stream mapping (reproducible on one device under any chunking, grouping or
batching): ``run`` keys history ``h`` (batch-major order) by ``h``; ``run_dij`` keys
the ``r``-th history of beamlet ``j`` by ``r`` (correlated, the default) or by its
global index ``h = offset_j + r`` (independent), where ``offset_j = j * n_per`` for
equal counts. ``run_dij`` additionally accepts a per-beamlet count array (uneven
beamlets); each count must divide by ``n_batches``.
"""

from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from pyradmc_amd import _native as nat
from pyradmc_amd.constants import (
    DIJ_TRUNCATION_RELATIVE,
    E_MAX_MARGIN,
    ECUT_MEV,
    ENERGY_QUANTUM_MEV,
    PCUT_MEV,
    VERSION,
    default_step_energy_fraction,
)
from pyradmc_amd.data.gs import build_gs_grid, gs_window_from_tables
from pyradmc_amd.progress import ProgressEmitter
from pyradmc_amd.scoring import DijResult, RunProvenance, ScoringGrid, TransportResult

_CHUNK_HISTORIES_PER_CU = 65536
_CHUNK_SIZE_CAP = 32 * 1024 * 1024
_CHUNK_MEMORY_FRACTION = 0.25
_QUEUE_BYTES_PER_SLOT = 4 * 48
_GROUP_MEMORY_FRACTION = 0.35
_SLOT_MEMORY_FRACTION = 0.25
# Host-sampled primaries: histories sampled per sample_batch call, and at most this many
# primaries (36 B each on the device, about twice that on the host) per transport call.
_SAMPLE_BLOCK = 1 << 20
_SAMPLE_CALL = 4 << 20


def _device_index(device) -> int:
    if isinstance(device, int):
        return device
    text = str(device)
    if text in ("hip", "cuda", "gpu"):
        return 0
    for prefix in ("hip:", "cuda:", "gpu:"):
        if text.startswith(prefix):
            return int(text[len(prefix):])
    raise ValueError(f"unsupported device {device!r}; expected 'hip:N'")


def _validate_msc(msc_model: str) -> int:
    if msc_model not in ("gaussian", "gs"):
        raise ValueError(f"unknown msc_model {msc_model!r}; expected 'gaussian' or 'gs'")
    return 1 if msc_model == "gs" else 0


def _validate_scoring_mode(scoring_mode: str, transport_electrons: bool) -> bool:
    if scoring_mode not in ("dose_to_medium", "dose_to_water"):
        raise ValueError(
            f"unknown scoring_mode {scoring_mode!r}; expected 'dose_to_medium' or 'dose_to_water'"
        )
    if scoring_mode == "dose_to_water" and not transport_electrons:
        raise ValueError("dose_to_water needs transport_electrons=True (no electron, no SPR)")
    return scoring_mode == "dose_to_water"


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


_BUILTIN_SOURCES = frozenset({
    "PencilBeamSource", "ParallelBeamSource", "SpectralParallelBeamSource",
    "SpectralBeamSource", "BeamletGridSource", "SpectralBeamletSource", "DevicePhaseSpace",
    "GaussianSpotBeamSource", "GaussianSpotBeamletSource",
})
_BUILTIN_MODULES = frozenset({"pyradmc_amd.geometry", "pyradmc.geometry.source", "pyradmc_amd.phasespace"})
_UPLOAD_COLUMNS = ("energy", "x", "y", "z", "ux", "uy", "uz", "weight")


# Host results at least this large (fetched CSC arrays, a run's dose and sigma) go to
# page-locked memory: copies into it run at the link's rate, into fresh pageable memory
# through staging buffers at the first-touch page-fault rate.
_PINNED_MIN_BYTES = 1 << 20


def host_array(n: int, dtype) -> np.ndarray:
    """An uninitialised 1-D host array for device results (page-locked when large)."""
    dt = np.dtype(dtype)
    if n * dt.itemsize >= _PINNED_MIN_BYTES:
        return nat.pinned_empty(n, dt)
    return np.empty(n, dtype=dt)


def _joined(parts: list, dtype) -> np.ndarray:
    """One array from the per-group pieces, without a copy when there is one group."""
    if not parts:
        return np.zeros(0, dtype)
    if len(parts) == 1:
        return parts[0]
    return np.concatenate(parts, out=host_array(sum(len(p) for p in parts), parts[0].dtype))


_WEIGHTED_SOURCES = frozenset({
    "CollimatedSource", "CollimatedBeamletSource", "TransmissionMaskSource", "TransmissionMaskBeamletSource",
})


def _weighted_parts(source) -> tuple:
    """(inner source, ray weight model) of this package's collimated and mask-shaped
    sources (exact types), else (source, None)."""
    t = type(source)
    if t.__name__ in _WEIGHTED_SOURCES and t.__module__ == "pyradmc_amd.collimation":
        return source.inner, source.model
    return source, None


def _in_kernel(source) -> bool:
    """Exact-type routing (pyradmc/backends/warp/engine.py:523-531): only the built-in
    source types themselves get the in-kernel generator; any other source, a subclass
    included (it may override emit or sample_batch), is host pre-sampled. A collimated
    or mask-shaped source around a built-in source (not around another wrapper) stays
    in-kernel: the generator applies the ray factor on the device."""
    inner, model = _weighted_parts(source)
    if model is not None:
        return _weighted_parts(inner)[1] is None and _in_kernel(inner)
    t = type(source)
    return t.__name__ in _BUILTIN_SOURCES and t.__module__ in _BUILTIN_MODULES


def _iaea_kinds(code) -> np.ndarray:
    """IAEA particle codes (1 photon, 2 electron, 3 positron) as the kernels' kinds."""
    code = np.asarray(code).reshape(-1)
    kind = np.full(code.shape, -1, dtype=np.int32)
    for iaea, k in ((1, 0), (2, 1), (3, 2)):
        kind[code == iaea] = k
    if np.any(kind < 0):
        bad = sorted(set(np.asarray(code[kind < 0]).tolist()))[:5]
        raise ValueError(f"sampled particle_type codes {bad} are not photon (1), electron (2) or positron (3)")
    return kind


def _type_name(obj) -> str:
    return type(obj).__name__


def generator_desc(source, primary_kind: str = "photon", beamlets: range | None = None):
    """The in-kernel generator's description of a built-in source (hmc_source):
    (descriptor, arrays it points into, analytic emitted energy?, mono energy)."""
    d = nat.SourceDesc()
    keep: list = []
    name = _type_name(source)
    d.particle = {"photon": 0, "electron": 1, "positron": 2}[primary_kind]
    analytic = False
    energy = 0.0

    def spectrum_arrays(spectrum):
        edges = np.ascontiguousarray(np.asarray(spectrum.edges, dtype=np.float64))
        cdf = np.ascontiguousarray(np.asarray(spectrum.cdf, dtype=np.float64))
        keep.extend([edges, cdf])
        d.n_bins = int(cdf.size)
        d.edges = nat.ptr(edges, nat.c_double)
        d.cdf = nat.ptr(cdf, nat.c_double)

    if name == "PencilBeamSource":
        d.kind = nat.SRC_PENCIL
        d.energy = float(source.energy)
        d.position[:] = [float(c) for c in source.position]
        d.direction[:] = [float(c) for c in source.direction]
        analytic, energy = True, float(source.energy)
    elif name == "ParallelBeamSource":
        d.kind = nat.SRC_PARALLEL
        d.energy = float(source.energy)
        d.z0 = float(source.z)
        d.x0, d.x_extent = float(source.x_range[0]), float(source.x_range[1] - source.x_range[0])
        d.y0, d.y_extent = float(source.y_range[0]), float(source.y_range[1] - source.y_range[0])
        analytic, energy = True, float(source.energy)
    elif name == "SpectralParallelBeamSource":
        d.kind = nat.SRC_PARALLEL_SPECTRAL
        d.particle = 0
        d.z0 = float(source.z)
        d.x0, d.x_extent = float(source.x_range[0]), float(source.x_range[1] - source.x_range[0])
        d.y0, d.y_extent = float(source.y_range[0]), float(source.y_range[1] - source.y_range[0])
        spectrum_arrays(source.spectrum)
    elif name in ("SpectralBeamSource", "SpectralBeamletSource", "GaussianSpotBeamSource",
                  "GaussianSpotBeamletSource"):
        spot = name.startswith("GaussianSpot")
        d.kind = nat.SRC_SPOT_FAN if spot else nat.SRC_SPECTRAL_FAN
        if spot:
            d.sigma_u, d.sigma_v = float(source.sigma_u), float(source.sigma_v)
        d.particle = 0
        spectrum_arrays(source.spectrum)
        d.focal[:] = [float(c) for c in source.focal_point]
        d.u_axis[:] = [float(c) for c in source.u_axis]
        d.v_axis[:] = [float(c) for c in source.v_axis]
        d.width_u, d.width_v = float(source.width_u), float(source.width_v)
        centers = source.centers if name.endswith("BeamletSource") else (source.center,)
        sel = list(beamlets) if beamlets is not None else list(range(len(centers)))
        params = np.ascontiguousarray(np.array([centers[j] for j in sel], dtype=np.float64).reshape(-1, 3))
        keep.append(params)
        d.n_beamlets = int(params.shape[0])
        d.beamlet_params = nat.ptr(params, nat.c_double)
    elif name == "BeamletGridSource":
        d.kind = nat.SRC_LATTICE
        d.particle = 0
        d.energy = float(source.energy)
        d.z0 = float(source.z)
        sel = list(beamlets) if beamlets is not None else list(range(source.n_beamlets))
        rows = []
        for j in sel:
            x_lo, x_hi, y_lo, y_hi = source.beamlet_bounds(j)
            rows.append((x_lo, x_hi - x_lo, y_lo, y_hi - y_lo))
        params = np.ascontiguousarray(np.array(rows, dtype=np.float64))
        keep.append(params)
        d.n_beamlets = int(params.shape[0])
        d.beamlet_params = nat.ptr(params, nat.c_double)
        analytic, energy = True, float(source.energy)
    else:
        raise ValueError(
            f"source type {name} has no in-kernel generator in this engine; supported: "
            + ", ".join(sorted(_BUILTIN_SOURCES))
        )
    return d, keep, analytic, energy


@dataclass
class HipEngine:
    """One-GPU transport-and-scoring engine over a voxel grid."""

    grid: object
    cross_sections: object
    device: object = "hip:0"
    chunk_size: int | None = None
    queue_factor: int = 16
    _ctx: object = field(default=None, init=False, repr=False)
    _table_cache: dict = field(default_factory=dict, init=False, repr=False)
    _lock: threading.Lock = field(default_factory=threading.Lock, init=False, repr=False)
    _grid_uploaded: bool = field(default=False, init=False, repr=False)
    _uploaded_tables: object = field(default=None, init=False, repr=False)
    _info: dict | None = field(default=None, init=False, repr=False)
    _scoring_default: object = field(default=None, init=False, repr=False)
    _scoring_last: object = field(default=None, init=False, repr=False)
    _map_bytes: int = field(default=0, init=False, repr=False)  # largest maps + sums set so far
    launch: dict = field(default_factory=dict, init=False, repr=False)  # set_launch's knobs

    def __post_init__(self) -> None:
        if self.queue_factor < 1:
            raise ValueError(f"queue_factor must be positive, got {self.queue_factor}")
        if self.chunk_size is not None and self.chunk_size < 1:
            raise ValueError(f"chunk_size must be positive, got {self.chunk_size}")
        if int(self.cross_sections.n_materials) > 256:
            raise ValueError(
                "the HIP engine addresses at most 256 materials (uint8 voxel map); this "
                f"source declares {self.cross_sections.n_materials}"
            )
        # Every voxel's material must have a table row: the kernels index the rows by the
        # material map without a bounds check.
        n_rows = int(self.cross_sections.n_materials)
        top = int(np.max(np.asarray(self.grid.material))) if np.asarray(self.grid.material).size else 0
        if top >= n_rows:
            raise ValueError(
                f"voxel material index {top} has no row in the cross-section source "
                f"({n_rows} material{'s' if n_rows != 1 else ''})"
            )
        self._index = _device_index(self.device)
        self._ctx = nat.Context(self._index)
        self._info = nat.device_info(self._index)
        self.launch = {"round_lanes_per_cu": 2048, "refill_block": 64, "push_chunk": 64, "use_graphs": True}
        self._population_token = None
        self._ray_model = None

    def close(self) -> None:
        if self._ctx is not None:
            self._ctx.close()

    # ------------------------------------------------------------------ helpers

    @property
    def device_name(self) -> str:
        return self._info["name"] if self._info else "?"

    def _provenance(self, seed, pcut, ecut, msc_model, step, deposit_resolution_cm) -> RunProvenance:
        return RunProvenance(
            version=VERSION,
            backend="hip",
            device=f"hip:{self._index}",
            seed=int(seed),
            pcut_mev=pcut,
            ecut_mev=ecut,
            msc_model=msc_model,
            step_energy_fraction=step,
            cross_sections=self.cross_sections.provenance,
            deposit_resolution_cm=deposit_resolution_cm,
        )

    def _chunk(self, total: int) -> int:
        if self.chunk_size is not None:
            return int(min(self.chunk_size, max(total, 1)))
        target = min(_CHUNK_HISTORIES_PER_CU * max(self._info["cu_count"], 1), _CHUNK_SIZE_CAP)
        # Queue memory is budgeted against what is free now (other engines may be alive).
        free = nat.device_free_bytes(self._index)
        if free > 0:
            target = min(target, int(free * _CHUNK_MEMORY_FRACTION) // (self.queue_factor * _QUEUE_BYTES_PER_SLOT))
        return int(max(1, min(target, max(total, 1))))

    def _upload_grid(self) -> None:
        if self._grid_uploaded:
            return
        g = self.grid
        density = _f32(g.density)
        material = np.ascontiguousarray(np.asarray(g.material).astype(np.uint8))
        origin = np.asarray(g.origin, dtype=np.float64)
        spacing = np.asarray(g.spacing, dtype=np.float64)
        self._ctx.call(
            "hmc_set_grid", int(g.shape[0]), int(g.shape[1]), int(g.shape[2]),
            nat.ptr(origin, nat.c_double), nat.ptr(spacing, nat.c_double),
            nat.ptr(density, nat.ctypes.c_float), nat.ptr(material, nat.c_uint8),
        )
        self._grid_uploaded = True

    def _host_tables(self, e_max: float, pcut: float, ecut: float):
        key = (ecut, pcut, e_max)
        with self._lock:
            host = self._table_cache.get(key)
            if host is None:
                host = self.cross_sections.build_tables(ecut=ecut, pcut=pcut, e_max=e_max * E_MAX_MARGIN)
                self._table_cache[key] = host
        return host

    def _upload_tables(self, e_max: float, pcut: float, ecut: float, with_gs: bool) -> None:
        host = self._host_tables(e_max, pcut, ecut)
        key = (ecut, pcut, e_max, with_gs)
        if self._uploaded_tables == key:
            return
        keep = {}

        def f(name, arr=None):
            a = _f32(getattr(host, name) if arr is None else arr)
            keep[name] = a
            return nat.ptr(a, nat.ctypes.c_float)

        n_mat = int(np.asarray(host.mu_compton).reshape(-1, host.n_points).shape[0])
        t = nat.Tables()
        t.n_points = int(host.n_points)
        t.n_mat = n_mat
        t.n_coherent = int(host.n_coherent)
        for name in (
            "majorant", "mu_compton", "mu_photo", "mu_pair", "mu_rayleigh", "csda_range", "moller",
            "restricted_range", "stopping_radiative", "scattering_power", "log_eta",
            "stopping_restricted", "energy_of_restricted_range", "coherent_x", "coherent_cumulative",
        ):
            setattr(t, name, f(name))
        t.p_log_min = float(host.photon_log_e_min)
        t.p_inv_dlog = float(host.photon_inv_dlog)
        t.e_log_min = float(host.electron_log_e_min)
        t.e_inv_dlog = float(host.electron_inv_dlog)
        t.r_log_min = float(host.range_log_r_min)
        t.r_inv_dlog = float(host.range_inv_dlog)
        self._ctx.call("hmc_set_tables", nat.ctypes.byref(t))
        if with_gs:
            present = np.unique(np.asarray(self.grid.material)).astype(int).tolist()
            gs = build_gs_grid(*gs_window_from_tables(host, present))
            values = np.ascontiguousarray(gs.values, dtype=np.float32)
            self._ctx.call(
                "hmc_set_gs", nat.ptr(values, nat.ctypes.c_float), int(values.shape[0]),
                int(values.shape[1]), int(values.shape[2]), int(gs.ix0), int(gs.iy0),
            )
        self._uploaded_tables = key

    def _free_for_maps(self) -> int:
        """Device memory the tally maps and sums may take: what is free now (other
        engines, torch and this engine's queues may hold memory taken after
        construction) plus what this engine's current maps and sums hold, which a new
        layout reuses or releases."""
        free = nat.device_free_bytes(self._index)
        return free + self._map_bytes if free > 0 else 0

    def _batch_slots(self, n_batches: int, n_columns: int, n_vox: int) -> int:
        """Batches transported together: each holds its own int64 map of the group's
        columns, folded in batch order afterwards (so any slot count gives the same
        bits). As many as a quarter of the memory available for maps allows."""
        free = self._free_for_maps()
        per_slot = 8 * n_columns * n_vox
        fit = int(free * _SLOT_MEMORY_FRACTION) // per_slot if free > 0 else 1
        return max(1, min(n_batches, fit))

    def _default_scoring(self) -> ScoringGrid:
        if self._scoring_default is None:
            self._scoring_default = ScoringGrid.for_grid(self.grid)
        return self._scoring_default

    def _set_scoring(self, scoring, n_columns: int, n_slots: int = 1) -> None:
        # Same grid object and layout as the previous call: only clear the maps (the
        # voxel masses and buffers on the device are already right). Grids and scoring
        # grids are treated as immutable once handed to the engine.
        key = (n_columns, n_slots)
        if self._scoring_last is not None and self._scoring_last[0] is scoring and self._scoring_last[1] == key:
            self._ctx.call("hmc_reset_maps")
            return
        mass = np.ascontiguousarray(np.asarray(scoring.voxel_mass, dtype=np.float64))
        origin = np.asarray(scoring.origin, dtype=np.float64)
        spacing = np.asarray(scoring.spacing, dtype=np.float64)
        self._ctx.call(
            "hmc_set_scoring", int(scoring.shape[0]), int(scoring.shape[1]), int(scoring.shape[2]),
            nat.ptr(origin, nat.c_double), nat.ptr(spacing, nat.c_double),
            nat.ptr(mass, nat.c_double), int(n_columns), int(n_slots),
        )
        self._scoring_last = (scoring, key)
        # int64 maps of every slot plus the float64 sums s1 and s2
        self._map_bytes = max(self._map_bytes, 8 * (n_slots + 2) * n_columns * int(scoring.n_voxels))

    def _set_physics(self, pcut, ecut, step, resolution, transport_electrons, msc_gs, d2w) -> None:
        p = nat.Physics()
        p.pcut = float(pcut)
        p.ecut = float(ecut)
        p.step_energy_fraction = float(step)
        p.deposit_resolution_cm = float(resolution)
        p.transport_electrons = 1 if transport_electrons else 0
        p.msc_gs = int(msc_gs)
        p.dose_to_water = 1 if d2w else 0
        self._ctx.call("hmc_set_physics", nat.ctypes.byref(p))

    def _set_source(self, source, primary_kind: str, beamlets: range | None = None) -> tuple[bool, float]:
        """Describe the source to the generator; returns (analytic_emitted, energy).
        Sources without an in-kernel generator are set up for host pre-sampling: their
        primaries are uploaded per transport call (``_upload_primaries``)."""
        inner, model = _weighted_parts(source)
        if model is not None and _in_kernel(source):
            self._set_source(inner, primary_kind, beamlets)
            self._set_ray_weights(model)
            return False, 0.0  # emitted energy is booked per weighted primary on the device
        self._set_ray_weights(None)
        d = nat.SourceDesc()
        name = _type_name(source)
        if not _in_kernel(source):
            sampler = "sample_beamlet_batch" if beamlets is not None else "sample_batch"
            if not callable(getattr(source, sampler, None)):
                raise ValueError(
                    f"source type {name} has no in-kernel generator in this engine and no "
                    f"{sampler}() for host pre-sampling; in-kernel types: "
                    + ", ".join(sorted(_BUILTIN_SOURCES))
                )
            d.kind = nat.SRC_UPLOADED
            d.particle = 0
            self._ctx.call("hmc_set_source", nat.ctypes.byref(d))
            return False, 0.0
        if name == "DevicePhaseSpace":
            self._upload_population(source)
            d.kind = nat.SRC_RESIDENT
            d.particle = 0
            self._ctx.call("hmc_set_source", nat.ctypes.byref(d))
            return False, 0.0
        d, keep, analytic, energy = generator_desc(source, primary_kind, beamlets)
        self._ctx.call("hmc_set_source", nat.ctypes.byref(d))
        del keep  # the arrays only had to outlive the call
        return analytic, energy

    def _set_ray_weights(self, model) -> None:
        """Hand the generator a collimator (StackTransmission) or mask
        (MaskTransmission) factor model, or none; a model already set is not sent
        again."""
        if model is self._ray_model:
            return
        self._ray_model = None
        if model is None:
            self._ctx.call("hmc_set_ray_weights", None)
            return
        d = nat.RayWeightsDesc()
        keep = []

        def arr(a, dtype, ctype):
            a = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))
            keep.append(a)
            return nat.ptr(a, ctype)

        name = type(model).__name__
        if name == "StackTransmission":
            cs = model.stack.compile()
            d.kind = 1
            d.n_devices = int(cs.kind.size)
            d.origin[:] = [float(c) for c in cs.origin]
            d.basis[:] = [float(c) for c in cs.basis.reshape(-1)]
            for field_name, values in (("device_kind", cs.kind), ("flags", cs.flags), ("n_pairs", cs.n_pairs),
                                       ("edge_offset", cs.offset), ("pair_offset", cs.pair_offset)):
                setattr(d, field_name, arr(values, np.int32, nat.c_int32))
            for field_name, values in (("z_top", cs.z_top), ("z_bottom", cs.z_bottom), ("edge_neg", cs.edge_neg),
                                       ("edge_pos", cs.edge_pos), ("density", model.densities)):
                setattr(d, field_name, arr(values, np.float64, nat.c_double))
            d.n_leaf_edges, d.n_tips = int(cs.leaf_edges.size), int(cs.tips_pos.size)
            d.leaf_edges = arr(cs.leaf_edges, np.float64, nat.c_double)
            d.tips_neg = arr(cs.tips_neg, np.float64, nat.c_double)
            d.tips_pos = arr(cs.tips_pos, np.float64, nat.c_double)
            d.n_table = int(model.log_energies.size)
            d.log_energies = arr(model.log_energies, np.float64, nat.c_double)
            d.log_mu = arr(model.log_mu, np.float64, nat.c_double)
        elif name == "MaskTransmission":
            d.kind = 2
            d.center[:] = [float(c) for c in model.center]
            d.u_axis[:] = [float(c) for c in model.u_axis]
            d.v_axis[:] = [float(c) for c in model.v_axis]
            d.width_u, d.width_v = model.widths
            d.mask_nu, d.mask_nv = (int(n) for n in model.mask.shape)
            d.mask = arr(model.mask, np.float64, nat.c_double)
        else:
            raise ValueError(f"no device evaluation for the ray weight model {name}")
        self._ctx.call("hmc_set_ray_weights", nat.ctypes.byref(d))
        self._ray_model = model

    def _upload_population(self, source) -> None:
        """Copy a DevicePhaseSpace's records to the device, once per population: the
        context keeps the last population written, so running the same one again (a
        loop over seeds) uploads nothing."""
        if self._population_token == source.token:
            return
        self._population_token = None
        n, offset = len(source), 0
        for cols in source.population_blocks():
            kind = _iaea_kinds(cols["particle_type"])
            data = [np.ascontiguousarray(np.asarray(cols[c], dtype=np.float32).reshape(-1)) for c in _UPLOAD_COLUMNS]
            if any(c.shape != kind.shape for c in data) or offset + kind.size > n:
                raise ValueError("phase-space population blocks differ from its length")
            self._ctx.call("hmc_set_population", int(n), int(offset), int(kind.size), nat.ptr(kind, nat.c_int32),
                           *[nat.ptr(c, nat.ctypes.c_float) for c in data])
            offset += kind.size
        if offset != n:
            raise ValueError(f"phase-space population gave {offset} records, expected {n}")
        self._population_token = source.token

    def _upload_primaries(self, samples: list[dict]) -> None:
        """Upload host-sampled primaries (``sample_batch`` columns, IAEA particle codes
        1 photon, 2 electron, 3 positron), concatenated in work-item order."""
        kind = _iaea_kinds(np.concatenate([np.asarray(b["particle_type"]).reshape(-1) for b in samples]))
        cols = [np.ascontiguousarray(np.concatenate(
            [np.asarray(b[c], dtype=np.float32).reshape(-1) for b in samples])) for c in _UPLOAD_COLUMNS]
        if any(c.shape != kind.shape for c in cols):
            raise ValueError("sampled primary columns differ in length")
        self._ctx.call("hmc_set_uploaded", int(kind.size), nat.ptr(kind, nat.c_int32),
                       *[nat.ptr(c, nat.ctypes.c_float) for c in cols])

    def _presampled_transport(self, seed: int, items: list[tuple[int, int, int, int]], chunk: int,
                              sample, books: dict) -> None:
        """Transport host-sampled primaries a bounded piece at a time: each work item is
        cut into sampling blocks of ``_SAMPLE_BLOCK`` histories aligned to its first key
        (so what is sampled depends on the keys only, not on slots or chunk size), and
        blocks go to the device in calls of at most ``_SAMPLE_CALL`` primaries, all
        tallying into the same slots before their fold. Host and device memory for
        primaries stay bounded whatever the run's size (the reference samples one chunk
        at a time, pyradmc/backends/warp/engine.py:2506-2521). ``sample(key, n, beamlet)``
        returns one block's columns."""
        pieces = []
        for key, count, column, beamlet in items:
            for start in range(0, int(count), _SAMPLE_BLOCK):
                pieces.append((int(key) + start, min(_SAMPLE_BLOCK, int(count) - start), column, beamlet))
        call, size = [], 0
        for piece in pieces + [None]:
            if call and (piece is None or size + piece[1] > _SAMPLE_CALL):
                self._upload_primaries([sample(key, n, beamlet) for key, n, _, beamlet in call])
                self._merge(books, self._transport(seed, call, chunk))
                call, size = [], 0
            if piece is not None:
                call.append(piece)
                size += piece[1]

    def _transport(self, seed: int, items: list[tuple[int, int, int, int]], chunk: int) -> dict:
        """items: (key_base, count, flat tally column, beamlet parameter index)."""
        arr = (nat.WorkItem * len(items))()
        total = 0
        for k, (key_base, count, column, beamlet) in enumerate(items):
            arr[k].key_base = int(key_base)
            arr[k].count = int(count)
            arr[k].column = int(column)
            arr[k].beamlet = int(beamlet)
            total += int(count)
        ledger = nat.Ledger()
        self._ctx.call(
            "hmc_transport", nat.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), arr, len(items),
            int(chunk), int(self.queue_factor), nat.ctypes.byref(ledger),
        )
        return ledger.as_dict()

    def _fold(self, n_slots: int, nhist: np.ndarray) -> None:
        """Fold slots 0..n_slots-1 in order; nhist is [slot][column]. Returns once the
        fold is queued: the folded quanta come back with the next finalize
        (``_folded_quanta``)."""
        nh = np.ascontiguousarray(np.asarray(nhist, dtype=np.float64).reshape(-1))
        self._ctx.call("hmc_fold", int(n_slots), nat.ptr(nh, nat.c_double), None)

    def _folded_quanta(self) -> int:
        """Quanta folded since the last call (read back by the finalize after them)."""
        total = nat.c_int64()
        self._ctx.call("hmc_folded_quanta", nat.ctypes.byref(total))
        return int(total.value)

    @staticmethod
    def _merge(books: dict, ledger: dict) -> None:
        for k, v in ledger.items():
            books[k] = books.get(k, 0) + v

    def set_profiling(self, enabled: bool) -> None:
        """Time every transport launch with HIP events (per-kernel ``*_kernel_ms`` in
        the result's ``counters``); a measurement hook, off by default."""
        self._ctx.call("hmc_set_profiling", 1 if enabled else 0)

    def set_launch(self, round_lanes_per_cu: int = 2048, refill_block: int = 64, push_chunk: int = 64,
                   use_graphs: bool = True) -> None:
        """Launch shape of the transport rounds: lanes per CU of each round's grid,
        queue indices a wave claims at a time, output slots it reserves at a time, and
        whether groups of rounds replay a captured HIP graph. Results do not depend on
        them (each particle carries its own random stream; tallies are integers);
        ``pyradmc_amd.autotune`` picks them."""
        self._ctx.call("hmc_set_launch", int(round_lanes_per_cu), int(refill_block), int(push_chunk),
                       1 if use_graphs else 0)
        self.launch = {"round_lanes_per_cu": int(round_lanes_per_cu), "refill_block": int(refill_block),
                       "push_chunk": int(push_chunk), "use_graphs": bool(use_graphs)}

    def autotune(self, source, n_histories: int, n_batches: int, seed: int = 1, **kw):
        """Tune chunk size and launch shape on a timed ``run`` of this workload and apply
        the best (``pyradmc_amd.autotune.tune_run``); returns the chosen configuration."""
        from pyradmc_amd import autotune

        return autotune.tune_run(self, source, n_histories, n_batches, seed=seed, **kw)

    def autotune_dij(self, source, n_histories_per_beamlet, n_batches: int, seed: int = 1, **kw):
        """Tune on a timed ``run_dij`` of this workload (``pyradmc_amd.autotune.tune_dij``)
        and apply the best; returns the chosen configuration."""
        from pyradmc_amd import autotune

        return autotune.tune_dij(self, source, n_histories_per_beamlet, n_batches, seed=seed, **kw)

    # --------------------------------------------------------------------- run

    def run(
        self,
        source,
        n_histories: int,
        n_batches: int,
        seed: int,
        pcut: float = PCUT_MEV,
        ecut: float = ECUT_MEV,
        transport_electrons: bool = True,
        primary_kind: str = "photon",
        scoring_grid: ScoringGrid | None = None,
        scoring_mode: str = "dose_to_medium",
        step_energy_fraction: float | None = None,
        deposit_resolution_cm: float | None = None,
        msc_model: str = "gs",
        progress=None,
        concurrent_batches: int = 1,
    ) -> TransportResult:
        """Transport ``n_histories`` primaries in ``n_batches`` equal batches.

        ``concurrent_batches`` keeps the reference's signature: this engine already
        transports as many batches side by side as the tally memory allows (one map slot
        each, folded in batch order), so it is validated and otherwise has no effect on
        the schedule or the results."""
        if n_histories < 1:
            raise ValueError(f"need at least one history, got {n_histories}")
        if concurrent_batches < 1:
            raise ValueError(f"need at least one lane, got concurrent_batches={concurrent_batches}")
        if n_batches < 1:
            raise ValueError(f"need at least one batch, got {n_batches}")
        if n_histories % n_batches != 0:
            raise ValueError(
                f"n_histories={n_histories} not divisible by n_batches={n_batches}; "
                "unequal batches would weight batch means inconsistently"
            )
        if primary_kind not in ("photon", "electron"):
            raise ValueError(f"unknown primary_kind {primary_kind!r}")
        d2w = _validate_scoring_mode(scoring_mode, transport_electrons)
        msc_gs = _validate_msc(msc_model)
        step = default_step_energy_fraction(msc_model) if step_energy_fraction is None else float(step_energy_fraction)
        if deposit_resolution_cm is not None and deposit_resolution_cm <= 0.0:
            raise ValueError(f"deposit resolution must be positive, got {deposit_resolution_cm}")
        resolution = 0.0 if deposit_resolution_cm is None else float(deposit_resolution_cm)

        scoring = scoring_grid if scoring_grid is not None else self._default_scoring()
        with self._lock:
            self._upload_grid()
        self._upload_tables(float(source.max_energy), pcut, ecut, with_gs=msc_gs == 1)
        self._set_physics(pcut, ecut, step, resolution, transport_electrons, msc_gs, d2w)
        analytic, mono_energy = self._set_source(source, primary_kind)
        presample = not _in_kernel(source)
        if _type_name(source) == "DevicePhaseSpace":
            source.note_histories(n_histories)
        slots = self._batch_slots(n_batches, 1, scoring.n_voxels)
        self._set_scoring(scoring, 1, slots)

        per_batch = n_histories // n_batches
        chunk = self._chunk(per_batch * slots)
        emitter = ProgressEmitter(progress, n_histories)
        books: dict = {}
        self._folded_quanta()  # drop any count an interrupted earlier run left behind
        for b0 in range(0, n_batches, slots):
            block = range(b0, min(b0 + slots, n_batches))
            items = [(b * per_batch, per_batch, b - b0, 0) for b in block]
            if presample:
                self._presampled_transport(seed, items, chunk, lambda key, n, _: source.sample_batch(seed, key, n),
                                           books)
            else:
                self._merge(books, self._transport(seed, items, chunk))
            self._fold(len(block), np.full(len(block), float(per_batch)))
            for _ in block:
                emitter.tick(per_batch)
        n_vox = scoring.n_voxels
        mean = host_array(n_vox, np.float64)
        sigma = host_array(n_vox, np.float64)
        self._ctx.call("hmc_finalize", int(n_batches), nat.ptr(mean, nat.c_double), nat.ptr(sigma, nat.c_double))
        deposited_quanta = self._folded_quanta()

        if analytic:
            particle_rest = 0.0
            energy_emitted = n_histories * mono_energy + particle_rest
        else:
            energy_emitted = float(books.get("emitted", 0)) * ENERGY_QUANTUM_MEV
        energy_deposited = (
            float(books.get("deposited", 0)) * ENERGY_QUANTUM_MEV if d2w else float(deposited_quanta) * ENERGY_QUANTUM_MEV
        )
        return TransportResult(
            dose=mean.reshape(scoring.shape),
            dose_sigma=sigma.reshape(scoring.shape),
            energy_emitted=energy_emitted,
            energy_deposited=energy_deposited,
            energy_escaped=float(books.get("escaped", 0)) * ENERGY_QUANTUM_MEV,
            energy_unscored=float(books.get("unscored", 0)) * ENERGY_QUANTUM_MEV,
            n_histories=n_histories,
            n_batches=n_batches,
            scoring_mode=scoring_mode,
            provenance=self._provenance(seed, pcut, ecut, msc_model, step, deposit_resolution_cm),
            counters=books,
        )

    # ----------------------------------------------------------------- run_dij

    def _prepare(self, source, pcut: float = PCUT_MEV, ecut: float = ECUT_MEV, step: float | None = None,
                 resolution: float = 0.0, transport_electrons: bool = True, msc_gs: int = 1,
                 d2w: bool = False) -> None:
        """Grid, cross-section tables (and GS grid) and physics on the device for a
        source: what every transport needs before its first chunk."""
        if step is None:
            step = default_step_energy_fraction("gs" if msc_gs else "gaussian")
        with self._lock:
            self._upload_grid()
        self._upload_tables(float(source.max_energy), pcut, ecut, with_gs=msc_gs == 1)
        self._set_physics(pcut, ecut, step, resolution, transport_electrons, msc_gs, d2w)

    def _dij_group(self, source, group: range, counts: np.ndarray, n_batches: int, seed: int,
                   correlated: bool, scoring, books: dict) -> tuple[bool, float, int]:
        """Transport, fold and finalize one group of beamlet columns; the means and
        sigmas stay on the device (for truncation, or for the denoiser's windows).
        Returns (analytic emitted energy?, mono energy, deposited quanta)."""
        per_batch = counts // n_batches
        offsets = np.concatenate(([0], np.cumsum(counts)[:-1]))
        g0 = group.start
        n_cols = len(group)
        analytic, mono_energy = self._set_source(source, "photon", beamlets=group)
        slots = self._batch_slots(n_batches, n_cols, scoring.n_voxels)
        self._set_scoring(scoring, n_cols, slots)
        chunk = self._chunk(int(per_batch[g0:group.stop].sum()) * slots)
        self._folded_quanta()  # drop any count an interrupted earlier call left behind
        for b0 in range(0, n_batches, slots):
            block = range(b0, min(b0 + slots, n_batches))
            items = []
            for b in block:
                for j in group:
                    r0 = b * int(per_batch[j])
                    key = r0 if correlated else int(offsets[j]) + r0
                    items.append((key, int(per_batch[j]), (b - b0) * n_cols + (j - g0), j - g0))
            if not _in_kernel(source):
                self._presampled_transport(
                    seed, items, chunk, lambda key, n, local: source.sample_beamlet_batch(seed, key, n, g0 + local),
                    books)
            else:
                self._merge(books, self._transport(seed, items, chunk))
            nhist = np.tile(per_batch[g0:group.stop].astype(np.float64), len(block))
            self._fold(len(block), nhist)
        self._ctx.call("hmc_finalize", int(n_batches), None, None)
        return analytic, mono_energy, self._folded_quanta()

    def run_dij(
        self,
        source,
        n_histories_per_beamlet,
        n_batches: int,
        seed: int,
        pcut: float = PCUT_MEV,
        ecut: float = ECUT_MEV,
        transport_electrons: bool = True,
        truncation: float = DIJ_TRUNCATION_RELATIVE,
        correlated: bool = True,
        beamlet_group_size: int | None = None,
        scoring_grid: ScoringGrid | None = None,
        scoring_mode: str = "dose_to_medium",
        step_energy_fraction: float | None = None,
        deposit_resolution_cm: float | None = None,
        msc_model: str = "gs",
        devices=None,
        concurrent_batches: int = 1,
        progress=None,
        value_dtype=np.float64,
    ) -> DijResult:
        """Beamlet-resolved dose matrix over the source's beamlets, CSC by column.

        ``value_dtype`` (an extension; float64 as in the reference by default) may be
        ``np.float32`` for the CSC dose and sigma arrays, halving a large Dij's host
        footprint; indices stay int64 and the device statistics stay float64. Arrays of
        1 MiB or more are page-locked host memory (released with their last view).

        ``counters`` holds the summed transport ledgers (integer event counts and
        quanta) and, as floats, the host wall split of this call: ``transport_s``
        (transport, fold and finalize), ``assemble_s`` (truncation, download, join) and
        ``groups`` (beamlet groups run)."""
        if not hasattr(source, "n_beamlets"):
            raise TypeError(
                f"run_dij needs a beamlet source; {_type_name(source)} has no beamlets "
                "(phase-space and other whole-field sources go through run)")
        value_dtype = np.dtype(value_dtype)
        if value_dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError(f"value_dtype must be float64 or float32, got {value_dtype}")
        narrow = value_dtype == np.dtype(np.float32)
        n_beamlets = int(source.n_beamlets)
        counts = np.asarray(n_histories_per_beamlet, dtype=np.int64)
        uneven = counts.ndim == 1
        if uneven:
            if counts.shape != (n_beamlets,):
                raise ValueError(f"per-beamlet counts shape {counts.shape} != ({n_beamlets},)")
        else:
            counts = np.full(n_beamlets, int(counts), dtype=np.int64)
        if n_batches < 1:
            raise ValueError(f"need at least one batch, got {n_batches}")
        if np.any(counts < 1):
            raise ValueError(f"need at least one history per beamlet, got {int(counts.min())}")
        if np.any(counts % n_batches != 0):
            raise ValueError(
                f"n_histories_per_beamlet not divisible by n_batches={n_batches}; "
                "unequal batches would weight batch means inconsistently"
            )
        if beamlet_group_size is not None and beamlet_group_size < 1:
            raise ValueError(f"need a positive beamlet group size, got {beamlet_group_size}")
        if concurrent_batches < 1:
            raise ValueError(f"need at least one lane, got concurrent_batches={concurrent_batches}")
        if truncation < 0.0:
            raise ValueError(f"negative truncation {truncation}")
        if devices is not None:
            wanted = [_device_index(d) for d in devices]
            if len(set(wanted)) != 1 or wanted[0] != self._index:
                raise ValueError("this engine runs on one GPU; devices must name only its own device")
        d2w = _validate_scoring_mode(scoring_mode, transport_electrons)
        msc_gs = _validate_msc(msc_model)
        step = default_step_energy_fraction(msc_model) if step_energy_fraction is None else float(step_energy_fraction)
        if deposit_resolution_cm is not None and deposit_resolution_cm <= 0.0:
            raise ValueError(f"deposit resolution must be positive, got {deposit_resolution_cm}")
        resolution = 0.0 if deposit_resolution_cm is None else float(deposit_resolution_cm)

        scoring = scoring_grid if scoring_grid is not None else self._default_scoring()
        n_vox = scoring.n_voxels
        if beamlet_group_size is None:
            # float64 sums plus one int64 map per group column; batch slots come after.
            free = self._free_for_maps()
            fit = int(free * _GROUP_MEMORY_FRACTION) // (24 * n_vox) if free > 0 else 128
            beamlet_group_size = max(1, min(fit, n_beamlets, 4096))

        self._prepare(source, pcut, ecut, step, resolution, transport_electrons, msc_gs, d2w)

        emitter = ProgressEmitter(progress, int(counts.sum()))
        books: dict = {}
        deposited_quanta = 0
        col_counts, col_idx, col_dose, col_sigma = [], [], [], []
        analytic, mono_energy = False, 0.0
        wall = {"groups": 0, "transport_s": 0.0, "assemble_s": 0.0}
        for g0 in range(0, n_beamlets, beamlet_group_size):
            group = range(g0, min(g0 + beamlet_group_size, n_beamlets))
            group_hist = int(counts[g0:group.stop].sum())
            t0 = time.perf_counter()
            analytic, mono_energy, q = self._dij_group(source, group, counts, n_batches, seed,
                                                      correlated, scoring, books)
            t1 = time.perf_counter()
            deposited_quanta += q
            kept = np.zeros(len(group), dtype=np.int64)
            total = nat.c_int64()
            self._ctx.call("hmc_truncate", float(truncation), nat.ptr(kept, nat.c_int64), nat.ctypes.byref(total))
            m = int(total.value)
            idx = host_array(m, np.int64)
            if narrow:  # converted on the device: half the value bytes over PCIe
                dose = host_array(m, np.float32)
                sig = host_array(m, np.float32)
                if m:
                    self._ctx.call("hmc_fetch_csc_f32", nat.ptr(idx, nat.c_int64), nat.ptr(dose, nat.ctypes.c_float),
                                   nat.ptr(sig, nat.ctypes.c_float))
            else:
                dose = host_array(m, np.float64)
                sig = host_array(m, np.float64)
                if m:
                    self._ctx.call("hmc_fetch_csc", nat.ptr(idx, nat.c_int64), nat.ptr(dose, nat.c_double),
                                   nat.ptr(sig, nat.c_double))
            col_counts.append(kept)
            col_idx.append(idx)
            col_dose.append(dose.astype(value_dtype, copy=False))
            col_sigma.append(sig.astype(value_dtype, copy=False))
            wall["groups"] += 1
            wall["transport_s"] += t1 - t0
            wall["assemble_s"] += time.perf_counter() - t1
            emitter.tick(group_hist)

        t2 = time.perf_counter()
        indptr = np.zeros(n_beamlets + 1, dtype=np.int64)
        np.cumsum(np.concatenate(col_counts), out=indptr[1:])
        indices, dose_all, sigma_all = _joined(col_idx, np.int64), _joined(col_dose, value_dtype), \
            _joined(col_sigma, value_dtype)
        wall["assemble_s"] += time.perf_counter() - t2
        books.update(wall)
        if analytic:
            energy_emitted = float(counts.sum()) * mono_energy
        else:
            energy_emitted = float(books.get("emitted", 0)) * ENERGY_QUANTUM_MEV
        energy_deposited = (
            float(books.get("deposited", 0)) * ENERGY_QUANTUM_MEV if d2w else float(deposited_quanta) * ENERGY_QUANTUM_MEV
        )
        return DijResult(
            grid_shape=tuple(scoring.shape),
            n_beamlets=n_beamlets,
            n_histories_per_beamlet=counts.copy() if uneven else int(counts[0]),
            n_batches=n_batches,
            truncation=float(truncation),
            indptr=indptr,
            indices=indices,
            dose=dose_all,
            sigma=sigma_all,
            energy_emitted=energy_emitted,
            energy_deposited=energy_deposited,
            energy_escaped=float(books.get("escaped", 0)) * ENERGY_QUANTUM_MEV,
            energy_unscored=float(books.get("unscored", 0)) * ENERGY_QUANTUM_MEV,
            correlated=bool(correlated),
            scoring_mode=scoring_mode,
            provenance=self._provenance(seed, pcut, ecut, msc_model, step, deposit_resolution_cm),
            counters=books,
        )
