"""ctypes binding of ``libhipmc.so`` (declarations in ``include/hipmc.h``).

Loads the in-tree library and fails loudly when it is missing: the engine has no
CPU fallback. Status codes map onto the exception types the reference engine raises
(ValueError for bad arguments, RuntimeError for in-kernel failures and call order,
MemoryError for device allocation).

The PyTorch wheel carries its own HIP runtime. A process that loads this library first
and torch afterwards ends up with two runtimes, and torch then sees no device. So when
torch is installed, it is imported before the library is loaded: the library then binds
to torch's runtime and the two share devices, streams and memory.
"""

from __future__ import annotations

import atexit
import ctypes
import threading
import weakref
from ctypes import POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_uint8, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhipmc.so"

HMC_OK = 0
HMC_EINVAL = -1
HMC_EHIP = -2
HMC_EOVERFLOW = -3
HMC_EMAJORANT = -4
HMC_ENOMEM = -5
HMC_ESTATE = -6

SRC_PENCIL = 0
SRC_PARALLEL = 1
SRC_PARALLEL_SPECTRAL = 2
SRC_SPECTRAL_FAN = 3
SRC_LATTICE = 4
SRC_UPLOADED = 5
SRC_RESIDENT = 6
SRC_SPOT_FAN = 7

_f32p = POINTER(ctypes.c_float)
_f64p = POINTER(c_double)


class Tables(Structure):
    _fields_ = [
        ("n_points", c_int),
        ("n_mat", c_int),
        ("n_coherent", c_int),
        ("majorant", _f32p),
        ("mu_compton", _f32p),
        ("mu_photo", _f32p),
        ("mu_pair", _f32p),
        ("mu_rayleigh", _f32p),
        ("csda_range", _f32p),
        ("moller", _f32p),
        ("restricted_range", _f32p),
        ("stopping_radiative", _f32p),
        ("scattering_power", _f32p),
        ("log_eta", _f32p),
        ("stopping_restricted", _f32p),
        ("energy_of_restricted_range", _f32p),
        ("coherent_x", _f32p),
        ("coherent_cumulative", _f32p),
        ("p_log_min", c_double),
        ("p_inv_dlog", c_double),
        ("e_log_min", c_double),
        ("e_inv_dlog", c_double),
        ("r_log_min", c_double),
        ("r_inv_dlog", c_double),
    ]


_i32p = POINTER(c_int32)


class RayWeightsDesc(Structure):
    """hmc_ray_weights: the generator's collimator or mask factor model."""

    _fields_ = [
        ("kind", c_int),
        ("n_devices", c_int),
        ("origin", c_double * 3),
        ("basis", c_double * 9),
        ("device_kind", _i32p),
        ("flags", _i32p),
        ("n_pairs", _i32p),
        ("edge_offset", _i32p),
        ("pair_offset", _i32p),
        ("z_top", _f64p),
        ("z_bottom", _f64p),
        ("edge_neg", _f64p),
        ("edge_pos", _f64p),
        ("density", _f64p),
        ("n_leaf_edges", c_int),
        ("n_tips", c_int),
        ("leaf_edges", _f64p),
        ("tips_neg", _f64p),
        ("tips_pos", _f64p),
        ("n_table", c_int),
        ("log_energies", _f64p),
        ("log_mu", _f64p),
        ("center", c_double * 3),
        ("u_axis", c_double * 3),
        ("v_axis", c_double * 3),
        ("width_u", c_double),
        ("width_v", c_double),
        ("mask_nu", c_int),
        ("mask_nv", c_int),
        ("mask", _f64p),
    ]


class PresolveDesc(Structure):
    """hmc_presolve: the head pre-solve's settings and attenuation tables."""

    _fields_ = [
        ("first_compton", c_int),
        ("has_air", c_int),
        ("exit_z", c_double),
        ("air_z_start", c_double),
        ("air_z_end", c_double),
        ("pcut", c_double),
        ("ecut", c_double),
        ("weight_floor", c_double),
        ("origin", c_double * 3),
        ("basis", c_double * 9),
        ("n_table", c_int),
        ("n_rows", c_int),
        ("log_energies", _f64p),
        ("log_mu_total", _f64p),
        ("log_mu_compton", _f64p),
    ]


class SourceDesc(Structure):
    _fields_ = [
        ("kind", c_int),
        ("particle", c_int),
        ("energy", c_double),
        ("position", c_double * 3),
        ("direction", c_double * 3),
        ("z0", c_double),
        ("x0", c_double),
        ("x_extent", c_double),
        ("y0", c_double),
        ("y_extent", c_double),
        ("focal", c_double * 3),
        ("u_axis", c_double * 3),
        ("v_axis", c_double * 3),
        ("width_u", c_double),
        ("width_v", c_double),
        ("n_bins", c_int),
        ("edges", _f64p),
        ("cdf", _f64p),
        ("n_beamlets", c_int),
        ("beamlet_params", _f64p),
        ("sigma_u", c_double),
        ("sigma_v", c_double),
    ]


class Physics(Structure):
    _fields_ = [
        ("pcut", c_double),
        ("ecut", c_double),
        ("step_energy_fraction", c_double),
        ("deposit_resolution_cm", c_double),
        ("transport_electrons", c_int),
        ("msc_gs", c_int),
        ("dose_to_water", c_int),
        ("reserved", c_int),
    ]


class WorkItem(Structure):
    _fields_ = [("key_base", c_int64), ("count", c_int64), ("column", c_int32), ("beamlet", c_int32)]


class SamplerArgs(Structure):
    """Inputs of ``hmc_sample_physics`` (include/hipmc.h documents the output columns)."""

    _fields_ = [
        ("path_mu", c_double), ("channel_mu", c_double * 4), ("compton_energy", c_double * 2),
        ("moller_energy", c_double), ("moller_cut", c_double), ("hinge_ms", c_double),
        ("brems_energy", c_double), ("brems_radiative_loss", c_double), ("brems_pcut", c_double),
        ("coherent_energy", c_double), ("coherent_material", c_int), ("reserved", c_int),
        ("gs_log_eta", c_double), ("gs_ms", c_double * 2), ("rotate_cos", c_double), ("rotate_phi", c_double),
    ]


SAMPLE_COLUMNS = 22


class Ledger(Structure):
    _fields_ = [
        (name, c_int64)
        for name in (
            "escaped", "unscored", "deposited", "emitted", "violations", "trapped", "overflow",
            "rounds", "photon_voxels", "photon_deposits", "photon_pushes", "photon_pops",
            "photon_steps", "electron_voxels", "electron_deposits", "electron_pushes",
            "electron_pops", "electron_steps", "photon_launches", "electron_launches",
        )
    ] + [("photon_kernel_ms", c_double), ("electron_kernel_ms", c_double), ("graph_launches", c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib():
    """Load ``libhipmc.so`` once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (one HIP runtime per process; see the module docstring)
    except ImportError:
        pass
    if not LIB_PATH.is_file():
        raise RuntimeError(
            f"native library {LIB_PATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (or pyradmc_amd.build.build())"
        )
    h = ctypes.CDLL(str(LIB_PATH))
    ctx_p = c_void_p
    sig = {
        "hmc_create": (c_int, [c_int, POINTER(ctx_p)]),
        "hmc_destroy": (None, [ctx_p]),
        "hmc_last_error": (c_char_p, [ctx_p]),
        "hmc_device_info": (c_int, [c_int, c_char_p, c_int, POINTER(c_int), POINTER(c_int64), POINTER(c_int64)]),
        "hmc_set_grid": (c_int, [ctx_p, c_int, c_int, c_int, _f64p, _f64p, _f32p, POINTER(c_uint8)]),
        "hmc_set_scoring": (c_int, [ctx_p, c_int, c_int, c_int, _f64p, _f64p, _f64p, c_int64, c_int]),
        "hmc_set_tables": (c_int, [ctx_p, POINTER(Tables)]),
        "hmc_set_gs": (c_int, [ctx_p, _f32p, c_int, c_int, c_int, c_int, c_int]),
        "hmc_set_source": (c_int, [ctx_p, POINTER(SourceDesc)]),
        "hmc_set_uploaded": (c_int, [ctx_p, c_int64, POINTER(c_int32)] + [_f32p] * 8),
        "hmc_set_population": (c_int, [ctx_p, c_int64, c_int64, c_int64, POINTER(c_int32)] + [_f32p] * 8),
        "hmc_set_ray_weights": (c_int, [ctx_p, POINTER(RayWeightsDesc)]),
        "hmc_presolve_head": (c_int, [ctx_p, c_uint64, c_int64, POINTER(SourceDesc)] + [_f32p] * 8
                              + [POINTER(RayWeightsDesc), POINTER(PresolveDesc), POINTER(c_int64)]),
        "hmc_get_population": (c_int, [ctx_p, c_int64, c_int64, POINTER(c_int32)] + [_f32p] * 8),
        "hmc_ray_weight_factors": (c_int, [ctx_p, c_int64, _f32p, _f32p, _f32p, _f64p]),
        "hmc_set_physics": (c_int, [ctx_p, POINTER(Physics)]),
        "hmc_set_profiling": (c_int, [ctx_p, c_int]),
        "hmc_set_launch": (c_int, [ctx_p, c_int, c_int, c_int, c_int]),
        "hmc_transport": (c_int, [ctx_p, c_uint64, POINTER(WorkItem), c_int, c_int64, c_int, POINTER(Ledger)]),
        "hmc_fold": (c_int, [ctx_p, c_int, _f64p, POINTER(c_int64)]),
        "hmc_folded_quanta": (c_int, [ctx_p, POINTER(c_int64)]),
        "hmc_finalize": (c_int, [ctx_p, c_int, _f64p, _f64p]),
        "hmc_truncate": (c_int, [ctx_p, c_double, POINTER(c_int64), POINTER(c_int64)]),
        "hmc_fetch_csc": (c_int, [ctx_p, POINTER(c_int64), _f64p, _f64p]),
        "hmc_fetch_csc_f32": (c_int, [ctx_p, POINTER(c_int64), _f32p, _f32p]),
        "hmc_host_alloc": (c_int, [ctypes.c_size_t, POINTER(c_void_p)]),
        "hmc_host_free": (None, [c_void_p]),
        "hmc_reset_maps": (c_int, [ctx_p]),
        "hmc_set_quanta": (c_int, [ctx_p, POINTER(c_int64)]),
        "hmc_get_quanta": (c_int, [ctx_p, POINTER(c_int64)]),
        "hmc_build_gs_grid": (c_int, [c_int, c_int, c_int, c_int, c_int, c_double, c_int, _f64p]),
        "hmc_gather_windows": (c_int, [ctx_p, POINTER(c_int32), c_int64, c_int, c_int, c_void_p]),
        "hmc_scatter_windows": (c_int, [ctx_p, POINTER(c_int32), c_int64, c_int, c_int, c_void_p, c_void_p]),
        "hmc_set_column_entries": (c_int, [ctx_p, c_int64, c_int64, POINTER(c_int64), _f64p, _f64p]),
        "hmc_sample_physics": (c_int, [ctx_p, c_uint64, c_int64, POINTER(SamplerArgs), _f32p]),
    }
    for name, (restype, argtypes) in sig.items():
        fn = getattr(h, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = h
    # Registered after torch's import: exit handlers run last-in first-out, so this one
    # runs before torch's and before the HIP runtime's own teardown.
    atexit.register(shutdown)
    return h


# Native handles still open (contexts, device-resident matrices): anything with a
# close() that releases device memory, streams or graphs.
_open_handles: "weakref.WeakSet" = weakref.WeakSet()
_exiting = False


def register_handle(obj) -> None:
    """Have ``shutdown`` close ``obj`` (it must have ``close()``) if it is still open
    at process exit."""
    _open_handles.add(obj)


def shutdown() -> None:
    """Release every native resource in a fixed order before the process exits: open
    contexts and matrices first (their streams, graphs, device and page-locked
    buffers), then the cached page-locked result blocks. Blocks that arrays still view
    are freed as soon as those arrays go, instead of returning to the cache. Runs at
    exit by itself. Called earlier, the library stays usable (new handles and blocks
    are made as before), but released blocks are freed from then on, not cached."""
    global _exiting
    _exiting = True
    for obj in list(_open_handles):
        try:
            obj.close()
        except Exception:
            pass
    _open_handles.clear()
    release_pinned_cache()


def exported_symbols() -> list[str]:
    """Every ``hmc_*`` / ``hdn_*`` / ``hsp_*`` entry point the headers declare."""
    import re

    include = Path(__file__).resolve().parent.parent / "include"
    names = set()
    for header, prefix in (("hipmc.h", "hmc"), ("hipdn.h", "hdn"), ("hipsp.h", "hsp")):
        names |= set(re.findall(rf"\b({prefix}_[a-z0-9_]+)\s*\(", (include / header).read_text()))
    return sorted(names)


def ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(POINTER(ctype))


class NativeError(RuntimeError):
    pass


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc == HMC_OK:
        return
    message = ""
    if ctx is not None:
        raw = lib().hmc_last_error(ctx)
        message = raw.decode() if raw else ""
    text = f"{what}: {message}" if what else message
    if rc == HMC_EINVAL:
        raise ValueError(text)
    if rc == HMC_ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text or f"native status {rc}")


class Context:
    """Owns one ``hmc_ctx`` on one device."""

    def __init__(self, device: int = 0) -> None:
        self._lib = lib()
        handle = c_void_p()
        rc = self._lib.hmc_create(int(device), ctypes.byref(handle))
        if rc != HMC_OK:
            raise RuntimeError(
                f"hmc_create(device={device}) failed with status {rc}: no usable HIP device"
            )
        self.handle = handle
        self.device = device
        register_handle(self)

    def close(self) -> None:
        if self.handle:
            self._lib.hmc_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def call(self, name: str, *args) -> None:
        rc = getattr(self._lib, name)(self.handle, *args)
        check(rc, self.handle, name)


# Released page-locked blocks kept for reuse. Pinning and unpinning memory costs tens of
# microseconds per MiB and the unpin waits for the device, so a loop of runs that each
# return a large result would otherwise pay both per run. At most this many bytes stay
# cached (the oldest blocks go first); ``release_pinned_cache`` returns them to the
# system.
PINNED_CACHE_MAX_BYTES = 256 << 20
_PAGE = 4096
_pinned_lock = threading.RLock()  # reentrant: a block's __del__ may run inside a take
_pinned_cache: list[tuple[int, int]] = []  # (bytes, address)
_pinned_cached_bytes = 0


def _pinned_take(nbytes: int) -> tuple[int, int]:
    """A page-locked block of at least ``nbytes``: a cached one no more than twice that
    size, else a new allocation. Returns (bytes, address)."""
    global _pinned_cached_bytes
    with _pinned_lock:
        best = -1
        for k in reversed(range(len(_pinned_cache))):  # newest first among equal sizes
            size = _pinned_cache[k][0]
            if nbytes <= size <= 2 * nbytes and (best < 0 or size < _pinned_cache[best][0]):
                best = k
        if best >= 0:
            size, address = _pinned_cache.pop(best)
            _pinned_cached_bytes -= size
            return size, address
    size = -(-max(1, nbytes) // _PAGE) * _PAGE
    ptr = c_void_p()
    rc = lib().hmc_host_alloc(size, ctypes.byref(ptr))
    if rc != 0 or not ptr.value:
        release_pinned_cache()  # cached blocks may be what the system is short of
        rc = lib().hmc_host_alloc(size, ctypes.byref(ptr))
        if rc != 0 or not ptr.value:
            raise MemoryError(f"page-locked host allocation of {nbytes} bytes failed")
    return size, ptr.value


def _pinned_give(size: int, address: int) -> None:
    """Cache a released block; when it would pass the byte cap, the oldest cached blocks
    are freed first (the sizes a loop of runs asks for next are its most recent ones).
    A block larger than the cap itself, or released while the process exits, is freed."""
    global _pinned_cached_bytes
    if _lib is None:
        return
    evicted = [address]
    with _pinned_lock:
        if not _exiting and size <= PINNED_CACHE_MAX_BYTES:
            evicted = []
            while _pinned_cached_bytes + size > PINNED_CACHE_MAX_BYTES:
                old_size, old_address = _pinned_cache.pop(0)
                _pinned_cached_bytes -= old_size
                evicted.append(old_address)
            _pinned_cache.append((size, address))
            _pinned_cached_bytes += size
    for a in evicted:
        _lib.hmc_host_free(a)


def release_pinned_cache() -> int:
    """Free the cached page-locked blocks that no array uses any more; returns the bytes
    released."""
    global _pinned_cached_bytes
    with _pinned_lock:
        blocks = list(_pinned_cache)
        _pinned_cache.clear()
        _pinned_cached_bytes = 0
    if _lib is not None:
        for _, address in blocks:
            _lib.hmc_host_free(address)
    return sum(size for size, _ in blocks)


class _PinnedBlock:
    """One page-locked host block, exposed to NumPy; handed back to the cache (or freed)
    when the last array viewing it goes away."""

    def __init__(self, n: int, dtype) -> None:
        dt = np.dtype(dtype)
        self._ptr = None
        self._size, address = _pinned_take(n * dt.itemsize)
        self._ptr = address
        self.__array_interface__ = {"shape": (n,), "typestr": dt.str, "data": (self._ptr, False), "version": 3}

    def __del__(self) -> None:
        if self._ptr:
            address, self._ptr = self._ptr, None
            try:
                _pinned_give(self._size, address)
            except Exception:  # interpreter shutdown: the process releases its memory
                pass


def pinned_empty(n: int, dtype) -> np.ndarray:
    """An uninitialised 1-D array of ``n`` elements in page-locked host memory (the
    destination of a fetched CSC: device-to-host copies into it run at the link's rate).
    It behaves as any NumPy array; when the last view of it goes away its memory goes
    back to a bounded cache of page-locked blocks (``PINNED_CACHE_MAX_BYTES``) for the
    next result."""
    return np.asarray(_PinnedBlock(int(n), dtype))


def device_free_bytes(device: int = 0) -> int:
    """Free device memory now (the cheap part of ``device_info``, for per-run sizing)."""
    free = c_int64()
    rc = lib().hmc_device_info(int(device), None, 0, None, None, ctypes.byref(free))
    check(rc, None, "hmc_device_info")
    return int(free.value)


def device_info(device: int = 0) -> dict:
    h = lib()
    name = ctypes.create_string_buffer(256)
    cu = c_int()
    total = c_int64()
    free = c_int64()
    rc = h.hmc_device_info(int(device), name, 256, ctypes.byref(cu), ctypes.byref(total), ctypes.byref(free))
    check(rc, None, "hmc_device_info")
    return {
        "name": name.value.decode(),
        "cu_count": int(cu.value),
        "total_bytes": int(total.value),
        "free_bytes": int(free.value),
    }


def build_gs_grid_native(ix0: int, ix1: int, iy0: int, iy1: int, n_u: int, bins: float, threads: int = 0) -> np.ndarray:
    out = np.empty((ix1 - ix0 + 1, iy1 - iy0 + 1, n_u), dtype=np.float64)
    rc = lib().hmc_build_gs_grid(int(ix0), int(ix1), int(iy0), int(iy1), int(n_u), float(bins), int(threads), ptr(out, c_double))
    check(rc, None, "hmc_build_gs_grid")
    return out
