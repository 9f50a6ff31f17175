"""Dose for many weight vectors at once: D = Dij @ W on the device.

The reference's planning loop calls ``DijResult.dose_for_weights(w)`` once per weight
vector, on the host (pyradmc/scoring/dij.py:251-260). ``DeviceDij`` uploads a Dij once
(CSC, as ``run_dij`` returns it), keeps it on the GPU in row-major forms, and multiplies
it with a whole matrix of weight vectors in one pass (``include/hipsp.h``). The values
stay in the Dij's own type (float64 as the reference stores them, or float32). In
``"f64"`` each output is the reference's float64 sum bit for bit; ``"f32"`` halves
the bytes written.
"""

from __future__ import annotations

import ctypes

import numpy as np

from pyradmc_amd import _native

_c_void = ctypes.c_void_p
_PRECISION = {"f64": (0, np.float64), "f32": (1, np.float32)}


def _lib():
    h = _native.lib()
    if getattr(h, "_hsp_bound", False):
        return h
    i64p = ctypes.POINTER(ctypes.c_int64)
    f64p = ctypes.POINTER(ctypes.c_double)
    ll = ctypes.c_longlong
    sig = {
        "hsp_create": (ctypes.c_int, [ctypes.c_int, ll, ll, i64p, i64p, _c_void, ctypes.c_int,
                                      ctypes.POINTER(_c_void)]),
        "hsp_destroy": (None, [_c_void]),
        "hsp_last_error": (ctypes.c_char_p, [_c_void]),
        "hsp_apply": (ctypes.c_int, [_c_void, f64p, ctypes.c_int, ll, ll, ctypes.c_int, _c_void, f64p]),
        "hsp_apply_device": (ctypes.c_int, [_c_void, _c_void, ctypes.c_int, ll, ll, ctypes.c_int, _c_void, f64p]),
        "hsp_fetch": (ctypes.c_int, [_c_void, ll, ll, _c_void]),
        "hsp_csr": (ctypes.c_int, [_c_void, i64p, ctypes.POINTER(ctypes.c_int32), _c_void]),
        "hsp_block_entries": (ll, [_c_void]),
        "hsp_apply_transpose": (ctypes.c_int, [_c_void, f64p, f64p, f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    h._hsp_bound = True
    return h


def _raise(rc: int, msg: str) -> None:
    if rc == -1:
        raise ValueError(msg)
    if rc == -5:
        raise MemoryError(msg)
    raise RuntimeError(msg)


class DeviceDij:
    """A Dij resident on one HIP device, for repeated products with weight vectors.

    ``dij`` is anything with ``indptr`` / ``indices`` / ``dose`` (CSC, rows = voxels),
    ``n_beamlets`` and ``n_voxels`` or ``grid_shape``: a ``DijResult`` from either
    engine, for instance. float32 values stay float32 on the device; any other type is
    kept as float64.
    """

    def __init__(self, dij, device: int = 0) -> None:
        indptr = np.ascontiguousarray(dij.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(dij.indices, dtype=np.int64)
        dose = np.asarray(dij.dose)
        value_dtype = np.float32 if dose.dtype == np.float32 else np.float64
        values = np.ascontiguousarray(dose, dtype=value_dtype)
        self.value_dtype = np.dtype(value_dtype)
        n_voxels = getattr(dij, "n_voxels", None)
        self.n_rows = int(n_voxels if n_voxels is not None else np.prod(dij.grid_shape))
        self.n_cols = int(dij.n_beamlets)
        self.grid_shape = tuple(getattr(dij, "grid_shape", (self.n_rows,)))
        if indptr.shape != (self.n_cols + 1,):
            raise ValueError(f"indptr has {indptr.size} entries, expected n_beamlets + 1 = {self.n_cols + 1}")
        if indices.shape != (int(indptr[-1]),) or values.shape != indices.shape:
            raise ValueError("indices and dose must hold indptr[-1] entries")
        self.nnz = int(indptr[-1])
        self._h = None
        handle = _c_void()
        rc = _lib().hsp_create(int(device), self.n_rows, self.n_cols,
                               indptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               indices.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               values.ctypes.data, _PRECISION["f32" if value_dtype == np.float32 else "f64"][0],
                               ctypes.byref(handle))
        if rc != 0:
            _raise(rc, f"hsp_create failed with status {rc} (invalid CSC or device error)")
        self._h = handle
        _native.register_handle(self)
        self.last_kernel_ms: float | None = None

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raw = _lib().hsp_last_error(self._h)
            _raise(rc, f"{what}: {raw.decode() if raw else ''} (status {rc})")

    def close(self) -> None:
        if self._h:
            _lib().hsp_destroy(self._h)
            self._h = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    @property
    def block_entries(self) -> int:
        """Entries of the row-block form (the union of columns over each 8 rows)."""
        return int(_lib().hsp_block_entries(self._h))

    def apply(self, weights, rows: tuple[int, int] | None = None, precision: str = "f64",
              fetch: bool = True):
        """D[rows] = Dij[rows] @ W for W [n_beamlets][n_w] (or one vector [n_beamlets]).

        Returns the product [rows][n_w] (or [rows] for one vector) as float64
        (``precision="f64"``) or float32 (``"f32"``) when ``fetch``; otherwise None and
        the product stays on the device. ``last_kernel_ms`` holds the kernel's time.
        """
        code, dtype = _PRECISION[precision]
        w = np.asarray(weights, dtype=np.float64)
        single = w.ndim == 1
        w2 = np.ascontiguousarray(w[:, None] if single else w)
        if w2.ndim != 2 or w2.shape[0] != self.n_cols:
            raise ValueError(f"weights shape {w.shape}: expected ({self.n_cols},) or ({self.n_cols}, n_w)")
        r0, r1 = (0, self.n_rows) if rows is None else (int(rows[0]), int(rows[1]))
        ms = ctypes.c_double()
        self._check(_lib().hsp_apply(self._h, w2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     int(w2.shape[1]), r0, r1, code, None, ctypes.byref(ms)), "hsp_apply")
        self.last_kernel_ms = float(ms.value)
        if not fetch:
            return None
        out = np.empty((r1 - r0, w2.shape[1]), dtype=dtype)
        self._check(_lib().hsp_fetch(self._h, r0, r1, out.ctypes.data), "hsp_fetch")
        return out[:, 0] if single else out

    def apply_device(self, weights, out, rows: tuple[int, int] | None = None) -> float:
        """Product of device-resident torch tensors: ``weights`` [n_beamlets][n_w] and
        ``out`` [rows][n_w], both float64 or both float32, contiguous, on this matrix's
        device. Returns the kernel time in ms."""
        import torch

        if weights.dtype != out.dtype or weights.dtype not in (torch.float64, torch.float32):
            raise ValueError("weights and out must both be float64 or both float32")
        if not (weights.is_contiguous() and out.is_contiguous()) or weights.dim() != 2 or out.dim() != 2:
            raise ValueError("weights and out must be contiguous 2-D tensors")
        r0, r1 = (0, self.n_rows) if rows is None else (int(rows[0]), int(rows[1]))
        n_w = int(weights.shape[1])
        if tuple(weights.shape) != (self.n_cols, n_w) or tuple(out.shape) != (r1 - r0, n_w):
            raise ValueError("weights must be [n_beamlets][n_w] and out [rows][n_w]")
        code = 0 if weights.dtype == torch.float64 else 1
        torch.cuda.current_stream(weights.device).synchronize()  # weights written on torch's stream
        ms = ctypes.c_double()
        self._check(_lib().hsp_apply_device(self._h, weights.data_ptr(), n_w, r0, r1, code, out.data_ptr(),
                                            ctypes.byref(ms)), "hsp_apply_device")
        self.last_kernel_ms = float(ms.value)
        return self.last_kernel_ms

    def apply_transpose(self, vector) -> np.ndarray:
        """``Dij.T @ v`` for one float64 vector over the rows: one value per beamlet
        (the gradient's product in the planning objective). Each column is summed in a
        fixed order on the device: reproducible, and equal to the sequential sum to
        rounding, not bit for bit. ``last_kernel_ms`` holds the kernel's time."""
        v = np.ascontiguousarray(np.asarray(vector, dtype=np.float64).reshape(-1))
        if v.shape != (self.n_rows,):
            raise ValueError(f"vector shape {v.shape} != ({self.n_rows},)")
        out = np.empty(self.n_cols, dtype=np.float64)
        ms = ctypes.c_double()
        f64p = ctypes.POINTER(ctypes.c_double)
        self._check(_lib().hsp_apply_transpose(self._h, v.ctypes.data_as(f64p), out.ctypes.data_as(f64p),
                                               ctypes.byref(ms)), "hsp_apply_transpose")
        self.last_kernel_ms = float(ms.value)
        return out

    def dose_for_weights(self, weights) -> np.ndarray:
        """The reference's ``DijResult.dose_for_weights``: a float64 dose grid."""
        w = np.asarray(weights, dtype=np.float64)
        if w.shape != (self.n_cols,):
            raise ValueError(f"weights shape {w.shape} != ({self.n_cols},)")
        return self.apply(w).reshape(self.grid_shape)

    def csr(self):
        """(row_ptr, col, val) of the device CSR."""
        row_ptr = np.empty(self.n_rows + 1, dtype=np.int64)
        col = np.empty(max(self.nnz, 1), dtype=np.int32)
        val = np.empty(max(self.nnz, 1), dtype=self.value_dtype)
        self._check(_lib().hsp_csr(self._h, row_ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                   col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), val.ctypes.data), "hsp_csr")
        return row_ptr, col[:self.nnz], val[:self.nnz]
