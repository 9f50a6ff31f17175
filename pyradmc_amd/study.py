"""The study's fluence optimiser with its Dij on the GPU (pyradmc/study.py).

The problem (pyradmc/study.py:10-19), deliberately minimal and deterministic:

    minimize_{w >= 0}  mean_{v in PTV} (d_v(w) - Rx)^2
                     + lambda * mean_{v in OAR} max(0, d_v(w) - c Rx)^2,   d(w) = Dij @ w,

solved with L-BFGS-B under nonnegativity. The Dij rows of the PTV and OAR stay on the
device for the whole solve. Every objective evaluation is one device product
``d = D_roi @ w`` (float64, the reference's sum bit for bit), a residual on the host
over the region's rows only, and one device transpose product ``D_roi.T @ a`` for the
gradient (summed in a fixed order: reproducible, equal to the sequential sum to
rounding). Only ``w``, the region's dose and the gradient cross the link.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from functools import cached_property

import numpy as np

__all__ = ["ToyPlanProblem", "dose_at_volume", "objective_and_gradient", "optimize_weights"]


def _csc_parts(dij):
    """(indptr, indices, values, n_rows, n_cols) of a scipy sparse array or matrix, or
    of a DijResult (CSC by column, rows the scoring voxels)."""
    if hasattr(dij, "tocsc"):
        m = dij.tocsc()
        m.sort_indices()
        return (np.asarray(m.indptr, np.int64), np.asarray(m.indices, np.int64), np.asarray(m.data),
                int(m.shape[0]), int(m.shape[1]))
    n_rows = int(getattr(dij, "n_voxels", None) or np.prod(dij.grid_shape))
    return (np.asarray(dij.indptr, np.int64), np.asarray(dij.indices, np.int64), np.asarray(dij.dose), n_rows,
            int(dij.n_beamlets))


class _Rows:
    """A CSC restricted to chosen rows, renumbered, in the shape DeviceDij reads."""

    def __init__(self, indptr, indices, values, keep_rows: np.ndarray, n_cols: int) -> None:
        new_index = np.full(int(keep_rows.size), -1, dtype=np.int64)
        sel = np.flatnonzero(keep_rows)
        new_index[sel] = np.arange(sel.size)
        keep = keep_rows[indices]
        kept_before = np.concatenate(([0], np.cumsum(keep, dtype=np.int64)))
        self.indptr = kept_before[indptr]  # entries kept before each column's start
        self.indices = new_index[indices[keep]]
        self.dose = values[keep]
        self.n_voxels = int(sel.size)
        self.n_beamlets = n_cols


@dataclass(frozen=True)
class ToyPlanProblem:
    """One instance of the study's problem (pyradmc/study.py:40-90): ``dij`` a sparse
    ``(n_voxels, n_beamlets)`` dose matrix (a scipy sparse array or a DijResult),
    ``ptv`` a non-empty flat mask, ``oar`` a flat mask (may be empty), the
    prescription ``Rx``, the OAR limit ``c`` as a fraction of it, and the penalty
    weight ``lambda``. ``device`` is the HIP device holding the region's rows."""

    dij: object
    ptv: np.ndarray
    oar: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=bool))
    prescription: float = 1.0
    oar_max: float = 0.3
    oar_penalty: float = 1.0
    device: int = 0

    def __post_init__(self) -> None:
        n_voxels = self._csc[3]
        ptv = np.asarray(self.ptv, dtype=bool)
        oar = np.asarray(self.oar, dtype=bool)
        if ptv.shape != (n_voxels,) or not bool(ptv.any()):
            raise ValueError(f"ptv must be a nonempty flat mask of {n_voxels} voxels")
        if oar.size and oar.shape != (n_voxels,):
            raise ValueError(f"oar mask shape {oar.shape} != ({n_voxels},)")
        object.__setattr__(self, "ptv", ptv)
        object.__setattr__(self, "oar", oar)

    @cached_property
    def _csc(self):
        return _csc_parts(self.dij)

    @property
    def n_beamlets(self) -> int:
        return int(self._csc[4])

    @cached_property
    def _roi(self):
        """The region's rows on the device, and each region row's role."""
        from pyradmc_amd.spmm import DeviceDij

        indptr, indices, values, n_voxels, n_cols = self._csc
        oar = self.oar if self.oar.size else np.zeros(n_voxels, dtype=bool)
        keep = self.ptv | oar
        rows = _Rows(indptr, indices, values, keep, n_cols)
        return DeviceDij(rows, device=self.device), self.ptv[keep], oar[keep]


def objective_and_gradient(problem: ToyPlanProblem, weights) -> tuple[float, np.ndarray]:
    """Objective and gradient at ``weights`` (pyradmc/study.py:93-114): the PTV term's
    gradient ``2/n D_ptv.T (d - Rx)`` plus the one-sided OAR term's where its residual
    is positive."""
    dev, in_ptv, in_oar = problem._roi
    w = np.asarray(weights, dtype=np.float64)
    d = dev.apply(w)
    r_ptv = d[in_ptv] - problem.prescription
    n_ptv = r_ptv.size
    f = float(r_ptv @ r_ptv) / n_ptv
    a = np.zeros(d.size)
    a[in_ptv] = 2.0 / n_ptv * r_ptv
    if in_oar.any():
        r_oar = np.maximum(0.0, d[in_oar] - problem.oar_max * problem.prescription)
        n_oar = r_oar.size
        f += problem.oar_penalty * float(r_oar @ r_oar) / n_oar
        a[in_oar] += 2.0 * problem.oar_penalty / n_oar * r_oar
    return f, dev.apply_transpose(a)


def optimize_weights(problem: ToyPlanProblem, w0=None, info: dict | None = None) -> np.ndarray:
    """L-BFGS-B under ``w >= 0`` (pyradmc/study.py:117-138): from uniform weights scaled
    to the prescription's mean PTV dose unless ``w0`` is given, with the reference's
    iteration limit and tolerances. Repeat calls return identical weights. ``info``, if
    given, receives the solver's iterations, evaluations and final objective."""
    from scipy.optimize import minimize

    if w0 is None:
        dev, in_ptv, _ = problem._roi
        uniform = np.ones(problem.n_beamlets)
        mean_dose = float(np.mean(dev.apply(uniform)[in_ptv]))
        w0 = uniform * (problem.prescription / mean_dose if mean_dose > 0.0 else 1.0)
    result = minimize(lambda w: objective_and_gradient(problem, w), np.asarray(w0, dtype=np.float64), jac=True,
                      method="L-BFGS-B", bounds=[(0.0, None)] * problem.n_beamlets,
                      options={"maxiter": 500, "ftol": 1.0e-12, "gtol": 1.0e-10})
    if info is not None:
        info.update(iterations=int(result.nit), evaluations=int(result.nfev), objective=float(result.fun))
    return np.asarray(result.x, dtype=np.float64)


def dose_at_volume(dose_in_roi, volume_percent: float) -> float:
    """D(v%): the dose at least ``v`` percent of the region receives, the
    (100 - v)-th percentile (pyradmc/study.py:141-150)."""
    d = np.asarray(dose_in_roi)
    if d.size == 0:
        raise ValueError("empty ROI has no DVH")
    return float(np.percentile(d, 100.0 - volume_percent))
