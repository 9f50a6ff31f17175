"""Scoring grid and the result types of ``run`` and ``run_dij``.

Restatements of the reference's boundary types with the same fields and exports:
``ScoringGrid`` (``pyradmc/scoring/grid.py:66-219``), ``RunProvenance`` and
``TransportResult`` (``pyradmc/backends/results.py:19-116``), ``DijResult``
(``pyradmc/scoring/dij.py:194-318``). Doses are MeV/g per emitted history.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

from pyradmc_amd.constants import GY_PER_MEV_PER_G


def _axis_overlap(s_origin, s_spacing, s_n, t_origin, t_spacing, t_n) -> np.ndarray:
    s_lo = s_origin + s_spacing * np.arange(s_n, dtype=np.float64)
    s_hi = s_lo + s_spacing
    t_lo = t_origin + t_spacing * np.arange(t_n, dtype=np.float64)
    t_hi = t_lo + t_spacing
    overlap = np.minimum(s_hi[:, None], t_hi[None, :]) - np.maximum(s_lo[:, None], t_lo[None, :])
    return np.maximum(overlap, 0.0)


@dataclass(frozen=True)
class ScoringGrid:
    """Rectilinear dose grid with per-voxel mass rebinned from the transport grid."""

    shape: tuple[int, int, int]
    spacing: tuple[float, float, float]
    origin: tuple[float, float, float]
    voxel_mass: np.ndarray

    def __post_init__(self) -> None:
        if any(n < 1 for n in self.shape):
            raise ValueError(f"empty scoring grid shape {self.shape}")
        if any(s <= 0.0 for s in self.spacing):
            raise ValueError(f"non-positive scoring spacing {self.spacing}")
        if self.voxel_mass.shape != tuple(self.shape):
            raise ValueError(
                f"voxel_mass shape {self.voxel_mass.shape} != scoring grid shape {self.shape}"
            )
        if np.any(self.voxel_mass < 0.0):
            raise ValueError("negative scoring voxel mass")

    @classmethod
    def for_grid(cls, grid) -> "ScoringGrid":
        """Score on the transport grid: mass = density * voxel volume (same arithmetic)."""
        return cls(
            shape=tuple(grid.shape),
            spacing=tuple(grid.spacing),
            origin=tuple(grid.origin),
            voxel_mass=np.asarray(grid.density, dtype=np.float64) * grid.voxel_volume,
        )

    @classmethod
    def rebin(cls, grid, shape, spacing, origin) -> "ScoringGrid":
        """Arbitrary geometry; mass by exact separable voxel overlap."""
        if any(n < 1 for n in shape):
            raise ValueError(f"empty scoring grid shape {shape}")
        if any(s <= 0.0 for s in spacing):
            raise ValueError(f"non-positive scoring spacing {spacing}")
        wx = _axis_overlap(origin[0], spacing[0], shape[0], grid.origin[0], grid.spacing[0], grid.shape[0])
        wy = _axis_overlap(origin[1], spacing[1], shape[1], grid.origin[1], grid.spacing[1], grid.shape[1])
        wz = _axis_overlap(origin[2], spacing[2], shape[2], grid.origin[2], grid.spacing[2], grid.shape[2])
        partial_x = np.tensordot(wx, np.asarray(grid.density, dtype=np.float64), axes=(1, 0))
        partial_xy = np.einsum("bj,ajk->abk", wy, partial_x)
        mass = np.einsum("ck,abk->abc", wz, partial_xy)
        return cls(shape=tuple(shape), spacing=tuple(spacing), origin=tuple(origin), voxel_mass=mass)

    @property
    def n_voxels(self) -> int:
        return int(self.shape[0] * self.shape[1] * self.shape[2])


@dataclass(frozen=True)
class RunProvenance:
    version: str
    backend: str
    device: str
    seed: int
    pcut_mev: float
    ecut_mev: float
    msc_model: str
    step_energy_fraction: float
    cross_sections: str
    deposit_resolution_cm: float | None = None

    def summary(self) -> str:
        return (
            f"pyradmc_amd {self.version} {self.backend}/{self.device} seed={self.seed} "
            f"pcut={self.pcut_mev} ecut={self.ecut_mev} msc={self.msc_model} "
            f"step={self.step_energy_fraction} deposit_res={self.deposit_resolution_cm} "
            f"xs=[{self.cross_sections}]"
        )


@dataclass(frozen=True)
class TransportResult:
    dose: np.ndarray
    dose_sigma: np.ndarray
    energy_emitted: float
    energy_deposited: float
    energy_escaped: float
    n_histories: int
    n_batches: int
    energy_unscored: float = 0.0
    scoring_mode: str = "dose_to_medium"
    provenance: RunProvenance | None = None
    counters: dict | None = None
    """Device event counts of the run (voxel fetches, deposits, pushes, steps), the
    denominator of the algorithmic-bytes roofline; not part of the reference type."""


@dataclass(frozen=True)
class DijResult:
    """Sparse beamlet-resolved dose matrix in CSC layout (rows = flat scoring voxels)."""

    grid_shape: tuple[int, int, int]
    n_beamlets: int
    n_histories_per_beamlet: object  # int, or the per-beamlet array of the uneven extension
    n_batches: int
    truncation: float
    indptr: np.ndarray
    indices: np.ndarray
    dose: np.ndarray
    sigma: np.ndarray
    energy_emitted: float
    energy_deposited: float
    energy_escaped: float
    energy_unscored: float = 0.0
    correlated: bool = False
    scoring_mode: str = "dose_to_medium"
    provenance: RunProvenance | None = None
    counters: dict | None = None

    @property
    def n_voxels(self) -> int:
        return int(np.prod(self.grid_shape))

    def column_dense(self, beamlet: int) -> np.ndarray:
        return self._scatter(self.dose, beamlet)

    def sigma_dense(self, beamlet: int) -> np.ndarray:
        return self._scatter(self.sigma, beamlet)

    def _scatter(self, values: np.ndarray, beamlet: int) -> np.ndarray:
        if not 0 <= beamlet < self.n_beamlets:
            raise IndexError(f"beamlet {beamlet} outside 0..{self.n_beamlets - 1}")
        lo, hi = int(self.indptr[beamlet]), int(self.indptr[beamlet + 1])
        out = np.zeros(self.n_voxels, dtype=np.float64)
        out[self.indices[lo:hi]] = values[lo:hi]
        return out.reshape(self.grid_shape)

    def dose_for_weights(self, weights: np.ndarray) -> np.ndarray:
        """Host form of the container's product, as the reference's
        (pyradmc/scoring/dij.py:251-260). For many products use ``to_device()``."""
        w = np.asarray(weights, dtype=np.float64)
        if w.shape != (self.n_beamlets,):
            raise ValueError(f"weights shape {w.shape} != ({self.n_beamlets},)")
        out = np.zeros(self.n_voxels, dtype=np.float64)
        for j in range(self.n_beamlets):
            lo, hi = int(self.indptr[j]), int(self.indptr[j + 1])
            np.add.at(out, self.indices[lo:hi], w[j] * self.dose[lo:hi])
        return out.reshape(self.grid_shape)

    def to_device(self, device: int = 0):
        """This matrix on a HIP device (``pyradmc_amd.spmm.DeviceDij``), for products
        with many weight vectors."""
        from pyradmc_amd.spmm import DeviceDij

        return DeviceDij(self, device)

    @staticmethod
    def _unit_factor(unit: str) -> float:
        if unit == "mev_per_g":
            return 1.0
        if unit == "gy":
            return GY_PER_MEV_PER_G
        raise ValueError(f"unknown unit {unit!r}; expected 'mev_per_g' or 'gy'")

    def _csc(self, values: np.ndarray, factor: float):
        from scipy.sparse import csc_array

        data = values if factor == 1.0 else values * factor
        return csc_array((data, self.indices, self.indptr), shape=(self.n_voxels, self.n_beamlets))

    def dose_csc(self, unit: str = "mev_per_g"):
        return self._csc(self.dose, self._unit_factor(unit))

    def sigma_csc(self, unit: str = "mev_per_g"):
        return self._csc(self.sigma, self._unit_factor(unit))

    def variance_csc(self, unit: str = "mev_per_g"):
        if self.correlated:
            warnings.warn(
                "correlated sampling: the beamlet columns share random streams, so each "
                "variance describes its own entry only; sums of variances over several "
                "columns (a plan dose's variance, for one) leave out their covariance",
                UserWarning,
                stacklevel=2,
            )
        return self._csc(self.sigma**2, self._unit_factor(unit) ** 2)
