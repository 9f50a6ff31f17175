"""Builders for the benchmark workloads (phantoms, sources, run parameters).

* ``small_pencil`` - 16^3 uniform water, 6 MeV photon pencil along +z, 80,000
  histories in 10 batches, seed 20260726. ``units="cm"`` is the pyRadMC quick-start
  reading (1 cm voxels, pencil at (8, 8, -1) cm); ``units="mm"`` the 1 mm reading
  (0.1 cm voxels, pencil at (0.8, 0.8, -0.1) cm).
* ``slab_phantom`` - 64x64x96 voxels at 2 mm: water, lung slab (0.26 g/cc) at depth
  3-7 cm, bone slab (1.85 g/cc) at 9-11 cm; 10x10 cm parallel field of 6 MV-like
  spectral photons; 1e7 histories in 20 batches.
* ``ct_like_phantom`` - 128x128x96 voxels at 2.5 mm: water body ellipse with seeded
  lung / soft-tissue / bone ellipsoids; 100 spectral beamlets of 5x5 mm on a
  10x10 cm grid from a point source at 100 cm SAD; lognormal per-beamlet counts
  (median 2e5, sigma 0.8).
* ``full_plan`` - 256x256x128 voxels at 2 mm, the same kind of CT-like phantom; five
  coplanar beams at gantry 0/72/144/216/288 degrees about the y axis through the
  isocentre (the grid centre), each 15x15 spectral beamlets of 5x5 mm from 100 cm SAD
  (1125 columns). Each beam is its own source (the gantry rotation is applied to the
  focal point, beamlet centres and aperture axes on the host); ``hstack_dij`` joins
  the five beams' Dijs column-wise, in gantry order.

All heterogeneity is density-only water (the analytic source is water-only).
"""

from __future__ import annotations

import math

import numpy as np

from pyradmc_amd.data.analytic import AnalyticCrossSections
from pyradmc_amd.scoring import DijResult
from pyradmc_amd.geometry import (
    PencilBeamSource,
    SpectralBeamletSource,
    SpectralParallelBeamSource,
    VoxelGrid,
    bremsstrahlung_6mv_like,
)


def small_pencil(units: str = "cm"):
    if units == "cm":
        spacing, position = 1.0, (8.0, 8.0, -1.0)
    elif units == "mm":
        spacing, position = 0.1, (0.8, 0.8, -0.1)
    else:
        raise ValueError(f"units must be 'cm' or 'mm', got {units!r}")
    grid = VoxelGrid.uniform_water((16, 16, 16), (spacing, spacing, spacing))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    source = PencilBeamSource(6.0, position, (0.0, 0.0, 1.0))
    return grid, xs, source, dict(n_histories=80_000, n_batches=10, seed=20260726)


def slab_grid() -> VoxelGrid:
    shape = (64, 64, 96)
    density = np.ones(shape, dtype=np.float64)
    density[:, :, 15:35] = 0.26  # lung, z in [3, 7) cm
    density[:, :, 45:55] = 1.85  # bone, z in [9, 11) cm
    return VoxelGrid(shape=shape, spacing=(0.2, 0.2, 0.2), density=density,
                     material=np.zeros(shape, dtype=np.int32))


def slab_phantom(n_histories: int = 10_000_000, n_batches: int = 20, seed: int = 20260711):
    grid = slab_grid()
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    source = SpectralParallelBeamSource(bremsstrahlung_6mv_like(), -0.1, (1.4, 11.4), (1.4, 11.4))
    return grid, xs, source, dict(n_histories=n_histories, n_batches=n_batches, seed=seed)


def ct_like_grid(seed: int = 20260711, shape=(128, 128, 96), spacing: float = 0.25) -> VoxelGrid:
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) * spacing
    y = (np.arange(ny) + 0.5) * spacing
    z = (np.arange(nz) + 0.5) * spacing
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    cx, cy = nx * spacing / 2, ny * spacing / 2
    # Body outline: elliptic cylinder along z; air-like outside is not supported
    # (vacuum-free grids), so the outside is low-density water at 0.05 g/cc.
    body = ((X - cx) / (0.45 * nx * spacing)) ** 2 + ((Y - cy) / (0.35 * ny * spacing)) ** 2 <= 1.0
    density = np.where(body, 1.0, 0.05)
    kinds = [("lung", 0.2, 0.35, 2), ("soft", 1.02, 1.06, 4), ("bone", 1.4, 1.9, 3)]
    for _, lo, hi, count in kinds:
        for _ in range(count):
            c = np.array([rng.uniform(0.3, 0.7) * nx * spacing, rng.uniform(0.35, 0.65) * ny * spacing,
                          rng.uniform(0.2, 0.8) * nz * spacing])
            r = rng.uniform(1.0, 4.0, size=3)
            inside = ((X - c[0]) / r[0]) ** 2 + ((Y - c[1]) / r[1]) ** 2 + ((Z - c[2]) / r[2]) ** 2 <= 1.0
            density = np.where(inside & body, rng.uniform(lo, hi), density)
    return VoxelGrid(shape=shape, spacing=(spacing,) * 3, density=density,
                     material=np.zeros(shape, dtype=np.int32))


def ct_like_beamlets(n_side: int = 10, pitch: float = 1.0, width: float = 0.5,
                     iso=(16.0, 16.0, 12.0), sad: float = 100.0):
    centers = []
    half = (n_side - 1) / 2.0
    for i in range(n_side):
        for j in range(n_side):
            centers.append((iso[0] + (i - half) * pitch, iso[1] + (j - half) * pitch, iso[2]))
    focal = (iso[0], iso[1], iso[2] - sad)
    return SpectralBeamletSource(bremsstrahlung_6mv_like(), focal, centers, width, width)


def lognormal_counts(n_beamlets: int, n_batches: int, median: float = 2e5, sigma: float = 0.8,
                     seed: int = 20260711) -> np.ndarray:
    rng = np.random.default_rng(seed)
    raw = np.round(median * np.exp(sigma * rng.standard_normal(n_beamlets)))
    counts = np.maximum(n_batches, (np.ceil(raw / n_batches) * n_batches)).astype(np.int64)
    return counts


def ct_like_dij(n_batches: int = 10, median: float = 2e5, seed: int = 20260711):
    grid = ct_like_grid(seed)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    source = ct_like_beamlets()
    counts = lognormal_counts(source.n_beamlets, n_batches, median=median, seed=seed)
    return grid, xs, source, dict(n_histories_per_beamlet=counts, n_batches=n_batches, seed=seed,
                                  truncation=1e-4)


def dose_weighted_rms(dose: np.ndarray, reference: np.ndarray) -> float:
    """``sqrt(sum R (D-R)^2 / sum R) / max R`` over voxels with R > 0.1 max R."""
    r_max = float(reference.max())
    mask = reference > 0.1 * r_max
    r = reference[mask]
    d = dose[mask]
    return math.sqrt(float(np.sum(r * (d - r) ** 2) / np.sum(r))) / r_max


FULL_PLAN_GANTRY_DEG = (0.0, 72.0, 144.0, 216.0, 288.0)


def full_plan_grid(seed: int = 20260711) -> VoxelGrid:
    return ct_like_grid(seed, shape=(256, 256, 128), spacing=0.2)


def full_plan_beams(grid: VoxelGrid, n_side: int = 15, pitch: float = 0.5, sad: float = 100.0,
                    gantry_deg=FULL_PLAN_GANTRY_DEG) -> list:
    """One SpectralBeamletSource per gantry angle. Gantry 0 points along +z; a gantry
    angle t turns the beam about the y axis through the isocentre, to direction
    (sin t, 0, cos t), with the aperture's u axis (cos t, 0, -sin t) and v axis +y."""
    iso = np.array([grid.shape[i] * grid.spacing[i] / 2.0 for i in range(3)])
    half = (n_side - 1) / 2.0
    beams = []
    for deg in gantry_deg:
        t = math.radians(deg)
        d = np.array([math.sin(t), 0.0, math.cos(t)])
        u = np.array([math.cos(t), 0.0, -math.sin(t)])
        v = np.array([0.0, 1.0, 0.0])
        centers = [tuple(iso + (i - half) * pitch * u + (j - half) * pitch * v)
                   for i in range(n_side) for j in range(n_side)]
        beams.append(SpectralBeamletSource(bremsstrahlung_6mv_like(), tuple(iso - sad * d), centers, pitch, pitch,
                                           u_axis=tuple(u), v_axis=tuple(v)))
    return beams


def full_plan(seed: int = 20260711):
    grid = full_plan_grid(seed)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    return grid, xs, full_plan_beams(grid), dict(n_batches=10, seed=seed, truncation=1e-4)


def hstack_dij(parts: list) -> DijResult:
    """Column-wise join of Dijs over one grid (beam after beam); energies add."""
    if not parts:
        raise ValueError("nothing to join")
    first = parts[0]
    if any(tuple(p.grid_shape) != tuple(first.grid_shape) or p.n_batches != first.n_batches for p in parts):
        raise ValueError("Dijs to join must share the grid and the batch count")
    indptr = [np.zeros(1, dtype=np.int64)]
    offset = 0
    for p in parts:
        indptr.append(np.asarray(p.indptr[1:], dtype=np.int64) + offset)
        offset += int(p.indptr[-1])
    counts = [np.broadcast_to(np.asarray(p.n_histories_per_beamlet, dtype=np.int64), (p.n_beamlets,))
              for p in parts]
    counts = np.concatenate(counts)
    return DijResult(
        grid_shape=tuple(first.grid_shape),
        n_beamlets=sum(p.n_beamlets for p in parts),
        n_histories_per_beamlet=int(counts[0]) if np.all(counts == counts[0]) else counts,
        n_batches=first.n_batches,
        truncation=first.truncation,
        indptr=np.concatenate(indptr),
        indices=np.concatenate([p.indices for p in parts]),
        dose=np.concatenate([p.dose for p in parts]),
        sigma=np.concatenate([p.sigma for p in parts]),
        energy_emitted=sum(p.energy_emitted for p in parts),
        energy_deposited=sum(p.energy_deposited for p in parts),
        energy_escaped=sum(p.energy_escaped for p in parts),
        energy_unscored=sum(p.energy_unscored for p in parts),
        correlated=first.correlated,
        scoring_mode=first.scoring_mode,
        provenance=first.provenance,
    )
