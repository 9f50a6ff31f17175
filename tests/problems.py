"""Problem set-ups shared by the CPU and GPU parity tests (inputs of the golden runs)."""

from __future__ import annotations

import numpy as np


def fan_dij_problem(golden):
    """The divergent spectral beamlet Dij of ``ref_dij_fan.npz``: 16x16x24 at 5 mm with a
    0.3 g/cc slab at voxels 6-11 in depth, four 1.0 x 0.8 cm parallelogram apertures
    (v axis (0.3, 1, 0)) at depth 4 cm from a focal point 100 cm upstream, 6 MV-like
    spectrum; 16000 histories per beamlet in 8 batches, correlated, truncation 1e-3."""
    from pyradmc_amd import SpectralBeamletSource, Spectrum, VoxelGrid
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    shape = (16, 16, 24)
    grid = VoxelGrid(shape=shape, spacing=(0.5, 0.5, 0.5), density=np.asarray(golden["density"]),
                     material=np.zeros(shape, dtype=np.int32))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    edges = np.linspace(0.1, 6.0, 61)
    mid = 0.5 * (edges[:-1] + edges[1:])
    spectrum = Spectrum(edges, mid * np.exp(-mid / 1.8), convention="number")
    centers = [(3.5, 3.5, 4.0), (4.5, 3.5, 4.0), (3.5, 4.5, 4.0), (4.5, 4.8, 4.0)]
    src = SpectralBeamletSource(spectrum, (4.0, 4.0, -96.0), centers, 1.0, 0.8,
                                u_axis=(1.0, 0.0, 0.0), v_axis=(0.3, 1.0, 0.0))
    return grid, xs, src


FAN_HISTORIES, FAN_BATCHES, FAN_TRUNCATION = 16000, 8, 1e-3


def csc_dense(indptr, indices, values, n_voxels: int) -> np.ndarray:
    """[columns, voxels] dense array of a CSC Dij."""
    out = np.zeros((len(indptr) - 1, n_voxels))
    for c in range(len(indptr) - 1):
        out[c, indices[indptr[c]:indptr[c + 1]]] = values[indptr[c]:indptr[c + 1]]
    return out
