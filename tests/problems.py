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


# ------------------------------------------------------------------ two materials
#
# The reference's synthetic two-material instrument (its tests/unit/
# test_tabulated_source.py:29-78): a water row with known analytic forms on 32-node
# geometric grids, and an 'air' row at half the water values (twice the CSDA range).
# A test instrument for the material-row plumbing, not physics.

TWO_MATERIAL_DELTA_CUT = 0.2
TWO_MATERIAL_SHAPE = (8, 8, 16)
TWO_MATERIAL_SPACING = (2.0, 2.0, 1.0)
TWO_MATERIAL_SLAB = slice(5, 9)  # z voxels of material 1


def two_material_data(n_p: int = 32, n_e: int = 32) -> dict:
    """Fields of the compiled container (photon process -> [2][n_p] rows)."""
    pe = np.geomspace(0.01, 10.0, n_p)
    ee = np.geomspace(0.01, 10.0, n_e)
    water_mu = {0: 0.15 * np.ones_like(pe), 1: 1.0e-3 / pe**3,
                2: np.maximum(pe - 1.022, 0.0) * 1.0e-3, 3: 0.01 / pe}
    water_e = {"restricted_stopping": 2.0 / ee, "radiative_stopping": 0.01 * ee, "moller": 0.1 / ee,
               "csda_range": 0.5 * ee, "scattering_power": 1.0 / ee}
    scale = {"csda_range": 2.0}
    data = {
        "photon_energies": pe,
        "mu_over_rho": {p: np.stack([t, 0.5 * t]) for p, t in water_mu.items()},
        "electron_energies": ee,
        "delta_cut": TWO_MATERIAL_DELTA_CUT,
        "materials": ("water", "air"),
        "provenance": "synthetic two-material",
    }
    for name, t in water_e.items():
        data[name] = np.stack([t, scale.get(name, 0.5) * t])
    return data


def two_material_slab():
    """(grid, cross sections) of the slab: 8x8x16 voxels of 2x2x1 cm, unit density
    everywhere, material 1 at z voxels 5-8 (the reference's
    tests/integration/test_multimaterial_transport.py:27-48)."""
    from pyradmc_amd import VoxelGrid
    from pyradmc_amd.data.tabulated import TabulatedCrossSections, TabulatedData

    material = np.zeros(TWO_MATERIAL_SHAPE, dtype=np.int32)
    material[:, :, TWO_MATERIAL_SLAB] = 1
    grid = VoxelGrid(shape=TWO_MATERIAL_SHAPE, spacing=TWO_MATERIAL_SPACING,
                     density=np.ones(TWO_MATERIAL_SHAPE), material=material)
    xs = TabulatedCrossSections(TabulatedData(**two_material_data()),
                                geometry_densities=grid.max_density_by_material())
    return grid, xs


def two_material_beam(energy: float = 2.0):
    """The slab's 16 x 16 cm parallel beam along +z from z = -1 cm."""
    from pyradmc_amd import ParallelBeamSource

    return ParallelBeamSource(energy=energy, z=-1.0, x_range=(0.0, 16.0), y_range=(0.0, 16.0))
