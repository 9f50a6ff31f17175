"""CT adapter: HU calibration known answers, grid assembly and the ITK frame conversion
(pyradmc/adapters/ct.py; cases after tests/unit/test_ct_calibration.py), and a CT-derived
grid transported on the GPU against the oracle."""

from __future__ import annotations

import numpy as np
import pytest

from pyradmc_amd.adapters import ct
from pyradmc_amd.data.materials import ADIPOSE, AIR, CORTICAL_BONE, LUNG, N_REGISTRY, WATER

CAL = ct.DEFAULT_CALIBRATION


def test_density_ramp_known_answers():
    assert CAL.density(-1000.0) == pytest.approx(0.00121)
    assert CAL.density(0.0) == 1.0
    assert CAL.density(500.0) == pytest.approx(1.3)            # halfway between 1.0 and 1.6
    assert CAL.density(1400.0) == pytest.approx(1.84)          # bone segment, about 1.85
    assert CAL.density(-5000.0) == CAL.density(-1000.0)        # flat beyond the ends
    assert CAL.density(9000.0) == CAL.density(3000.0) == pytest.approx(2.8)
    rho = CAL.density(np.linspace(-2000.0, 5000.0, 701))
    assert np.all(np.diff(rho) >= 0.0) and np.all(rho > 0.0)


@pytest.mark.parametrize("hu,expected", [(-1000.0, AIR), (-950.0, LUNG), (-500.0, LUNG), (-120.0, ADIPOSE),
                                         (-60.0, ADIPOSE), (-20.0, WATER), (0.0, WATER), (124.9, WATER),
                                         (125.0, CORTICAL_BONE), (800.0, CORTICAL_BONE)])
def test_material_bins_closed_on_the_left(hu, expected):
    assert CAL.material(hu) == expected


def test_vectorised_and_registry_bound():
    mats = CAL.material(np.linspace(-2000.0, 4000.0, 997))
    assert mats.dtype.kind == "i" and np.all((mats >= 0) & (mats < N_REGISTRY))
    assert np.all(ct.WATER_CALIBRATION.material(np.linspace(-2000.0, 4000.0, 97)) == WATER)
    np.testing.assert_array_equal(ct.WATER_CALIBRATION.density([-700.0, 40.0]), CAL.density([-700.0, 40.0]))


@pytest.mark.parametrize("kw,match", [
    (dict(density_hu=(0.0, -1000.0), density_values=(1.0, 0.0012)), "ascending"),
    (dict(density_values=(0.0, 1.0)), "positive"),
    (dict(material_thresholds=(-950.0, 100.0), material_indices=(AIR, WATER)), "bins"),
    (dict(material_thresholds=(100.0, -950.0), material_indices=(AIR, WATER, CORTICAL_BONE)), "ascending"),
    (dict(material_indices=(WATER, N_REGISTRY)), "registry"),
])
def test_calibration_validation(kw, match):
    args = dict(density_hu=(-1000.0, 0.0), density_values=(0.0012, 1.0), material_thresholds=(0.0,),
                material_indices=(WATER, CORTICAL_BONE))
    args.update(kw)
    with pytest.raises(ValueError, match=match):
        ct.HounsfieldCalibration(**args)


def test_grid_from_hu():
    hu = np.array([[[-1000.0, 0.0], [-500.0, 800.0]]])
    g = ct.grid_from_hu(hu, spacing=(0.3, 0.3, 0.3), origin=(-1.0, -2.0, -3.0))
    assert g.shape == (1, 2, 2) and g.spacing == (0.3, 0.3, 0.3) and g.origin == (-1.0, -2.0, -3.0)
    assert [int(g.material[0, i, j]) for i in range(2) for j in range(2)] == [AIR, WATER, LUNG, CORTICAL_BONE]
    assert g.density[0, 0, 1] == 1.0


def test_itk_frame_conversion():
    """(z, y, x) mm image, origin at the first voxel's centre, with the y axis reversed
    (direction -1): values land at ascending coordinates, origin at the lower corner."""
    rng = np.random.default_rng(1)
    nz, ny, nx = 3, 4, 5
    zyx = rng.uniform(-900.0, 900.0, (nz, ny, nx))
    spacing, origin = (2.0, 2.5, 3.0), (10.0, 40.0, -6.0)
    g = ct.grid_from_array(zyx, spacing, origin, np.diag([1.0, -1.0, 1.0]).ravel())
    assert g.shape == (nx, ny, nz)
    assert g.spacing == pytest.approx((0.2, 0.25, 0.3))
    # x: centre 10 mm -> corner 9 mm; y: last voxel's centre 40 - 3*2.5 = 32.5 mm -> corner 31.25 mm
    assert g.origin == pytest.approx((0.9, 3.125, -0.75))
    for x, y, z in ((0, 0, 0), (4, 3, 2), (2, 1, 1)):
        assert g.density[x, y, z] == CAL.density(zyx[z, ny - 1 - y, x])
    with pytest.raises(ValueError, match="oblique"):
        ct.grid_from_array(zyx, spacing, origin, [1, 0.1, 0, 0, 1, 0, 0, 0, 1])


@pytest.mark.gpu
def test_ct_grid_transport_matches_oracle(oracle_lib):
    """A synthetic HU volume (air shell, soft tissue, a lung block, a bone rod) through the
    water calibration: density-only heterogeneity, transported against the oracle."""
    from tests.stats import batched_chi2_pvalue

    from pyradmc_amd import HipEngine, PencilBeamSource
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    hu = np.full((20, 20, 24), 40.0)
    hu[:, :, :2] = -1000.0
    hu[4:16, 4:16, 6:12] = -750.0
    hu[8:12, 8:12, 14:20] = 1200.0
    grid = ct.grid_from_hu(hu, (0.5, 0.5, 0.5), calibration=ct.WATER_CALIBRATION)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    src = PencilBeamSource(6.0, (5.0, 5.0, -0.1), (0.0, 0.0, 1.0))
    eng = HipEngine(grid, xs)
    r = eng.run(src, 800_000, 8, seed=21)
    eng.close()
    ora = oracle_lib.for_grid(grid, src).run(200_000, 8, seed=22, threads=16)
    mask = ora["dose"] > 0.1 * ora["dose"].max()
    p, *_ = batched_chi2_pvalue(r.dose, r.dose_sigma, 8, ora["dose"], ora["dose_sigma"], 8, mask)
    assert p > 0.01
