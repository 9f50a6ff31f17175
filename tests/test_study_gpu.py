"""The study's fluence optimiser with its Dij on the GPU, against the scipy oracle
(oracle/study_ref.py).

* Objective equal bit for bit (the device product is the reference's float64 sum),
  gradient to 1e-12 relative (the transpose sums each column in a fixed order).
* On an engine-made Dij (float64 and float32 values), L-BFGS-B reaches the oracle's
  objective, and repeat solves return identical weights.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.sparse import csc_array, random as sparse_random

from oracle import study_ref as S

pytestmark = pytest.mark.gpu


def test_objective_and_gradient_match_the_oracle():
    from pyradmc_amd.study import ToyPlanProblem, objective_and_gradient

    rng = np.random.default_rng(4)
    n_vox, n_beam = 5000, 60
    dij = csc_array(sparse_random(n_vox, n_beam, density=0.2, random_state=4, format="csc"))
    ptv = np.zeros(n_vox, bool)
    ptv[1000:2500] = True
    oar = np.zeros(n_vox, bool)
    oar[2300:3500] = True
    p = ToyPlanProblem(dij, ptv, oar, prescription=2.0, oar_max=0.4, oar_penalty=3.0)
    for _ in range(3):
        w = rng.uniform(0.0, 0.5, n_beam)
        f, g = objective_and_gradient(p, w)
        f_ref, g_ref = S.objective_and_gradient(dij, ptv, oar, 2.0, 0.4, 3.0, w)
        assert f == f_ref
        np.testing.assert_allclose(g, g_ref, rtol=1e-12, atol=1e-15 * np.abs(g_ref).max())


@pytest.fixture(scope="module")
def engine_dij():
    from pyradmc_amd import AnalyticCrossSections, BeamletGridSource, HipEngine, VoxelGrid, _native

    _native.lib()
    grid = VoxelGrid.uniform_water((16, 16, 16), (0.5, 0.5, 0.5))
    eng = HipEngine(grid, AnalyticCrossSections(geometry_densities=grid.max_density_by_material()))
    src = BeamletGridSource(6.0, -0.1, (1.0, 7.0), (1.0, 7.0), 6, 6)
    d64 = eng.run_dij(src, 60_000, 6, seed=5, truncation=1e-4)
    d32 = eng.run_dij(src, 60_000, 6, seed=5, truncation=1e-4, value_dtype=np.float32)
    eng.close()
    ix, iy, iz = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    ptv = ((np.abs(ix - 7.5) < 3) & (np.abs(iy - 7.5) < 3) & (iz >= 4) & (iz < 10)).ravel()
    oar = ((ix >= 10) & (ix < 13) & (np.abs(iy - 7.5) < 4) & (iz >= 3) & (iz < 11)).ravel()
    return d64, d32, ptv, oar


@pytest.mark.parametrize("which", ["f64", "f32"])
def test_engine_dij_optimisation_matches_the_oracle(engine_dij, which):
    from pyradmc_amd.study import ToyPlanProblem, objective_and_gradient, optimize_weights

    d64, d32, ptv, oar = engine_dij
    d = d64 if which == "f64" else d32
    dij = d.dose_csc()
    rx = float(np.median(csc_array(dij)[ptv, :] @ np.ones(dij.shape[1])))
    p = ToyPlanProblem(d, ptv, oar, prescription=rx, oar_max=0.3, oar_penalty=10.0)
    w = optimize_weights(p)
    w_ref = S.optimize(dij, ptv, oar, prescription=rx, oar_max=0.3, oar_penalty=10.0)
    f, _ = objective_and_gradient(p, w)
    f_ref, _ = S.objective_and_gradient(dij, ptv, oar, rx, 0.3, 10.0, w_ref)
    print(f"  {which}: objective {f:.8g} (oracle {f_ref:.8g})", flush=True)
    # L-BFGS-B stops where its relative decrease falls below ftol; gradients that differ
    # in rounding stop it at slightly different points of a flat optimum
    assert f == pytest.approx(f_ref, rel=1e-4)
    f0, _ = S.objective_and_gradient(dij, ptv, oar, rx, 0.3, 10.0, np.ones(dij.shape[1]) * 1e-9)
    assert f < 0.2 * f0
    np.testing.assert_array_equal(w, optimize_weights(p))  # deterministic
    # the plans: PTV dose and its DVH endpoints agree
    dose, dose_ref = csc_array(dij)[ptv, :] @ w, csc_array(dij)[ptv, :] @ w_ref
    assert dose.mean() == pytest.approx(dose_ref.mean(), rel=1e-3)
    for v in (98.0, 50.0, 2.0):
        assert S.dose_at_volume(dose, v) == pytest.approx(S.dose_at_volume(dose_ref, v), rel=5e-3)
