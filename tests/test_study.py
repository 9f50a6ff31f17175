"""The study's fluence optimiser on the host (CPU): the oracle (oracle/study_ref.py)
pinned by finite differences and closed forms, the row restriction the device problem
uses, and the problem's argument contract."""

from __future__ import annotations

import numpy as np
import pytest
from scipy.sparse import csc_array, random as sparse_random

from oracle import study_ref as S


def _problem(seed=0, n_vox=400, n_beam=12):
    rng = np.random.default_rng(seed)
    dij = csc_array(sparse_random(n_vox, n_beam, density=0.3, random_state=seed, format="csc"))
    ptv = np.zeros(n_vox, bool)
    ptv[100:200] = True
    oar = np.zeros(n_vox, bool)
    oar[180:260] = True  # overlaps the PTV on purpose
    return dij, ptv, oar, rng


def test_gradient_matches_central_differences():
    dij, ptv, oar, rng = _problem()
    w = rng.uniform(0.5, 3.0, dij.shape[1])
    f, g = S.objective_and_gradient(dij, ptv, oar, 1.0, 0.3, 2.0, w)
    h = 1e-6
    fd = np.array([(S.objective_and_gradient(dij, ptv, oar, 1.0, 0.3, 2.0, w + h * e)[0]
                    - S.objective_and_gradient(dij, ptv, oar, 1.0, 0.3, 2.0, w - h * e)[0]) / (2 * h)
                   for e in np.eye(w.size)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    assert f > 0.0


def test_one_beamlet_closed_form_optimum():
    # f(w) = mean((a w - Rx)^2) is minimised at w = Rx sum(a) / sum(a^2)
    a = np.array([0.5, 1.0, 2.0, 0.0, 1.5])
    dij = csc_array(a[:, None])
    ptv = np.array([1, 1, 1, 0, 1], bool)
    w = S.optimize(dij, ptv, np.zeros(0, bool), prescription=2.0)
    ap = a[ptv]
    assert w[0] == pytest.approx(2.0 * ap.sum() / (ap @ ap), rel=1e-8)


def test_oar_term_vanishes_below_its_limit():
    dij, ptv, oar, rng = _problem(1)
    w = np.full(dij.shape[1], 1e-3)  # far below any limit in the OAR
    f_oar, g_oar = S.objective_and_gradient(dij, ptv, oar, 1.0, 0.3, 5.0, w)
    f0, g0 = S.objective_and_gradient(dij, ptv, np.zeros(0, bool), 1.0, 0.3, 5.0, w)
    assert f_oar == f0
    np.testing.assert_array_equal(g_oar, g0)


def test_dose_at_volume_on_a_ramp():
    from pyradmc_amd.study import dose_at_volume

    d = np.arange(101, dtype=float)  # 0..100
    assert dose_at_volume(d, 98.0) == pytest.approx(2.0)
    assert dose_at_volume(d, 2.0) == pytest.approx(98.0)
    assert S.dose_at_volume(d, 50.0) == dose_at_volume(d, 50.0) == pytest.approx(50.0)
    with pytest.raises(ValueError):
        dose_at_volume(np.zeros(0), 50.0)


def test_row_restriction_is_the_scipy_submatrix():
    from pyradmc_amd.study import _Rows

    dij, ptv, oar, _ = _problem(2)
    dij = csc_array(dij)
    dij.sort_indices()
    keep = ptv | oar
    keep[-1] = True  # a row at the end
    r = _Rows(np.asarray(dij.indptr, np.int64), np.asarray(dij.indices, np.int64), dij.data, keep, dij.shape[1])
    sub = dij[keep, :].tocsc()
    sub.sort_indices()
    np.testing.assert_array_equal(r.indptr, sub.indptr)
    np.testing.assert_array_equal(r.indices, sub.indices)
    np.testing.assert_array_equal(r.dose, sub.data)
    assert r.n_voxels == int(keep.sum())
    # trailing empty columns
    z = csc_array((np.array([1.0]), np.array([0]), np.array([0, 1, 1, 1])), shape=(3, 3))
    rz = _Rows(np.asarray(z.indptr, np.int64), np.asarray(z.indices, np.int64), z.data, np.array([1, 0, 1], bool), 3)
    np.testing.assert_array_equal(rz.indptr, [0, 1, 1, 1])


def test_problem_contract():
    from pyradmc_amd.study import ToyPlanProblem

    dij, ptv, oar, _ = _problem(3)
    with pytest.raises(ValueError, match="nonempty"):
        ToyPlanProblem(dij, np.zeros(dij.shape[0], bool))
    with pytest.raises(ValueError, match="oar mask shape"):
        ToyPlanProblem(dij, ptv, np.zeros(7, bool))
    p = ToyPlanProblem(dij, ptv, oar)
    assert p.n_beamlets == dij.shape[1]
