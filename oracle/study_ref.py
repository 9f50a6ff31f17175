"""ORACLE (test infrastructure only): the study's fluence optimiser in scipy float64.

Only tests/ use this module, as the checker of ``pyradmc_amd.study``. It restates
pyradmc/study.py: the region-restricted Dij (:82-90), ``objective_and_gradient``
(:93-114: the PTV mean-square term and the one-sided OAR term with their analytic
gradients), ``optimize_weights`` (:117-138: L-BFGS-B from uniform weights scaled to the
prescription, maxiter 500, ftol 1e-12, gtol 1e-10) and ``dose_at_volume`` (:141-150).

Pinned in tests/test_study.py: the gradient against central finite differences, the
one-beamlet problem's closed-form optimum, the OAR term vanishing below its limit, and
D(v%) on a known ramp.
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import minimize
from scipy.sparse import csc_array


def objective_and_gradient(dij, ptv, oar, prescription, oar_max, oar_penalty, w):
    d_ptv = csc_array(dij)[ptv, :]
    r = d_ptv @ w - prescription
    n = r.size
    f = float(r @ r) / n
    g = 2.0 / n * (d_ptv.T @ r)
    if oar.size and oar.any():
        d_oar = csc_array(dij)[oar, :]
        ro = np.maximum(0.0, d_oar @ w - oar_max * prescription)
        f += oar_penalty * float(ro @ ro) / ro.size
        g = g + 2.0 * oar_penalty / ro.size * (d_oar.T @ ro)
    return f, np.asarray(g, dtype=np.float64)


def optimize(dij, ptv, oar, prescription=1.0, oar_max=0.3, oar_penalty=1.0, w0=None):
    n = dij.shape[1]
    if w0 is None:
        mean = float(np.mean(csc_array(dij)[ptv, :] @ np.ones(n)))
        w0 = np.ones(n) * (prescription / mean if mean > 0.0 else 1.0)
    res = minimize(lambda w: objective_and_gradient(dij, ptv, oar, prescription, oar_max, oar_penalty, w), w0,
                   jac=True, method="L-BFGS-B", bounds=[(0.0, None)] * n,
                   options={"maxiter": 500, "ftol": 1.0e-12, "gtol": 1.0e-10})
    return np.asarray(res.x, dtype=np.float64)


def dose_at_volume(dose, volume_percent):
    return float(np.percentile(np.asarray(dose), 100.0 - volume_percent))
