"""ORACLE (test infrastructure only): NumPy restatement of the GS node construction.

Restates ``pyradmc/data/goudsmit_saunderson.py``: screened-Rutherford Legendre
moments (:402-453), the delta-subtracted series CDF (:456-516), Poisson composition
with self-convolution doublings (:519-604), the series/composition split at
Lambda = 300 and the 512 unit-mean bin averages (:652-759), and the window rule of
``build_gs_grid`` (:1108-1219). The composition branch draws from
``numpy.random.default_rng(20260720)`` through the same Generator calls in the same
order (poisson, random, uniform, permutation), which the reference's MC nodes depend
on (numpy 2.2 here).

Pinned by: the reference's own node golden vectors (``tests/golden/
gs_nodes_water.npz``; series and composition nodes at three screening values), the
exact unit mean of every row, and the closed-form first moment
``<1 - mu> = 1 - exp(-Lambda G_1)`` of the series CDF (``tests/test_oracle_pins.py``).
"""

from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

RAW_SAMPLES = 1 << 19
MU_BINS = 4096
QUADRATURE = 1 << 16
L_MAX = 1024
SERIES_MIN_LAMBDA = 300.0
SERIES_CUTOFF = 1.0e-12
SEED = 20260720


def g1(eta: float) -> float:
    return float(2.0 * eta * (1.0 + eta) * np.log((1.0 + eta) / eta) - 2.0 * eta)


def moments(eta: float, l_max: int = L_MAX) -> np.ndarray:
    t_max = np.log((1.0 + eta) / eta)
    t = (np.arange(QUADRATURE) + 0.5) * (t_max / QUADRATURE)
    et = np.exp(t)
    xi = eta * (et - 1.0)
    w = eta * et * (t_max / QUADRATURE)
    mu = 1.0 - 2.0 * eta * (1.0 - xi) / (xi + eta)
    out = np.empty(l_max + 1)
    out[0] = 0.0
    total = w.sum()
    p0, p1 = np.ones_like(mu), mu.copy()
    for ell in range(1, l_max + 1):
        out[ell] = float(total - np.dot(p1, w))
        p0, p1 = p1, ((2 * ell + 1) * mu * p1 - ell * p0) / (ell + 1)
    return out


def series_cdf(lam: float, g: np.ndarray, mu: np.ndarray) -> np.ndarray:
    unscattered = math.exp(-lam)
    coef = np.exp(-lam * g) - unscattered
    sig = np.flatnonzero(np.abs(coef) > SERIES_CUTOFF)
    l_max = int(sig[-1]) if sig.size else 0
    cdf = (1.0 - unscattered) * 0.5 * (1.0 + mu)
    if l_max >= 1:
        p_prev, p_curr = np.ones_like(mu), mu.copy()
        p_next = 1.5 * mu * p_curr - 0.5 * p_prev
        for ell in range(1, l_max + 1):
            cdf += 0.5 * coef[ell] * (p_next - p_prev)
            p_prev, p_curr = p_curr, p_next
            n = ell + 1
            p_next = ((2 * n + 1) * mu * p_curr - n * p_prev) / (n + 1)
    np.clip(cdf, 0.0, 1.0, out=cdf)
    np.maximum.accumulate(cdf, out=cdf)
    cdf[0] = 0.0
    cdf[-1] = 1.0
    return cdf


def _compose(a: np.ndarray, b: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    sa = np.sqrt(np.maximum(0.0, 1.0 - a * a))
    sb = np.sqrt(np.maximum(0.0, 1.0 - b * b))
    phi = rng.uniform(0.0, 2.0 * np.pi, size=a.size)
    return np.clip(a * b + sa * sb * np.cos(phi), -1.0, 1.0)


def mc_cosines(eta: float, lam: float, n: int, rng: np.random.Generator) -> np.ndarray:
    doublings = max(0, math.ceil(math.log2(lam)) if lam > 1.0 else 0)
    lam0 = lam / (2.0**doublings)
    counts = rng.poisson(lam0, size=n)
    acc = np.ones(n)
    for step in range(int(counts.max()) if counts.size else 0):
        active = counts > step
        k = int(active.sum())
        if k == 0:
            break
        xi = rng.random(k)
        inv = xi / (2.0 * eta * (1.0 + eta)) + 1.0 / (2.0 + 2.0 * eta)
        acc[active] = _compose(acc[active], 1.0 + 2.0 * eta - 1.0 / inv, rng)
    for _ in range(doublings):
        acc = _compose(acc, rng.permutation(acc), rng)
    return acc


def raw_cosines(eta: float, theta2: float, g_cache: dict) -> np.ndarray:
    lam = 0.5 * theta2 / g1(eta)
    if lam < SERIES_MIN_LAMBDA:
        return np.sort(mc_cosines(eta, lam, RAW_SAMPLES, np.random.default_rng(SEED)))
    g = g_cache.get(eta)
    if g is None:
        g = moments(eta)
        g_cache[eta] = g
    omm = np.geomspace(1.0e-12, 2.0, MU_BINS)
    mu = np.clip(1.0 - omm, -1.0, 1.0)[::-1]
    mu[0], mu[-1] = -1.0, 1.0
    cdf = series_cdf(lam, g, mu)
    u = (np.arange(RAW_SAMPLES) + 0.5) / RAW_SAMPLES
    lo = int(np.flatnonzero(cdf <= 0.0)[-1]) if np.any(cdf <= 0.0) else 0
    hi = int(np.flatnonzero(cdf >= 1.0)[0]) if np.any(cdf >= 1.0) else cdf.size - 1
    if hi <= lo:
        lo, hi = 0, cdf.size - 1
    sc, sm = cdf[lo : hi + 1], mu[lo : hi + 1]
    keep = np.concatenate(([True], np.diff(sc) > 0.0))
    return np.interp(u, sc[keep], sm[keep])


def node(eta: float, theta2: float, n_u: int = 512, g_cache: dict | None = None) -> np.ndarray:
    raw = raw_cosines(eta, theta2, {} if g_cache is None else g_cache)
    omm = (1.0 - raw).reshape(n_u, raw.size // n_u).mean(axis=1)
    mean = float(omm.mean())
    return np.zeros(n_u) if mean <= 0.0 else omm / mean


def window(log_eta_min: float, log_eta_max: float, theta2_max: float, bins: float = 4.0):
    return (math.floor(log_eta_min * bins), math.floor(log_eta_max * bins) + 1,
            math.floor(math.log(1.0e-12) * bins), math.floor(math.log(theta2_max) * bins) + 1)


def grid(ix0: int, ix1: int, iy0: int, iy1: int, n_u: int = 512, bins: float = 4.0,
         workers: int = 8) -> np.ndarray:
    def column(i: int) -> np.ndarray:
        eta = math.exp(i / bins)
        cache: dict = {}
        return np.stack([node(eta, math.exp(j / bins), n_u, cache) for j in range(iy0, iy1 + 1)])

    with ThreadPoolExecutor(max_workers=workers) as pool:
        cols = list(pool.map(column, range(ix0, ix1 + 1)))
    return np.stack(cols)
