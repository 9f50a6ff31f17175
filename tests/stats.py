"""Statistical consistency oracles for Monte Carlo dose comparisons.

The z-score binomial test, the exact-sigma chi-squared test, and the batched
(Welch-Satterthwaite F-moment) chi-squared test of the reference's test suite
(pyRadMC ``tests/conftest.py:45-196``), restated for this repository's tests.
"""

from __future__ import annotations

import numpy as np
from scipy import stats


def z_scores(a, sigma_a, b, sigma_b, mask):
    denom = np.sqrt(sigma_a[mask] ** 2 + sigma_b[mask] ** 2)
    if np.any(denom <= 0.0):
        raise ValueError("zero uncertainty in the compared region")
    return (a[mask] - b[mask]) / denom


def assert_z_consistent(a, sigma_a, b, sigma_b, k=3.0, alpha=0.01, mask=None):
    if mask is None:
        mask = np.ones(a.shape, dtype=bool)
    z = np.abs(z_scores(a, sigma_a, b, sigma_b, mask))
    n = int(z.size)
    n_out = int(np.count_nonzero(z > k))
    p_expected = float(2.0 * stats.norm.sf(k))
    p_value = float(stats.binomtest(n_out, n, p_expected, alternative="greater").pvalue)
    assert p_value > alpha, (
        f"{n_out}/{n} voxels exceed z={k} (expected ~{p_expected * n:.1f}); "
        f"binomial p={p_value:.2e} < alpha={alpha}"
    )


def assert_chi2_consistent(a, sigma_a, b, sigma_b, alpha=0.01, mask=None):
    if mask is None:
        mask = np.ones(a.shape, dtype=bool)
    denom = sigma_a[mask] ** 2 + sigma_b[mask] ** 2
    chi2 = float(np.sum((a[mask] - b[mask]) ** 2 / denom))
    dof = int(np.count_nonzero(mask))
    p_value = float(stats.chi2.sf(chi2, dof))
    assert p_value > alpha, f"chi2={chi2:.1f} on {dof} dof, p={p_value:.2e} < alpha={alpha}"


def batched_chi2_pvalue(a, sigma_a, n_a, b, sigma_b, n_b, mask=None):
    if min(n_a, n_b) < 6:
        raise ValueError("need at least 6 batches per side for a calibrated batched chi2")
    if mask is None:
        mask = np.ones(a.shape, dtype=bool)
    va = sigma_a[mask] ** 2
    vb = sigma_b[mask] ** 2
    if np.any(va + vb <= 0.0):
        raise ValueError("zero uncertainty in the compared region")
    nu = (va + vb) ** 2 / (va**2 / (n_a - 1) + vb**2 / (n_b - 1))
    z_sq = (a[mask] - b[mask]) ** 2 / (va + vb)
    chi2 = float(np.sum(z_sq))
    mean = float(np.sum(nu / (nu - 2.0)))
    variance = float(np.sum(2.0 * nu**2 * (nu - 1.0) / ((nu - 2.0) ** 2 * (nu - 4.0))))
    z_score = (chi2 - mean) / np.sqrt(variance)
    return float(stats.norm.sf(z_score)), chi2, mean, float(np.sqrt(variance))


def assert_chi2_consistent_batched(a, sigma_a, n_a, b, sigma_b, n_b, alpha=0.01, mask=None):
    p, chi2, mean, sd = batched_chi2_pvalue(a, sigma_a, n_a, b, sigma_b, n_b, mask)
    n = int(np.count_nonzero(mask)) if mask is not None else a.size
    assert p > alpha, (
        f"chi2={chi2:.1f} vs batched-null mean {mean:.1f} (sd {sd:.1f}) on {n} voxels, "
        f"p={p:.2e} < alpha={alpha}"
    )


def north_star_gate(a, sigma_a, b, sigma_b, frac_max=0.05, rel_floor=0.10):
    """At most ``frac_max`` of voxels above ``rel_floor`` of max dose beyond |z| > 2."""
    mask = b > rel_floor * b.max()
    mask &= (sigma_a > 0) | (sigma_b > 0)
    z = np.abs(z_scores(a, sigma_a, b, sigma_b, mask))
    frac = float(np.mean(z > 2.0))
    return frac, int(mask.sum())


def comparison_statistics(truth, truth_sigma, truth_batches, d, s, n_batches, mask) -> dict:
    """Summary of a sample ``d`` (sigma ``s``) against a precise run ``truth``."""
    _, chi2, mean, _ = batched_chi2_pvalue(truth, truth_sigma, truth_batches, d, s, n_batches, mask)
    z = z_scores(truth, truth_sigma, d, s, mask)
    return {"chi2_ratio": chi2 / mean, "frac_z2": float(np.mean(np.abs(z) > 2.0)),
            "mean_z": float(z.mean()), "total_ratio": float(d[mask].sum() / truth[mask].sum())}


def assert_within_replicates(ref: dict, replicates: list[dict], k: float = 4.0) -> None:
    """Each statistic of the reference sample within ``k`` replicate standard deviations
    of the replicate mean.

    Replicates are samples of the engine under test drawn with the reference sample's
    exact history/batch structure, each summarised against the same precise run as the
    reference sample. With few histories per batch, per-voxel batch doses are skewed and
    a low batch mean comes with a small batch sigma: normal-theory statistics then run
    biased (chi-squared a few percent high, more than 5 percent of voxels beyond
    |z| = 2) even between identical physics, and this calibration absorbs that.
    """
    for key, value in ref.items():
        v = np.array([rep[key] for rep in replicates])
        sd = v.std(ddof=1)
        t = (value - v.mean()) / sd if sd > 0 else 0.0
        assert abs(t) < k, f"{key}: reference {value:.4f}, replicates {v.mean():.4f} +- {sd:.4f}"
