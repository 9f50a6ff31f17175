"""Every scalar sampler of the transport rounds, drawn on the GPU, against the analytic
law it samples (the reference's device sampler tests, tests/physics/
test_warp_physics.py:183-263 and test_warp_gs_sampler.py:139-201).

``hmc_sample_physics`` runs the rounds' own device functions (hipmc_device.h) from
independent streams. Distributions are judged by KS or chi-squared tests at 1 percent
against closed forms (exponential free path, channel frequencies, the Klein-Nishina
spectrum, the Moller cross section, the Fermi-Eyges hinge, the log-uniform
bremsstrahlung spectrum, the Thomson coherent law, the uniform sphere); deterministic
companions by kinematic identities (Compton and Moller momentum conservation, unit
norms). The Goudsmit-Saunderson sampler is checked against a synthetic grid whose rows
are known, where every draw must be one of the 512 bilinear blends of its four rows,
each chosen with probability 1/512, and against a real window built by the product's
builder, where the mean of 1 - cos must equal the requested 1 - exp(-theta^2 / 2).
"""

from __future__ import annotations

import ctypes
import itertools
import math

import numpy as np
import pytest
from scipy import integrate, stats

pytestmark = pytest.mark.gpu

N = 200_000
ME = 0.51099895069
ARGS = dict(path_mu=0.05, channel_mu=(0.55, 0.25, 0.12, 0.08), compton_energy=(1.25, 6.0),
            moller_energy=2.0, moller_cut=0.2, hinge_ms=0.01, brems_energy=2.0,
            brems_radiative_loss=0.02, brems_pcut=0.05, coherent_energy=0.3, coherent_material=0,
            gs_log_eta=-10.9375, gs_ms=(1.0, math.exp(0.125)), rotate_cos=0.3, rotate_phi=1.1)
# the synthetic GS grid: 3 x 3 nodes at eta = e^(-11 + i/4), theta^2 = e^(j/4)
GS_IX0, GS_IY0, GS_NU = -44, 0, 512


def _gs_rows() -> np.ndarray:
    """Known rows of mean exactly 1 (to float32 rounding): 1 + a_ij cos(pi (k + 1/2) / 512)."""
    k = np.arange(GS_NU)
    wave = np.cos(np.pi * (k + 0.5) / GS_NU)
    rows = np.empty((3, 3, GS_NU))
    for i in range(3):
        for j in range(3):
            rows[i, j] = 1.0 + (0.3 + 0.1 * (i + 3 * j)) * wave
    return rows.astype(np.float32)


def _draw(ctx, args: dict, seed: int = 20261203, n: int = N) -> np.ndarray:
    from pyradmc_amd import _native as nat

    a = nat.SamplerArgs()
    for key, value in args.items():
        field = getattr(a, key)
        if isinstance(value, tuple):
            for k, v in enumerate(value):
                field[k] = v
        else:
            setattr(a, key, value)
    out = np.empty((n, nat.SAMPLE_COLUMNS), dtype=np.float32)
    ctx.call("hmc_sample_physics", ctypes.c_uint64(seed), n, ctypes.byref(a), nat.ptr(out, ctypes.c_float))
    return out


@pytest.fixture(scope="module")
def samples():
    """Draws with analytic water tables and the synthetic GS grid uploaded."""
    from pyradmc_amd import HipEngine, VoxelGrid, _native as nat
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    grid = VoxelGrid.uniform_water((4, 4, 4), (1.0, 1.0, 1.0))
    eng = HipEngine(grid, AnalyticCrossSections())
    eng._upload_tables(6.0, 0.05, 0.2, with_gs=False)
    rows = np.ascontiguousarray(_gs_rows())
    eng._ctx.call("hmc_set_gs", nat.ptr(rows, ctypes.c_float), 3, 3, GS_NU, GS_IX0, GS_IY0)
    out = _draw(eng._ctx, ARGS)
    yield out.astype(np.float64)
    eng.close()


def _histogram_chi2(x: np.ndarray, pdf, lo: float, hi: float, bins: int = 32) -> float:
    edges = np.linspace(lo, hi, bins + 1)
    observed, _ = np.histogram(x, bins=edges)
    norm = integrate.quad(pdf, lo, hi)[0]
    expected = np.array([integrate.quad(pdf, a, b)[0] for a, b in itertools.pairwise(edges)]) / norm * x.size
    assert np.all(expected > 20.0)
    return float(stats.chisquare(observed, expected * observed.sum() / expected.sum()).pvalue)


def test_free_path_is_exponential(samples):
    s = samples[:, 0]
    assert np.all(s > 0.0)
    assert stats.kstest(s, "expon", args=(0.0, 1.0 / ARGS["path_mu"])).pvalue > 0.01


def test_channel_frequencies_follow_the_coefficients(samples):
    c = samples[:, 1].astype(int)
    mu = np.array(ARGS["channel_mu"])
    observed = np.bincount(c, minlength=4)
    assert observed.size == 4
    assert stats.chisquare(observed, mu / mu.sum() * c.size).pvalue > 0.01


def _kn(r: float, alpha: float) -> float:
    cos_t = 1.0 + 1.0 / alpha - 1.0 / (alpha * r)
    return r + 1.0 / r - (1.0 - cos_t * cos_t)


@pytest.mark.parametrize("column,energy", [(2, 1.25), (5, 6.0)])
def test_compton_ratio_follows_klein_nishina(samples, column, energy):
    alpha = energy / ME
    r = samples[:, column]
    r_min = 1.0 / (1.0 + 2.0 * alpha)
    assert r.min() >= r_min - 1e-6 and r.max() <= 1.0 + 1e-6
    assert _histogram_chi2(r, lambda x: _kn(x, alpha), r_min, 1.0) > 0.01


def test_compton_kinematics_conserve_momentum(samples):
    e = ARGS["compton_energy"][0]
    alpha = e / ME
    r, cos_g, cos_e = samples[:, 2], samples[:, 3], samples[:, 4]
    # the photon angle reproduces the Compton relation
    np.testing.assert_allclose(1.0 / (1.0 + alpha * (1.0 - cos_g)), r, rtol=1e-4)
    # photon and electron momenta add up to the incoming photon's (where the recoil is
    # large enough for the electron angle to be well conditioned in float32)
    used = r < 1.0 - 0.05 / e
    assert used.sum() > 0.8 * r.size
    eg = e * r[used]
    t = e - eg
    pe = np.sqrt(t * (t + 2.0 * ME))
    longitudinal = eg * cos_g[used] + pe * cos_e[used]
    np.testing.assert_allclose(longitudinal, e, rtol=2e-4)
    transverse = eg * np.sqrt(1.0 - cos_g[used] ** 2) - pe * np.sqrt(np.clip(1.0 - cos_e[used] ** 2, 0.0, None))
    np.testing.assert_allclose(transverse, 0.0, atol=2e-3 * e)


def test_isotropic_directions_and_rotation(samples):
    np.testing.assert_allclose(samples[:, 6], 1.0, atol=1e-5)
    assert stats.kstest(samples[:, 7], "uniform", args=(-1.0, 2.0)).pvalue > 0.01
    np.testing.assert_allclose(samples[:, 8], 1.0, atol=1e-5)
    # a direction along the pole turns through the guarded branch to cos theta exactly
    np.testing.assert_allclose(samples[:, 21], ARGS["rotate_cos"], rtol=1e-6)


def _moller_dcs(w: float, e: float) -> float:
    """Moller cross section per unit energy transfer, up to a constant."""
    tau = e / ME
    x = w / e
    return (1.0 / (x * x)) * (1.0 + (x / (1.0 - x)) ** 2 + (tau / (tau + 1.0)) ** 2 * x * x
                              - (2.0 * tau + 1.0) / (tau + 1.0) ** 2 * x / (1.0 - x))


def test_moller_delta_energy_follows_the_cross_section(samples):
    e, cut = ARGS["moller_energy"], ARGS["moller_cut"]
    w = samples[:, 9]
    assert w.min() >= cut * (1 - 1e-6) and w.max() <= 0.5 * e * (1 + 1e-6)
    assert _histogram_chi2(w, lambda x: _moller_dcs(x, e), cut, 0.5 * e) > 0.01


def test_moller_kinematics_conserve_momentum(samples):
    e = ARGS["moller_energy"]
    two_m = 2.0 * ME
    w, cd, cp = samples[:, 9], samples[:, 10], samples[:, 11]
    p_in = math.sqrt(e * (e + two_m))
    p_d = np.sqrt(w * (w + two_m))
    p_p = np.sqrt((e - w) * (e - w + two_m))
    np.testing.assert_allclose(p_d * cd + p_p * cp, p_in, rtol=1e-4)
    np.testing.assert_allclose(p_d * np.sqrt(1 - cd**2) - p_p * np.sqrt(1 - cp**2), 0.0, atol=2e-3 * p_in)


def test_gaussian_hinge_angle_squared_is_exponential(samples):
    theta_sq = np.arccos(np.clip(samples[:, 12], -1.0, 1.0)) ** 2
    assert stats.kstest(theta_sq, "expon", args=(0.0, ARGS["hinge_ms"])).pvalue > 0.01


def test_bremsstrahlung_step_and_spectrum(samples):
    e, loss, pcut = ARGS["brems_energy"], ARGS["brems_radiative_loss"], ARGS["brems_pcut"]
    # photons of log-uniform energy in (pcut, E) carry the radiative loss above pcut:
    # the emission probability is that loss over their mean energy
    mean_k = (e - pcut) / math.log(e / pcut)
    np.testing.assert_allclose(samples[:, 13], loss * (1.0 - pcut / e) / mean_k, rtol=3e-5)
    np.testing.assert_allclose(samples[:, 14], loss * pcut / e, rtol=3e-5)
    k = samples[:, 15]
    assert k.min() >= pcut * (1 - 1e-6) and k.max() <= e * (1 + 1e-6)
    u = np.log(k / pcut) / math.log(e / pcut)
    assert stats.kstest(u, "uniform").pvalue > 0.01


def test_coherent_cosine_with_a_flat_form_factor_is_thomson(samples):
    mu = samples[:, 16]
    assert np.all(np.abs(mu) <= 1.0)
    assert _histogram_chi2(mu, lambda x: 1.0 + x * x, -1.0, 1.0) > 0.01


def _gs_expected(fx: float, fy: float) -> np.ndarray:
    """The 512 bilinear blends the sampler may return at (fx, fy) (float64)."""
    rows = _gs_rows().astype(np.float64)
    ix = int(math.floor(fx)) - GS_IX0
    iy = int(math.floor(fy)) - GS_IY0
    tx = fx - (GS_IX0 + ix)
    ty = fy - (GS_IY0 + iy)
    lo = rows[ix, iy] * (1 - tx) + rows[ix + 1, iy] * tx
    hi = rows[ix, iy + 1] * (1 - tx) + rows[ix + 1, iy + 1] * tx
    return lo * (1 - ty) + hi * ty


@pytest.mark.parametrize("column,which", [(17, 0), (18, 1)])
def test_gs_draws_are_the_grid_rows_blended(samples, column, which):
    ms = ARGS["gs_ms"][which]
    omc = -math.expm1(-0.5 * ms)
    w = (1.0 - samples[:, column]) / omc
    fx = ARGS["gs_log_eta"] * 4.0
    fy = float(np.float32(math.log(ms))) * 4.0
    expected = _gs_expected(fx, fy)
    # every draw is one of the 512 blends (float32 rounding of the blend and the cosine)
    order = np.argsort(expected)
    pos = np.clip(np.searchsorted(expected[order], w), 1, GS_NU - 1)
    left, right = expected[order][pos - 1], expected[order][pos]
    nearest = np.where(np.abs(w - left) <= np.abs(w - right), pos - 1, pos)
    gap = np.abs(w - expected[order][nearest])
    assert gap.max() < 2e-5 * 2.0 / omc
    # ... chosen uniformly over the 512 u-bins by one uniform
    counts = np.bincount(order[nearest], minlength=GS_NU)
    assert stats.chisquare(counts).pvalue > 0.01
    # and the first moment is the requested one (the rows have mean 1)
    one_minus = 1.0 - samples[:, column]
    sem = one_minus.std() / math.sqrt(one_minus.size)
    assert abs(one_minus.mean() - omc) < 4.0 * sem + 1e-5 * omc


def test_gs_and_gaussian_hinge_each_draw_one_uniform(samples):
    np.testing.assert_array_equal(samples[:, 19], samples[:, 20])


def test_gs_first_moment_on_a_built_window():
    # A real window (two eta columns around log eta = -11.45, theta^2 up to 0.08, the
    # low-MeV soft-tissue regime of the reference's own test) from the product's node
    # builder: the mean of 1 - cos equals 1 - exp(-theta^2 / 2) off the nodes.
    from pyradmc_amd import HipEngine, VoxelGrid, _native as nat
    from pyradmc_amd.data.analytic import AnalyticCrossSections
    from pyradmc_amd.data.gs import build_gs_grid

    g = build_gs_grid(-11.45, -11.45, 0.08)
    values = np.ascontiguousarray(g.values, dtype=np.float32)
    grid = VoxelGrid.uniform_water((4, 4, 4), (1.0, 1.0, 1.0))
    eng = HipEngine(grid, AnalyticCrossSections())
    eng._upload_tables(6.0, 0.05, 0.2, with_gs=False)
    eng._ctx.call("hmc_set_gs", nat.ptr(values, ctypes.c_float), values.shape[0], values.shape[1],
                  values.shape[2], g.ix0, g.iy0)
    args = dict(ARGS, gs_log_eta=-11.45, gs_ms=(0.02, 0.05))
    out = _draw(eng._ctx, args, seed=7).astype(np.float64)
    sub = _draw(eng._ctx, dict(ARGS, gs_log_eta=-11.45, gs_ms=(1.0e-13, 1.0e-13)), seed=8, n=20_000)
    eng.close()
    for column, theta2 in ((17, 0.02), (18, 0.05)):
        one_minus = 1.0 - out[:, column]
        expected = -math.expm1(-0.5 * theta2)
        sem = one_minus.std() / math.sqrt(one_minus.size)
        assert abs(one_minus.mean() - expected) < 4.0 * sem + 1e-3 * expected, (theta2, one_minus.mean(), expected)
        assert np.all(out[:, column] >= -1.0) and np.all(out[:, column] <= 1.0)
    # a sub-floor step is exactly forward: never NaN, never backwards
    assert np.all(sub[:, 17] == 1.0) and np.all(sub[:, 18] == 1.0)


def test_sampler_arguments_are_checked():
    from pyradmc_amd import _native as nat

    ctx = nat.Context(0)
    with pytest.raises(ValueError):
        _draw(ctx, dict(ARGS, moller_energy=0.3), n=10)  # below twice the cut
    with pytest.raises(ValueError):
        _draw(ctx, dict(ARGS, channel_mu=(0.0, 0.0, 0.0, 0.0)), n=10)
    out = _draw(ctx, ARGS, n=64)  # no tables: the table columns are NaN, the rest drawn
    ctx.close()
    assert np.isnan(out[:, 16:20]).all() and np.isfinite(out[:, :16]).all()
