"""Treatment-head pre-solve on the host (CPU): the oracle (oracle/head_ref.py) pinned by
closed forms, and the pre-solve's argument contract (pyradmc's tests/unit/
test_head_presolve.py cases)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import head_ref as H
from pyradmc_amd.collimation import BeamFrame, BeamLimitingStack, JawPair
from pyradmc_amd.head import AirColumn, resolve_presolve

ME = 0.51099895069


def _kn_pdf_ratio(e, r):
    """Klein-Nishina dsigma/dr (unnormalised) in the energy ratio r = E'/E."""
    a = e / ME
    cos_t = 1.0 - (1.0 / r - 1.0) / a
    return (r + 1.0 / r - (1.0 - cos_t ** 2)) / a


@pytest.mark.parametrize("energy", [0.3, 2.0, 6.0])
def test_kahn_sampler_follows_klein_nishina(energy):
    gen = np.random.Generator(np.random.PCG64(1))
    n = 200_000
    r = H.kahn_ratios(np.full(n, energy), gen)
    r_min = 1.0 / (1.0 + 2.0 * energy / ME)
    assert r.min() >= r_min - 1e-12 and r.max() <= 1.0
    edges = np.linspace(r_min, 1.0, 21)
    grid = np.linspace(r_min, 1.0, 20_001)
    pdf = _kn_pdf_ratio(energy, grid)
    cdf = np.concatenate(([0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))))
    cdf /= cdf[-1]
    expect = n * np.diff(np.interp(edges, grid, cdf))
    counts = np.histogram(r, bins=edges)[0]
    chi2 = float(np.sum((counts - expect) ** 2 / expect))
    assert chi2 < 50.0  # 19 degrees of freedom: p ~ 1e-4


def test_rotation_turns_by_the_stated_angle():
    rng = np.random.default_rng(2)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:3] = [[0, 0, 1], [0, 0, -1], [1, 0, 0]]  # both poles
    cos_t = rng.uniform(-1, 1, 500)
    phi = rng.uniform(0, 2 * np.pi, 500)
    out = H.rotate(d, cos_t, phi)
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.sum(out * d, axis=1), cos_t, atol=1e-12)
    # azimuths are uniform: the turned directions average to cos_t * d
    many = H.rotate(np.tile([[0.6, 0.0, 0.8]], (100_000, 1)), np.full(100_000, 0.5),
                    rng.uniform(0, 2 * np.pi, 100_000))
    np.testing.assert_allclose(many.mean(axis=0), [0.3, 0.0, 0.4], atol=5e-3)


def test_electron_and_photon_angles_conserve_momentum():
    e = np.array([0.5, 2.0, 6.0])
    r = np.array([0.6, 0.3, 0.2])
    cg, ce = H.photon_cos(e, r), H.electron_cos(e, r)
    pe = np.sqrt(e * (1 - r) * (e * (1 - r) + 2 * ME))
    # along the incident direction and across it (opposite azimuths)
    np.testing.assert_allclose(r * e * cg + pe * ce, e, rtol=1e-12)
    np.testing.assert_allclose(r * e * np.sqrt(1 - cg ** 2), pe * np.sqrt(1 - ce ** 2), rtol=1e-10)


def test_roulette_preserves_the_expected_weight():
    gen = np.random.Generator(np.random.PCG64(3))
    w = np.full(400_000, 2.5e-4)
    out = H.roulette(w, 1e-3, gen)
    assert set(np.unique(out).tolist()) == {0.0, 1e-3}
    assert out.mean() == pytest.approx(2.5e-4, rel=0.02)
    big = np.full(10, 5e-3)
    np.testing.assert_array_equal(H.roulette(big, 1e-3, gen), big)
    np.testing.assert_array_equal(H.roulette(w[:10], 0.0, gen), w[:10])


def _setup():
    from pyradmc_amd import AnalyticCrossSections, GaussianSpotBeamSource, Spectrum

    xs = AnalyticCrossSections()
    frame = BeamFrame((0.0, 0.0, -100.0))
    stack = BeamLimitingStack(frame, (JawPair("u", 20.0, 28.0, -1.0, 1.2, 0, 19.3),
                                      JawPair("v", 30.0, 36.0, -1.5, 1.0, 0, 19.3, focused=True)))
    src = GaussianSpotBeamSource(Spectrum([1.0, 3.0], [1.0]), frame.origin, (0.0, 0.0, -80.0), 3.0, 3.0, 0.2, 0.2)
    return xs, stack, src


def test_attenuation_mode_is_beer_lambert_exactly():
    xs, stack, src = _setup()
    n = 5000
    cols = src.sample_batch(4, 0, n)
    out = H.presolve(cols, cross_sections=xs, max_energy=3.0, seed=4, exit_z=40.0, stack=stack, mode="attenuation",
                     n_table=512)
    pop = out["population"]
    assert pop["weight"].size == n  # every transmitted primary is forward and above the cut
    e = cols["energy"].astype(np.float64)
    p = np.column_stack([cols["x"], cols["y"], cols["z"]]).astype(np.float64)
    d = np.column_stack([cols["ux"], cols["uy"], cols["uz"]]).astype(np.float64)
    t = stack.path_lengths(p, d)
    mu = H.MuTables(xs, list(zip(stack.materials, stack.densities)), 0.010, 3.0, 512)
    expect = np.exp(-sum(mu.total(i, e) * t[:, i] for i in range(2)))
    np.testing.assert_allclose(pop["weight"], expect, rtol=1e-12)  # kept == n: scale 1
    # every kept particle sits on the exit plane, on its own ray
    pl, _ = stack.frame.to_local(np.column_stack([pop["x"], pop["y"], pop["z"]]), d)
    np.testing.assert_allclose(pl[:, 2], 40.0, atol=1e-9)
    # the mu table reproduces the cross-section source at its nodes
    nodes = np.exp(mu.log_e[[0, 100, 511]])
    np.testing.assert_allclose(mu.total(0, nodes), 19.3 * xs.mu_total_bulk(nodes, 0), rtol=1e-12)


def test_first_compton_population_structure_and_normalisation():
    # Records come in category order: stack photons, air photons, air electrons, then
    # one transmitted photon per primary, whose weight is the attenuation-mode weight
    # times the air column's transmission; the per-primary normalisation multiplies
    # every weight by kept / n.
    xs, stack, src = _setup()
    n = 20_000
    cols = src.sample_batch(5, 0, n)
    air = dict(z_start=40.0, z_end=58.0, material=0, density=0.05)
    kw = dict(cross_sections=xs, max_energy=3.0, seed=5, exit_z=60.0, stack=stack)
    att = H.presolve(cols, mode="attenuation", **kw)["population"]
    fc = H.presolve(cols, mode="first_compton", air=air, weight_floor=0.0, **kw)["population"]
    n_fc = fc["weight"].size
    assert n_fc > 2 * n
    raw = fc["weight"] * n / n_fc
    tail = slice(n_fc - n, n_fc)
    e = cols["energy"].astype(np.float64)
    np.testing.assert_array_equal(fc["energy"][tail], e)
    d = np.column_stack([cols["ux"], cols["uy"], cols["uz"]]).astype(np.float64)
    _, dl = stack.frame.to_local(np.zeros((n, 3)), d)
    mu = H.MuTables(xs, [(0, 19.3), (0, 19.3), (0, 0.05)], 0.010, 3.0, 512)
    tau_air = mu.total(2, e) * (air["z_end"] - air["z_start"]) / dl[:, 2]
    np.testing.assert_allclose(raw[tail], att["weight"] * np.exp(-tau_air), rtol=1e-12)
    # electrons only from the air, below the primary energy; no positrons
    assert set(np.unique(fc["particle_type"]).tolist()) == {1, 2}
    el = fc["particle_type"] == 2
    assert 0 < el.sum() <= n and np.all(fc["energy"][el] < 3.0)
    # without roulette the forced air pair carries the Compton share of the interacting
    # weight: photon and electron arms of one interaction have equal weights
    el_w = np.sort(raw[el])
    assert el_w.max() < 1.0 and el_w.min() > 0.0


def test_resolve_presolve_contract():
    _, stack, _ = _setup()
    assert resolve_presolve(stack, None, None, 40.0, "auto") == ("first_compton", stack.frame)
    with pytest.raises(ValueError, match="mode"):
        resolve_presolve(stack, None, None, 40.0, "fast")
    with pytest.raises(ValueError, match="frame"):
        resolve_presolve(None, None, None, 40.0, "auto")
    with pytest.raises(ValueError, match="above the stack exit"):
        resolve_presolve(stack, None, None, 30.0, "auto")
    with pytest.raises(ValueError, match="inside the stack"):
        resolve_presolve(stack, None, AirColumn(30.0, 50.0), 60.0, "auto")
    with pytest.raises(ValueError, match="air column end"):
        resolve_presolve(stack, None, AirColumn(40.0, 50.0), 45.0, "auto")
    with pytest.raises(ValueError):
        AirColumn(5.0, 4.0)
    frame = BeamFrame((0, 0, 0))
    assert resolve_presolve(None, frame, AirColumn(1.0, 2.0), 3.0, "attenuation") == ("attenuation", frame)


def test_presolve_head_refuses_non_photon_sources():
    from pyradmc_amd.head import presolve_head

    class Electrons:
        max_energy = 2.0

        def sample_batch(self, seed, offset, n):
            return dict(particle_type=np.full(n, 2, np.int32))

    _, stack, _ = _setup()
    with pytest.raises(ValueError, match="photon sources only"):
        presolve_head(Electrons(), cross_sections=None, n_histories=10, seed=1, exit_z=40.0, stack=stack)
    with pytest.raises(ValueError, match="resident"):
        presolve_head(Electrons(), cross_sections=None, n_histories=10, seed=1, exit_z=40.0, stack=stack,
                      resident=True)
    assert math.isfinite(ME)
