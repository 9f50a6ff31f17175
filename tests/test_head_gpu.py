"""Treatment-head pre-solve on the GPU.

* Attenuation mode is deterministic given the primaries: on host-sampled primaries the
  device population equals the oracle's (oracle/head_ref.py) record for record, to
  float32; primaries drawn in the kernel give the same population in distribution.
* First-Compton mode draws random numbers: the device and oracle populations agree in
  their weight totals per particle type and in the photons' weight-energy histogram.
* The same seed gives the same population, copied back or left resident.
* Transported, an attenuation-mode pre-solve gives the dose of the collimated source it
  restates, and the resident route the dose of the copied-back population; forced first
  Compton adds scatter dose under the jaws.
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest

from oracle import head_ref as H
from tests.stats import batched_chi2_pvalue, north_star_gate

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:.*latent variance:UserWarning")]

ALPHA = 0.01
FOCAL = (4.0, 4.0, -100.0)
EXIT_W = 97.0  # exit plane 3 cm above the phantom face


def _head(xs, host_sampled: bool = False):
    """The test head: jaws and an MLC, a spot fan above them (host-sampled through its
    sample_batch when asked, else generated in the pre-solve kernel)."""
    from pyradmc_amd import GaussianSpotBeamSource, Spectrum
    from pyradmc_amd.collimation import MLC, BeamFrame, BeamLimitingStack, JawPair

    class HostSampledSpot(GaussianSpotBeamSource):
        """A subclass: the pre-solve samples it on the host."""

    frame = BeamFrame(FOCAL)
    stack = BeamLimitingStack(frame, (
        JawPair("u", 20.0, 28.0, -0.6, 0.5, 0, 19.3),
        JawPair("v", 30.0, 36.0, -0.5, 0.6, 0, 19.3, focused=True),
        MLC(38.0, 45.0, tuple(np.linspace(-1.2, 1.2, 7)), (-0.5, -0.7, -0.9, -0.6, -0.4, -0.5),
            (0.3, 0.5, 0.6, 0.4, 0.2, 0.5), 8.0, 0, 18.0),
    ))
    cls = HostSampledSpot if host_sampled else GaussianSpotBeamSource
    src = cls(Spectrum([1.0, 2.0, 3.0], [2.0, 1.0]), FOCAL, (4.0, 4.0, -85.0), 1.5, 1.5, 0.2, 0.2)
    return stack, src


@pytest.fixture(scope="module")
def grid_xs():
    from pyradmc_amd import AnalyticCrossSections, VoxelGrid

    grid = VoxelGrid.uniform_water((16, 16, 16), (0.5, 0.5, 0.5))
    return grid, AnalyticCrossSections(geometry_densities=grid.max_density_by_material())


@pytest.fixture(scope="module")
def engine(grid_xs):
    from pyradmc_amd import HipEngine, _native

    _native.lib()
    eng = HipEngine(*grid_xs)
    yield eng
    eng.close()


def test_attenuation_mode_matches_the_oracle_record_for_record(grid_xs):
    from pyradmc_amd.head import presolve_head

    xs = grid_xs[1]
    stack, src = _head(xs, host_sampled=True)  # the oracle's primaries
    n = 200_000
    dev = presolve_head(src, cross_sections=xs, n_histories=n, seed=3, exit_z=EXIT_W, stack=stack,
                        mode="attenuation").columns()
    ora = H.presolve(src.sample_batch(3, 0, n), cross_sections=xs, max_energy=src.max_energy, seed=3, exit_z=EXIT_W,
                     stack=stack, mode="attenuation")["population"]
    assert dev["weight"].size == ora["weight"].size == n
    np.testing.assert_array_equal(dev["particle_type"], 1)
    np.testing.assert_allclose(dev["weight"], ora["weight"], rtol=2e-6, atol=1e-30)
    np.testing.assert_allclose(dev["energy"], ora["energy"], rtol=1e-6)
    for k in ("x", "y", "z"):
        np.testing.assert_allclose(dev[k], ora[k], atol=2e-4)
    assert 0.0 < dev["weight"].min() < 1e-3 and dev["weight"].max() == pytest.approx(1.0)


def _by_type(pop, n):
    raw = pop["weight"] * n / pop["weight"].size
    out = {}
    for code in (1, 2):
        m = pop["particle_type"] == code
        out[code] = (raw[m], pop["energy"][m])
    return out


def test_first_compton_population_matches_the_oracle_statistically(grid_xs):
    from pyradmc_amd.head import AirColumn, presolve_head

    xs = grid_xs[1]
    stack, src = _head(xs)
    n = 400_000
    air = AirColumn(50.0, 96.0, material=0, density=0.02)  # dense "air" (water) for a visible component
    dev = presolve_head(src, cross_sections=xs, n_histories=n, seed=4, exit_z=EXIT_W, stack=stack, air=air,
                        weight_floor=0.0).columns()
    ora = H.presolve(src.sample_batch(4, 0, n), cross_sections=xs, max_energy=src.max_energy, seed=4,
                     exit_z=EXIT_W, stack=stack, air=dict(z_start=50.0, z_end=96.0, material=0, density=0.02),
                     weight_floor=0.0)["population"]
    a, b = _by_type(dev, n), _by_type(ora, n)
    for code in (1, 2):
        (wa, ea), (wb, eb) = a[code], b[code]
        sa, sb = wa.sum(), wb.sum()
        sd = np.sqrt((wa ** 2).sum() + (wb ** 2).sum())
        print(f"  type {code}: device {sa:.2f} ({wa.size}) oracle {sb:.2f} ({wb.size}) sd {sd:.3f}", flush=True)
        assert abs(sa - sb) < 5.0 * sd
        ha = np.histogram(ea, bins=12, range=(0.0, 3.0), weights=wa)[0]
        hb = np.histogram(eb, bins=12, range=(0.0, 3.0), weights=wb)[0]
        va = np.histogram(ea, bins=12, range=(0.0, 3.0), weights=wa ** 2)[0]
        vb = np.histogram(eb, bins=12, range=(0.0, 3.0), weights=wb ** 2)[0]
        live = (va + vb) > 0
        chi2 = float(np.sum((ha - hb)[live] ** 2 / (va + vb)[live]))
        assert chi2 < 12 + 6 * np.sqrt(24), (code, chi2)
    assert a[2][0].size > 1000  # contaminant electrons are there


def test_generated_primaries_match_host_sampled_ones(grid_xs):
    # The pre-solve kernel draws a built-in source's primaries itself; the population
    # agrees in distribution with the one from the source's host sample_batch.
    from pyradmc_amd.head import presolve_head

    xs = grid_xs[1]
    n = 400_000
    pops = []
    for host in (False, True):
        stack, src = _head(xs, host_sampled=host)
        pops.append(presolve_head(src, cross_sections=xs, n_histories=n, seed=7, exit_z=EXIT_W, stack=stack,
                                  weight_floor=0.0).columns())
    a, b = _by_type(pops[0], n), _by_type(pops[1], n)
    (wa, ea), (wb, eb) = a[1], b[1]
    assert abs(wa.sum() - wb.sum()) < 5.0 * np.sqrt((wa ** 2).sum() + (wb ** 2).sum())
    ha = np.histogram(ea, bins=12, range=(0.0, 3.0), weights=wa)[0]
    hb = np.histogram(eb, bins=12, range=(0.0, 3.0), weights=wb)[0]
    va = np.histogram(ea, bins=12, range=(0.0, 3.0), weights=wa ** 2)[0]
    vb = np.histogram(eb, bins=12, range=(0.0, 3.0), weights=wb ** 2)[0]
    live = (va + vb) > 0
    assert float(np.sum((ha - hb)[live] ** 2 / (va + vb)[live])) < 12 + 6 * np.sqrt(24)
    # exit-plane positions over the field (photons scattered near 90 degrees reach the
    # plane metres away, so a mean would be all outliers: compare histograms)
    def hist(pop, weights):
        return np.histogram(pop["x"], bins=20, range=(-1.0, 9.0), weights=weights)[0]

    wa_all, wb_all = pops[0]["weight"] * n / pops[0]["weight"].size, pops[1]["weight"] * n / pops[1]["weight"].size
    hx = hist(pops[0], wa_all) - hist(pops[1], wb_all)
    vx = hist(pops[0], wa_all ** 2) + hist(pops[1], wb_all ** 2)
    live = vx > 0
    assert live.sum() >= 15
    assert float(np.sum(hx[live] ** 2 / vx[live])) < 20 + 6 * np.sqrt(40)


def test_same_seed_same_population_copied_back_or_resident(engine, grid_xs):
    from pyradmc_amd.head import presolve_head

    xs = grid_xs[1]
    stack, src = _head(xs)
    kw = dict(cross_sections=xs, n_histories=100_000, seed=5, exit_z=EXIT_W, stack=stack)
    a = presolve_head(src, **kw).columns()
    b = presolve_head(src, engine=engine, **kw).columns()
    dev = presolve_head(src, engine=engine, resident=True, **kw)
    c = dev.to_phase_space().columns()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    for k in ("particle_type", "energy", "x", "y", "z", "weight"):
        np.testing.assert_array_equal(a[k], c[k])
    assert len(dev) == a["weight"].size
    assert engine._population_token == dev.token


def test_presolved_transport_matches_the_collimated_source_and_the_resident_route(engine, grid_xs):
    from pyradmc_amd.collimation import CollimatedSource
    from pyradmc_amd.head import presolve_head

    print("pre-solve transport", flush=True)
    xs = grid_xs[1]
    stack, src = _head(xs)
    m = 2_000_000
    direct = engine.run(CollimatedSource(src, stack, xs), m, 20, seed=61)
    att = presolve_head(src, cross_sections=xs, n_histories=8_000_000, seed=62, exit_z=EXIT_W, stack=stack,
                        mode="attenuation", engine=engine, resident=True)
    ps = engine.run(att, m, 20, seed=63)
    # The batches resample one finite population: its own sampling noise, a variance of
    # about m / N of the run's, is common to every batch and missing from the sigmas.
    latent = np.sqrt(1.0 + m / len(att))
    mask = direct.dose > 0.1 * direct.dose.max()
    p, chi2, mean, sd = batched_chi2_pvalue(ps.dose, ps.dose_sigma * latent, 20, direct.dose, direct.dose_sigma, 20,
                                            mask)
    print(f"  attenuation pre-solve vs collimated: chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    frac, count = north_star_gate(ps.dose, ps.dose_sigma * latent, direct.dose, direct.dose_sigma)
    assert count > 100 and frac < 0.06
    assert ps.energy_emitted == pytest.approx(direct.energy_emitted, rel=0.01)
    # the copied-back population transported on the host route agrees with the resident
    # one (same population: no latent term between them)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = engine.run(att.to_phase_space(), m, 20, seed=64)
    p2, *_ = batched_chi2_pvalue(ps.dose, ps.dose_sigma, 20, host.dose, host.dose_sigma, 20, mask)
    assert p2 > ALPHA


def test_first_compton_adds_dose_under_the_block(engine, grid_xs):
    # A 10 cm water block over half the field: the photons it scatters carry a few
    # percent of the blocked weight into the shielded half (pyradmc's
    # tests/integration/test_head_presolve_transport.py:144-167).
    from pyradmc_amd import GaussianSpotBeamSource, Spectrum
    from pyradmc_amd.collimation import BeamFrame, BeamLimitingStack, JawPair
    from pyradmc_amd.head import presolve_head

    xs = grid_xs[1]
    # the block 10-20 cm above the phantom, so its forward scatter reaches it
    stack = BeamLimitingStack(BeamFrame(FOCAL), (JawPair("u", 80.0, 90.0, -1000.0, 0.0, 0, 1.0),))
    src = GaussianSpotBeamSource(Spectrum([1.5, 2.5], [1.0]), FOCAL, (4.0, 4.0, -70.0), 2.4, 2.4, 0.0, 0.0)
    kw = dict(cross_sections=xs, n_histories=8_000_000, exit_z=EXIT_W, stack=stack, engine=engine, resident=True)
    att = engine.run(presolve_head(src, mode="attenuation", seed=71, **kw), 4_000_000, 20, seed=72)
    fc = engine.run(presolve_head(src, mode="first_compton", seed=73, **kw), 4_000_000, 20, seed=74)
    shielded = np.s_[10:14, 4:12, :6]  # x 5-7 cm: under the block (u > 0 is x > 4)
    a, b = att.dose[shielded].mean(), fc.dose[shielded].mean()
    sa = np.sqrt(np.sum(att.dose_sigma[shielded] ** 2)) / att.dose[shielded].size
    sb = np.sqrt(np.sum(fc.dose_sigma[shielded] ** 2)) / fc.dose[shielded].size
    print(f"  shielded mean dose: attenuation {a:.4g} first Compton {b:.4g} (sd {sa:.2g}, {sb:.2g})", flush=True)
    assert b > a + 3.0 * np.hypot(sa, sb) and b > 1.01 * a
    opened = np.s_[2:6, 4:12, :6]
    assert fc.dose[opened].mean() == pytest.approx(att.dose[opened].mean(), rel=0.03)
