"""Gaussian focal-spot sources on the GPU (in-kernel generator).

* A spot fan against the oracle's spot generator: batched chi-squared, acceptance gate.
* The in-kernel generator against the host-sampled route of the same source (its
  ``sample_batch``, through a subclass), open and seen through a jaw stack.
* A spot beamlet fan's Dij against the oracle, column by column.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.stats import assert_z_consistent, batched_chi2_pvalue, north_star_gate

pytestmark = pytest.mark.gpu

ALPHA = 0.01
FOCAL = (4.0, 4.0, -100.0)


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


def _spot(cls=None, sigma=0.3):
    from pyradmc_amd import GaussianSpotBeamSource, bremsstrahlung_6mv_like

    cls = cls or GaussianSpotBeamSource
    return cls(bremsstrahlung_6mv_like(), FOCAL, (4.0, 4.0, -50.0), 2.0, 2.0, sigma, 0.5 * sigma)


def _host_spot_class():
    from pyradmc_amd import GaussianSpotBeamSource

    class HostSpot(GaussianSpotBeamSource):
        """A subclass: the engine host-samples it through sample_batch."""

    return HostSpot


def _compare(a, b, batches_a, batches_b, gate=True):
    bd, bs = (b.dose, b.dose_sigma) if hasattr(b, "dose") else (b["dose"], b["dose_sigma"])
    mask = bd > 0.1 * bd.max()
    p, chi2, mean, sd = batched_chi2_pvalue(a.dose, a.dose_sigma, batches_a, bd, bs, batches_b, mask)
    print(f"  chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    if gate:
        frac, count = north_star_gate(a.dose, a.dose_sigma, bd, bs)
        assert count > 200 and frac < 0.06
        assert_z_consistent(a.dose, a.dose_sigma, bd, bs, k=3.0, mask=mask)


def test_spot_fan_matches_oracle(engine, oracle_lib):
    print("spot fan vs oracle", flush=True)
    src = _spot()
    hip = engine.run(src, 4_000_000, 20, seed=21)
    ora = oracle_lib.for_grid(engine.grid, src).run(4_000_000, 20, seed=22, threads=16)
    _compare(hip, ora, 20, 20)
    assert hip.energy_emitted == pytest.approx(4_000_000 * src.spectrum.mean_energy, rel=2e-3)
    assert hip.energy_emitted == pytest.approx(ora["books"]["emitted"], rel=3e-3)


def test_in_kernel_spot_matches_host_sampling_open_and_collimated(engine, grid_xs):
    from pyradmc_amd.collimation import BeamFrame, BeamLimitingStack, CollimatedSource, JawPair

    print("spot fan: in-kernel vs host-sampled", flush=True)
    host = _host_spot_class()
    # ten batches a side: the batched chi-squared is calibrated for them, the z-score
    # tail counts (normal tails) are not, so only the former applies here
    a = engine.run(_spot(), 2_000_000, 10, seed=31)
    b = engine.run(_spot(host), 2_000_000, 10, seed=32)
    _compare(a, b, 10, 10, gate=False)
    stack = BeamLimitingStack(BeamFrame(FOCAL), (JawPair("u", 20.0, 27.0, -0.5, 0.3, 0, 19.3, focused=True),))
    c = engine.run(CollimatedSource(_spot(), stack, grid_xs[1]), 2_000_000, 10, seed=33)
    d = engine.run(CollimatedSource(_spot(host), stack, grid_xs[1]), 2_000_000, 10, seed=34)
    _compare(c, d, 10, 10, gate=False)
    assert c.energy_emitted < 0.8 * a.energy_emitted  # the jaws do cut the field
    assert c.energy_emitted == pytest.approx(d.energy_emitted, rel=0.01)


def test_spot_beamlet_dij_matches_oracle(engine, oracle_lib):
    from pyradmc_amd import GaussianSpotBeamletSource, Spectrum

    sp = Spectrum([1.0, 2.0, 3.0], [2.0, 1.0])
    centers = [(3.5, 4.0, -50.0), (4.5, 4.0, -50.0)]
    src = GaussianSpotBeamletSource(sp, FOCAL, centers, 0.5, 0.5, 0.2, 0.2)
    n = 800_000
    hip = engine.run_dij(src, n, 10, seed=41, truncation=0.0)
    ora = oracle_lib.for_grid(engine.grid, src).run_dij(np.full(2, n), 10, seed=42, truncation=0.0, threads=16)
    for j in range(2):
        da, sa = hip.column_dense(j).ravel(), hip.sigma_dense(j).ravel()
        db, sb = ora["dense_dose"][j], ora["dense_sigma"][j]
        mask = db > 0.1 * db.max()
        p, *_ = batched_chi2_pvalue(da, sa, 10, db, sb, 10, mask)
        assert p > ALPHA, (j, p)
