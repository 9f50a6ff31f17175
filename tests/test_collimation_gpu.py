"""Collimated and mask-shaped sources on the GPU: the generator's float64 evaluation of
the ray factors, and transport of in-kernel sources seen through jaws and an MLC.

* The device factor of random rays through every device variety (straight and focused
  jaws, MLCs with parallel and focused sides, energies past both table ends) equals the
  oracle's scalar factor (oracle/collim_ref.py) on the same float32 rays (optical depths
  to 1e-12 relative and 1e-11 absolute); the mask factor likewise.
* A collimated parallel field transported with the in-kernel generator agrees with the
  oracle's transport of independently sampled primaries carrying the same factors
  (batched chi-squared and the acceptance gate); the emitted energy is the sum of the
  weighted primaries.
* A collimated beamlet lattice's Dij (in-kernel) agrees column by column with the
  host-sampled route of the same wrapper around a host twin of the lattice.
* A stack wide open over the field changes no dose bit; a uniform 0.5 mask halves the
  emitted energy exactly.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import collim_ref as R
from tests.stats import assert_z_consistent, batched_chi2_pvalue, north_star_gate

pytestmark = pytest.mark.gpu

ALPHA = 0.01
FOCAL = (4.0, 4.0, -100.0)  # focal spot 100 cm above the phantom face


def _stack(edge_scale: float = 1.0):
    from pyradmc_amd.collimation import MLC, BeamFrame, BeamLimitingStack, JawPair

    s = edge_scale
    return BeamLimitingStack(BeamFrame(FOCAL), (
        JawPair("v", 20.0, 27.0, -1.6 * s, 1.5 * s, 0, 19.3),
        JawPair("u", 28.0, 35.0, -1.7 * s, 1.6 * s, 0, 19.3, focused=True),
        MLC(38.0, 45.0, tuple(np.linspace(-2.2, 2.2, 13) * s), tuple(np.linspace(-1.9, -0.8, 12) * s),
            tuple(np.linspace(0.7, 1.8, 12) * s), 8.0, 0, 18.0),
        MLC(46.0, 52.0, tuple(np.array([-2.0, -0.6, 0.4, 2.0]) * s), tuple(np.array([-1.2, -1.6, -1.0]) * s),
            tuple(np.array([1.3, 0.8, 1.5]) * s), 4.0, 0, 18.0, focused_sides=True),
    ))


def _oracle_devices(stack):
    from tests.test_collimation import _oracle_devices as od

    return od(stack.devices)


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


def _device_factors(engine, model, e, p, d):
    from pyradmc_amd import _native as nat

    engine._set_ray_weights(model)
    e = np.ascontiguousarray(e, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(e.size)
    engine._ctx.call("hmc_ray_weight_factors", int(e.size), nat.ptr(e, nat.ctypes.c_float),
                     nat.ptr(p, nat.ctypes.c_float), nat.ptr(d, nat.ctypes.c_float), nat.ptr(out, nat.c_double))
    return out


def _random_rays(n, rng):
    p = np.column_stack([rng.uniform(0, 8, n), rng.uniform(0, 8, n), rng.uniform(-1, 1, n)])
    aim = np.column_stack([rng.normal(FOCAL[0], 0.3, n), rng.normal(FOCAL[1], 0.3, n), np.full(n, FOCAL[2])])
    d = p - aim
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:4] = [[0, 0, 1], [1, 0, 0], [0, 0.6, 0.8], [0.6, 0, 0.8]]
    return p, d


def test_device_stack_factor_matches_the_oracle(engine, grid_xs):
    from pyradmc_amd.collimation import StackTransmission

    rng = np.random.default_rng(11)
    stack = _stack()
    model = StackTransmission(stack, grid_xs[1], 6.0, n_table=256)
    n = 3000
    p, d = _random_rays(n, rng)
    e = rng.uniform(0.004, 7.0, n)  # past both table ends
    e32, p32, d32 = e.astype(np.float32), p.astype(np.float32), d.astype(np.float32)
    got = _device_factors(engine, model, e32, p32, d32)
    dev = _oracle_devices(stack)
    basis = stack.frame.basis.tolist()
    ora = np.array([R.stack_factor(dev, stack.frame.origin, basis, model.log_energies.tolist(),
                                   model.log_mu.tolist(), model.densities.tolist(), float(e32[i]),
                                   p32[i].astype(np.float64), d32[i].astype(np.float64)) for i in range(n)])
    # Compare optical depths: the factor's relative error grows with tau (deep shadow),
    # and a chord grazing a tip cylinder loses digits to cancellation, where the device's
    # fused multiply-adds round differently from the oracle's separate operations.
    zero = ora == 0.0
    np.testing.assert_array_equal(got[zero], 0.0)
    np.testing.assert_allclose(-np.log(got[~zero]), -np.log(ora[~zero]), rtol=1e-12, atol=1e-11)
    partial = (ora > 1e-6) & (ora < 0.999)
    assert partial.sum() > 300 and (ora == 1.0).sum() > 100  # penumbra and open rays both sampled
    engine._set_ray_weights(None)


def test_device_mask_factor_matches_the_oracle(engine):
    from pyradmc_amd.collimation import MaskTransmission

    rng = np.random.default_rng(12)
    mask = rng.random((9, 7))
    model = MaskTransmission(mask, (4.0, 4.0, -50.0), 6.0, 5.0, (1, 1, 0), (-1, 1, 0))
    p, d = _random_rays(2000, rng)
    p32, d32 = p.astype(np.float32), d.astype(np.float32)
    got = _device_factors(engine, model, np.ones(2000, np.float32), p32, d32)
    ora = [R.mask_factor(mask.tolist(), model.center.tolist(), model.widths, model.u_axis.tolist(),
                         model.v_axis.tolist(), p32[i].astype(np.float64), d32[i].astype(np.float64))
           for i in range(2000)]
    np.testing.assert_allclose(got, ora, rtol=1e-12, atol=1e-14)
    assert np.count_nonzero(np.asarray(ora) > 0) > 500
    engine._set_ray_weights(None)


def _field_source():
    from pyradmc_amd import ParallelBeamSource

    return ParallelBeamSource(2.0, -0.1, (2.0, 6.0), (2.0, 6.0))


def test_collimated_field_matches_oracle(engine, grid_xs, oracle_lib):
    from pyradmc_amd.collimation import CollimatedSource

    print("collimated parallel field vs oracle", flush=True)
    grid, xs = grid_xs
    src = CollimatedSource(_field_source(), _stack(), xs)
    n, batches = 4_000_000, 20
    hip = engine.run(src, n, batches, seed=3)
    # the oracle transports independently sampled field primaries with the same factors
    rng = np.random.default_rng(4)
    x, y = rng.uniform(2.0, 6.0, n), rng.uniform(2.0, 6.0, n)
    ones = np.ones(n)
    pos = np.column_stack([x, y, np.full(n, -0.1)]).astype(np.float32).astype(np.float64)
    w = src.model.factors(2.0 * ones, pos, np.tile([0.0, 0.0, 1.0], (n, 1)))
    prim = dict(particle_type=np.ones(n, np.int32), energy=2.0 * ones, x=pos[:, 0], y=pos[:, 1], z=pos[:, 2],
                ux=0 * ones, uy=0 * ones, uz=ones, weight=w.astype(np.float32))
    given = dict(kind="given", max_energy=2.0, primaries=prim)
    ora = oracle_lib.Oracle(grid.shape, grid.spacing, grid.origin, grid.density, grid.material,
                            grid.max_density_by_material(), given).run(n, batches, seed=5, threads=16)
    mask = ora["dose"] > 0.1 * ora["dose"].max()
    assert mask.sum() > 200
    p, chi2, mean, sd = batched_chi2_pvalue(hip.dose, hip.dose_sigma, batches, ora["dose"], ora["dose_sigma"],
                                            batches, mask)
    print(f"  chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    frac, count = north_star_gate(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"])
    assert count > 200 and frac < 0.06
    assert_z_consistent(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"], k=3.0, mask=mask)
    # the field is shaped: well inside the aperture the dose is near the open field's,
    # far outside it is transmission only
    assert hip.energy_emitted == pytest.approx(2.0 * float(np.sum(w)), rel=4e-3)
    assert hip.energy_emitted < 0.5 * n * 2.0
    b = hip.energy_emitted - hip.energy_deposited - hip.energy_escaped - hip.energy_unscored
    assert abs(b) < 1e-6 * hip.energy_emitted


def test_collimated_dij_matches_the_host_sampled_route(engine, grid_xs):
    from pyradmc_amd import BeamletGridSource
    from pyradmc_amd.collimation import CollimatedBeamletSource
    from tests.test_gpu_parity import _HostLattice

    grid, xs = grid_xs
    from pyradmc_amd.collimation import BeamFrame, BeamLimitingStack, JawPair

    lattice = BeamletGridSource(2.0, -0.1, (1.0, 3.0), (1.0, 3.0), 2, 2)
    # asymmetric jaws over the lattice (focal spot above x = y = 2)
    stack = BeamLimitingStack(BeamFrame((2.0, 2.0, -100.0)), (
        JawPair("u", 20.0, 27.0, -0.6, 0.8, 0, 19.3), JawPair("v", 28.0, 35.0, -0.7, 0.5, 0, 19.3)))
    a = engine.run_dij(CollimatedBeamletSource(lattice, stack, xs), 400_000, 10, seed=8, truncation=0.0)
    b = engine.run_dij(CollimatedBeamletSource(_HostLattice(), stack, xs), 400_000, 10, seed=9, truncation=0.0)
    for j in range(4):
        da, sa = a.column_dense(j).ravel(), a.sigma_dense(j).ravel()
        db, sb = b.column_dense(j).ravel(), b.sigma_dense(j).ravel()
        pooled = 0.5 * (da + db)
        mask = pooled > 0.1 * pooled.max()
        p, *_ = batched_chi2_pvalue(da, sa, 10, db, sb, 10, mask)
        assert p > ALPHA, (j, p)
    # the jaws shape each beamlet differently
    col_sums = [a.column_dense(j).sum() for j in range(4)]
    assert max(col_sums) > 1.5 * min(col_sums)
    assert a.energy_emitted == pytest.approx(b.energy_emitted, rel=0.01)


def test_open_stack_changes_nothing_and_mask_halves_the_emitted_energy(engine, grid_xs):
    from pyradmc_amd.collimation import CollimatedSource, TransmissionMaskSource

    grid, xs = grid_xs
    n = 600_000
    plain = engine.run(_field_source(), n, 6, seed=13)
    wide = engine.run(CollimatedSource(_field_source(), _stack(edge_scale=50.0), xs), n, 6, seed=13)
    np.testing.assert_array_equal(plain.dose, wide.dose)
    assert wide.energy_emitted == pytest.approx(n * 2.0, rel=1e-12)
    half = engine.run(TransmissionMaskSource(_field_source(), mask=np.full((3, 3), 0.5), plane_center=(4, 4, -10.0),
                                             width_u=20.0, width_v=20.0), n, 6, seed=13)
    assert half.energy_emitted == pytest.approx(n * 1.0, rel=1e-12)
    mask = plain.dose > 0.2 * plain.dose.max()
    ratio = half.dose[mask].sum() / plain.dose[mask].sum()
    assert ratio == pytest.approx(0.5, rel=0.02)
    again = engine.run(_field_source(), n, 6, seed=13)  # the model is cleared for plain sources
    np.testing.assert_array_equal(plain.dose, again.dose)
