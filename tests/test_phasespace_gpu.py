"""Phase-space sources on the GPU: a recorded field of photons with electron and positron
contamination, read from an IAEA file and transported by the HIP engine.

* Against the oracle's given-primaries transport of records drawn from the same file
  (another seed, so the two samples are independent): the batched chi-squared, the
  acceptance gate and the z-score test over the field, and the deposited fraction.
* The emitted book is the sum of ``weight * (E + 2 m_e c^2 [positrons])`` over exactly
  the records the sampling law draws, and the ledger closes.
* Bit reproducibility: the same seed gives the same dose under another chunk size, and
  an in-memory phase space holding the file's transportable records in file order gives
  the same dose as the file (both sample the same records).
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest

from tests.stats import assert_z_consistent, batched_chi2_pvalue, north_star_gate

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:.*latent variance:UserWarning")]

ALPHA = 0.01
ME = 0.51099895069
SHAPE, SPACING = (16, 16, 24), (0.5, 0.5, 0.5)


def _field_columns(n: int, seed: int = 1) -> dict:
    """A 6 MeV photon field over x, y in (1, 7) cm entering at z = -0.5 cm with a 50 mrad
    spread, 4 percent electrons and 1 percent positrons of 0.5-4 MeV, weights 0.5-1.5,
    and a few neutrons (which the source must drop)."""
    rng = np.random.default_rng(seed)
    k = rng.random(n)
    code = np.ones(n, np.int32)
    code[k < 0.05] = 2
    code[k < 0.01] = 3
    code[rng.choice(n, 12, replace=False)] = 4
    theta = np.abs(rng.normal(0.0, 0.05, n))
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    return dict(particle_type=code, energy=np.where(code == 1, 6.0, rng.uniform(0.5, 4.0, n)),
                x=rng.uniform(1.0, 7.0, n), y=rng.uniform(1.0, 7.0, n), z=np.full(n, -0.5),
                ux=np.sin(theta) * np.cos(phi), uy=np.sin(theta) * np.sin(phi), uz=np.cos(theta),
                weight=rng.uniform(0.5, 1.5, n))


@pytest.fixture(scope="module")
def field(tmp_path_factory):
    from pyradmc_amd.phasespace import write_iaea_phase_space

    path = write_iaea_phase_space(tmp_path_factory.mktemp("phsp") / "field", _field_columns(20_000))
    return path


@pytest.fixture(scope="module")
def engine():
    from pyradmc_amd import AnalyticCrossSections, HipEngine, VoxelGrid, _native

    _native.lib()
    grid = VoxelGrid.uniform_water(SHAPE, SPACING)
    eng = HipEngine(grid, AnalyticCrossSections(geometry_densities=grid.max_density_by_material()))
    yield eng
    eng.close()


def _source(path):
    from pyradmc_amd.phasespace import PhaseSpaceSource

    with pytest.warns(UserWarning, match="unsupported particle types"):
        return PhaseSpaceSource(path)


def _emitted(columns: dict) -> float:
    latent = np.where(columns["particle_type"] == 3, 2.0 * ME, 0.0)
    w = columns["weight"].astype(np.float64)
    return float(np.sum(w * (columns["energy"].astype(np.float64) + latent)))


def _balance(r) -> float:
    return abs(r.energy_emitted - r.energy_deposited - r.energy_escaped - r.energy_unscored) / r.energy_emitted


def test_phase_space_field_matches_oracle(engine, field, oracle_lib):
    print("phase-space field vs oracle", flush=True)
    src = _source(field)
    n, batches = 4_000_000, 40
    hip = engine.run(src, n, batches, seed=11)
    drawn = src.sample_batch(11, 0, n)
    assert set(np.unique(drawn["particle_type"]).tolist()) == {1, 2, 3}
    assert hip.energy_emitted == pytest.approx(_emitted(drawn), rel=1e-6)
    assert _balance(hip) < 1e-6

    grid = engine.grid
    given = dict(kind="given", max_energy=src.max_energy, primaries=src.sample_batch(12, 0, n))
    ora = oracle_lib.Oracle(grid.shape, grid.spacing, grid.origin, grid.density, grid.material,
                            grid.max_density_by_material(), given).run(n, batches, seed=12, threads=16)
    src.close()
    mask = ora["dose"] > 0.1 * ora["dose"].max()
    assert mask.sum() > 1000
    p, chi2, mean, sd = batched_chi2_pvalue(hip.dose, hip.dose_sigma, batches, ora["dose"],
                                            ora["dose_sigma"], batches, mask)
    print(f"  chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    frac, count = north_star_gate(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"])
    assert count > 1000 and frac < 0.06
    assert_z_consistent(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"], k=2.0, mask=mask)
    assert_z_consistent(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"], k=3.0, mask=mask)
    b = ora["books"]
    assert hip.energy_deposited / hip.energy_emitted == pytest.approx(b["deposited"] / b["emitted"], rel=0.005)


def test_phase_space_runs_are_reproducible_and_chunk_inert(engine, field):
    from pyradmc_amd import HipEngine
    from pyradmc_amd.phasespace import IAEAPhaseSpace, InMemoryPhaseSpaceSource

    src = _source(field)
    a = engine.run(src, 600_000, 6, seed=21)
    b = engine.run(src, 600_000, 6, seed=21)
    np.testing.assert_array_equal(a.dose, b.dose)
    small = HipEngine(engine.grid, engine.cross_sections, chunk_size=50_000)
    c = small.run(src, 600_000, 6, seed=21)
    small.close()
    src.close()
    np.testing.assert_array_equal(a.dose, c.dose)
    assert a.energy_emitted == c.energy_emitted

    with IAEAPhaseSpace(field) as ps:
        cols = ps.columns()
    keep = np.isin(cols["particle_type"], (1, 2, 3))
    mem = InMemoryPhaseSpaceSource(*(cols[k][keep] for k in ("particle_type", "energy", "x", "y", "z",
                                                               "ux", "uy", "uz", "weight")))
    d = engine.run(mem, 600_000, 6, seed=21)
    np.testing.assert_array_equal(a.dose, d.dose)
    np.testing.assert_array_equal(a.dose_sigma, d.dose_sigma)


def test_resident_population_matches_host_sampling(engine, field):
    # The resident route draws its records in-kernel from its own stream: it must
    # estimate the same dose as the host-sampled route (independent samples, so the
    # batched chi-squared and the gate apply), book the energy of the records it drew,
    # and close the ledger.
    from pyradmc_amd.phasespace import DevicePhaseSpace

    print("resident population vs host sampling", flush=True)
    src = _source(field)
    dev = DevicePhaseSpace(src)
    n, batches = 4_000_000, 40
    a = engine.run(dev, n, batches, seed=31)
    b = engine.run(src, n, batches, seed=32)
    mask = b.dose > 0.1 * b.dose.max()
    assert mask.sum() > 1000
    p, chi2, mean, sd = batched_chi2_pvalue(a.dose, a.dose_sigma, batches, b.dose, b.dose_sigma, batches, mask)
    print(f"  chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    frac, count = north_star_gate(a.dose, a.dose_sigma, b.dose, b.dose_sigma)
    assert count > 1000 and frac < 0.06
    assert_z_consistent(a.dose, a.dose_sigma, b.dose, b.dose_sigma, k=3.0, mask=mask)
    assert _balance(a) < 1e-6
    # emitted energy per history: the population mean of w (E + latent), within 5 sigma
    pop = dev.to_phase_space().columns()
    per = pop["weight"] * (pop["energy"] + np.where(pop["particle_type"] == 3, 2.0 * ME, 0.0))
    assert abs(a.energy_emitted / n - per.mean()) < 5.0 * per.std() / np.sqrt(n)
    src.close()


def test_resident_runs_are_reproducible_and_uploaded_once(engine, field):
    from pyradmc_amd import HipEngine
    from pyradmc_amd.phasespace import DevicePhaseSpace

    src = _source(field)
    dev = DevicePhaseSpace(src, block=6_000)  # several upload calls
    a = engine.run(dev, 600_000, 6, seed=41)
    assert engine._population_token == dev.token
    mem = DevicePhaseSpace(dev.to_phase_space())  # the same rows, another population
    b = engine.run(mem, 600_000, 6, seed=41)
    assert engine._population_token == mem.token
    np.testing.assert_array_equal(a.dose, b.dose)
    c = engine.run(dev, 600_000, 6, seed=41)  # uploaded again, same draws
    np.testing.assert_array_equal(a.dose, c.dose)
    small = HipEngine(engine.grid, engine.cross_sections, chunk_size=50_000)
    d = small.run(dev, 600_000, 6, seed=41)
    small.close()
    src.close()
    np.testing.assert_array_equal(a.dose, d.dose)
    assert a.energy_emitted == d.energy_emitted
    e = engine.run(dev, 600_000, 6, seed=42)
    assert not np.array_equal(a.dose, e.dose)


def test_resident_population_argument_errors(engine):
    from pyradmc_amd import _native as nat

    n = 4
    kind = np.zeros(n, np.int32)
    cols = [np.ones(n, np.float32) for _ in range(8)]
    ptrs = [nat.ptr(c, nat.ctypes.c_float) for c in cols]
    lib = nat.lib()
    ctx = engine._ctx
    with pytest.raises(ValueError, match="in order"):
        ctx.call("hmc_set_population", n, 2, 2, nat.ptr(kind, nat.c_int32), *ptrs)
    with pytest.raises(ValueError, match="out of range"):
        ctx.call("hmc_set_population", n, 0, n + 1, nat.ptr(kind, nat.c_int32), *ptrs)
    bad = np.array([0, 1, 3, 0], np.int32)
    with pytest.raises(ValueError, match="bad particle kind"):
        ctx.call("hmc_set_population", n, 0, n, nat.ptr(bad, nat.c_int32), *ptrs)
    # a partly written population cannot back a resident source
    ctx.call("hmc_set_population", n, 0, 2, nat.ptr(kind, nat.c_int32), *ptrs)
    engine._population_token = None
    d = nat.SourceDesc()
    d.kind = nat.SRC_RESIDENT
    with pytest.raises(ValueError, match="complete population"):
        ctx.call("hmc_set_source", nat.ctypes.byref(d))
    assert lib is not None
