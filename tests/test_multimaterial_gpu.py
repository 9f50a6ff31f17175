"""Material-resolved transport on the GPU: the two-material slab against the oracle and
the reference.

The reference's synthetic two-material table (water, and 'air' at half the water
values, twice the CSDA range) in its slab (material 1 at z voxels 5-8 of 8x8x16
voxels of 2x2x1 cm, unit density), under a 2 MeV parallel beam: per-material table
rows, the Woodcock majorant over both materials, the GS window over both screening
rows, and dose to water weighted by the water/medium stopping ratio in the slab.

* HIP against the oracle at high statistics (40 batches a side) passes the acceptance
  gate in each mode: KERMA, electrons scored as dose to medium, and as dose to water.
* HIP replicates drawn with the reference sample's batch structure place the
  reference's ``ReferenceEngine`` doses (tests/golden/ref_two_material_slab.npz) within
  4 replicate standard deviations on every summary statistic.
* Same seed, dose to water against dose to medium: in the slab every deposit is
  weighted by exactly 2 (the stopping ratio of the two rows), in water by exactly 1,
  and the physical energy books are identical.
* The fast round build (dose to medium, no spreading) and the general build agree bit
  for bit on the slab.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests import problems as P
from tests.stats import (assert_within_replicates, assert_z_consistent, batched_chi2_pvalue,
                         comparison_statistics, north_star_gate)
from tests.test_multimaterial import MODES, oracle_slab_run

pytestmark = pytest.mark.gpu

ALPHA = 0.01


@pytest.fixture(scope="module")
def engine():
    from pyradmc_amd import HipEngine, _native

    _native.lib()
    grid, xs = P.two_material_slab()
    eng = HipEngine(grid, xs)
    yield eng
    eng.close()


def _run(engine, tag: str, n: int, batches: int, seed: int, **kw):
    electrons, mode = MODES[tag]
    return engine.run(P.two_material_beam(), n, batches, seed=seed, transport_electrons=electrons,
                      scoring_mode=mode, **kw)


def _balance(r) -> float:
    return abs(r.energy_emitted - r.energy_deposited - r.energy_escaped - r.energy_unscored) / r.energy_emitted


@pytest.mark.parametrize("tag", ["kerma", "medium", "water"])
def test_slab_matches_oracle(engine, oracle_lib, tag):
    print(f"two-material slab vs oracle: {tag}", flush=True)
    hip = _run(engine, tag, 4_000_000, 40, seed=41)
    ora = oracle_slab_run(tag, 2_000_000, 40, seed=42, threads=16)
    mask = ora["dose"] > 0.1 * ora["dose"].max()
    assert mask.sum() > 300
    p, chi2, mean, sd = batched_chi2_pvalue(hip.dose, hip.dose_sigma, 40, ora["dose"], ora["dose_sigma"], 40, mask)
    print(f"  chi2 {chi2:.1f} vs {mean:.1f} +- {sd:.1f}, p = {p:.3g}", flush=True)
    assert p > ALPHA
    frac, n = north_star_gate(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"])
    assert n > 300 and frac < 0.06
    assert_z_consistent(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"], k=2.0, mask=mask)
    assert_z_consistent(hip.dose, hip.dose_sigma, ora["dose"], ora["dose_sigma"], k=3.0, mask=mask)
    books = ora["books"]
    assert hip.energy_deposited / hip.energy_emitted == pytest.approx(books["deposited"] / books["emitted"],
                                                                      rel=0.005)
    assert _balance(hip) < 1e-6
    # the material row reaches the kernel: the slab's KERMA (or dose) dips
    pdd = hip.dose.sum(axis=(0, 1))
    assert pdd[P.TWO_MATERIAL_SLAB].mean() < 0.75 * pdd[3:5].mean() or tag == "water"


@pytest.mark.parametrize("tag", ["kerma", "medium", "water"])
def test_slab_matches_reference_engine(engine, golden_dir, tag):
    g = np.load(golden_dir / "ref_two_material_slab.npz")
    n, nb = int(g[f"{tag}_run"][0]), int(g[f"{tag}_run"][1])
    truth = _run(engine, tag, 8_000_000, 40, seed=1001)
    independent = _run(engine, tag, 2_400_000, 12, seed=2002).dose
    mask = independent > 0.1 * independent.max()

    def summary(dose, sigma):
        return comparison_statistics(truth.dose, truth.dose_sigma, 40, dose, sigma, nb, mask)

    reps = [summary(r.dose, r.dose_sigma) for r in (_run(engine, tag, n, nb, seed=3000 + k) for k in range(16))]
    ref = summary(g[f"{tag}_dose"], g[f"{tag}_sigma"])
    print(tag, {k: round(v, 4) for k, v in ref.items()}, flush=True)
    assert_within_replicates(ref, reps)


def test_dose_to_water_weighs_each_material_by_its_stopping_ratio(engine):
    medium = _run(engine, "medium", 400_000, 8, seed=7)
    water = _run(engine, "water", 400_000, 8, seed=7)
    # same histories: the weight draws no random number
    assert water.energy_deposited == medium.energy_deposited
    assert water.energy_escaped == medium.energy_escaped
    assert water.energy_unscored == medium.energy_unscored
    # Inner slab voxels (z 6-7): every deposit there comes from a sub-step or photon
    # interaction inside the slab (a half-step's deposit lands within half a sub-step,
    # at most half a voxel, of where the sub-step started). Each deposit's quanta round
    # once: 2 E rounded against 2 x (E rounded), at most one quantum (1e-9 MeV) apart.
    inner = np.zeros(P.TWO_MATERIAL_SHAPE, dtype=bool)
    inner[:, :, 6:8] = True
    hit = medium.dose > 0
    assert (inner & hit).sum() > 100
    ratio = water.dose[inner & hit] / medium.dose[inner & hit]
    np.testing.assert_allclose(ratio, 2.0, rtol=1e-6)
    far = np.zeros_like(inner)
    far[:, :, 11:] = True  # water voxels two or more voxels past the slab
    np.testing.assert_array_equal(water.dose[far], medium.dose[far])


def test_fast_and_general_builds_agree_on_two_materials(engine):
    # Dose to medium without spreading runs the fast build; a deposit resolution longer
    # than any sub-step runs the general one with every half-step in one piece.
    fast = _run(engine, "medium", 400_000, 8, seed=13)
    general = _run(engine, "medium", 400_000, 8, seed=13, deposit_resolution_cm=1.0e3)
    assert fast.dose.max() > 0.0
    np.testing.assert_array_equal(general.dose, fast.dose)
    np.testing.assert_array_equal(general.dose_sigma, fast.dose_sigma)
    assert general.energy_deposited == fast.energy_deposited
