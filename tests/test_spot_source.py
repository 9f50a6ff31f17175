"""Gaussian focal-spot sources on the host (CPU): the sampling contract of
``sample_batch`` (pyradmc's tests/unit/test_gaussian_spot_source.py cases) and the
oracle's spot generator pinned against its zero-sigma limit."""

from __future__ import annotations

import numpy as np
import pytest

from pyradmc_amd import GaussianSpotBeamletSource, GaussianSpotBeamSource, Spectrum, bremsstrahlung_6mv_like

FOCAL = (0.0, 0.0, -100.0)
PLANE_Z = -60.0


def _source(sigma=(0.2, 0.3), spectrum=None):
    return GaussianSpotBeamSource(spectrum or bremsstrahlung_6mv_like(), FOCAL, (0.0, 0.0, PLANE_Z), 4.0, 6.0,
                                  *sigma)


def test_spectrum_top_and_validation():
    assert _source().max_energy == pytest.approx(6.0)
    with pytest.raises(ValueError, match="non-negative"):
        _source(sigma=(-0.1, 0.2))


def test_primaries_start_on_the_plane_rectangle():
    b = _source().sample_batch(3, 0, 20_000)
    assert np.all(b["z"] == np.float32(PLANE_Z))
    assert np.all(np.abs(b["x"]) <= 2.0) and np.all(np.abs(b["y"]) <= 3.0)
    assert b["x"].std() == pytest.approx(4.0 / np.sqrt(12), rel=0.03)
    norm = b["ux"].astype(np.float64) ** 2 + b["uy"] ** 2 + b["uz"] ** 2
    np.testing.assert_allclose(norm, 1.0, atol=1e-6)
    assert np.all(b["particle_type"] == 1) and np.all(b["weight"] == 1.0)


def test_zero_sigma_gives_the_fan_lines():
    b = _source(sigma=(0.0, 0.0)).sample_batch(4, 0, 2000)
    start = np.column_stack([b["x"], b["y"], b["z"]]).astype(np.float64)
    line = start - np.asarray(FOCAL)
    line /= np.linalg.norm(line, axis=1, keepdims=True)
    d = np.column_stack([b["ux"], b["uy"], b["uz"]]).astype(np.float64)
    np.testing.assert_allclose(d, line, atol=2e-7)


def test_back_projected_spot_is_gaussian_with_the_stated_sigmas():
    # Carry each line back to the focal plane: the crossing is the sampled spot point.
    n = 200_000
    b = _source(sigma=(0.2, 0.5)).sample_batch(5, 0, n)
    start = np.column_stack([b["x"], b["y"], b["z"]]).astype(np.float64)
    d = np.column_stack([b["ux"], b["uy"], b["uz"]]).astype(np.float64)
    t = (FOCAL[2] - start[:, 2]) / d[:, 2]
    spot = start + t[:, None] * d
    for axis, sigma in ((0, 0.2), (1, 0.5)):
        s = spot[:, axis]
        assert abs(s.mean()) < 5 * sigma / np.sqrt(n)
        assert s.std() == pytest.approx(sigma, rel=0.01)
    # independent axes, Gaussian tails (kurtosis 3)
    assert abs(np.corrcoef(spot[:, 0], spot[:, 1])[0, 1]) < 0.01
    z = spot[:, 0] / 0.2
    assert np.mean(z ** 4) == pytest.approx(3.0, rel=0.05)


def test_energies_follow_the_spectrum():
    sp = Spectrum([0.5, 1.0, 2.0, 4.0], [1.0, 2.0, 1.0])
    n = 100_000
    e = _source(spectrum=sp).sample_batch(6, 0, n)["energy"]
    counts = np.histogram(e, bins=sp.edges)[0]
    expect = n * sp.bin_probabilities
    chi2 = np.sum((counts - expect) ** 2 / expect)
    assert chi2 < 16.0  # 2 degrees of freedom, p ~ 3e-4
    assert e.min() >= 0.5 and e.max() <= 4.0


def test_chunk_invariance_and_six_draws_per_history():
    src = _source()
    whole = src.sample_batch(7, 100, 500)
    parts = [src.sample_batch(7, 100 + a, b - a) for a, b in ((0, 1), (1, 250), (250, 500))]
    for k in whole:
        np.testing.assert_array_equal(whole[k], np.concatenate([p[k] for p in parts]))
    # history h uses draws 6h..6h+5: its offset in the rectangle is draws 6h+2, 6h+3
    u = np.random.Generator(np.random.PCG64(7)).random(6 * 600).reshape(600, 6)[100:]
    np.testing.assert_allclose(whole["x"], (u[:, 2] - 0.5) * 4.0, atol=1e-6)
    np.testing.assert_allclose(whole["y"], (u[:, 3] - 0.5) * 6.0, atol=1e-6)


def test_beamlets_replay_the_same_draws():
    centers = [(-2.0, 0.0, PLANE_Z), (0.0, 0.0, PLANE_Z), (3.0, 1.0, PLANE_Z)]
    src = GaussianSpotBeamletSource(bremsstrahlung_6mv_like(), FOCAL, centers, 1.0, 1.0, 0.2, 0.2)
    assert src.n_beamlets == 3
    cols = [src.sample_beamlet_batch(8, 40, 300, j) for j in range(3)]
    for j in range(3):
        np.testing.assert_allclose(cols[j]["x"] - cols[0]["x"], centers[j][0] - centers[0][0], atol=1e-5)
        np.testing.assert_array_equal(cols[j]["energy"], cols[0]["energy"])
    with pytest.raises(IndexError):
        src.sample_beamlet_batch(8, 0, 10, 3)


def test_oracle_spot_fan_reduces_to_the_fan_at_zero_sigma(oracle_lib):
    # The oracle's spot generator at zero sigma transports the fan's lines (started on
    # the aperture plane instead of the focal spot, both outside the phantom): the two
    # oracle sources agree statistically; a wide spot blurs the field edge.
    from pyradmc_amd import SpectralBeamSource, VoxelGrid
    from tests.stats import batched_chi2_pvalue

    grid = VoxelGrid.uniform_water((12, 12, 8), (0.5, 0.5, 0.5))
    sp = Spectrum([1.0, 2.0], [1.0])
    focal, center = (3.0, 3.0, -100.0), (3.0, 3.0, -50.0)
    fan = SpectralBeamSource(sp, focal, center, 1.0, 1.0)
    spot0 = GaussianSpotBeamSource(sp, focal, center, 1.0, 1.0, 0.0, 0.0)
    wide = GaussianSpotBeamSource(sp, focal, center, 1.0, 1.0, 1.0, 1.0)
    runs = [oracle_lib.for_grid(grid, s).run(1_000_000, 10, seed=s_, threads=8)
            for s, s_ in ((fan, 10), (spot0, 100), (wide, 3))]
    a, b, c = runs
    mask = a["dose"] > 0.1 * a["dose"].max()
    p, *_ = batched_chi2_pvalue(a["dose"], a["dose_sigma"], 10, b["dose"], b["dose_sigma"], 10, mask)
    assert p > 0.01
    assert c["dose"].max() < 0.9 * a["dose"].max()  # the wide spot spreads the field
    for r in runs:
        bk = r["books"]
        assert abs(bk["emitted"] - bk["deposited"] - bk["escaped"] - bk["unscored"]) < 1e-9 * bk["emitted"]
