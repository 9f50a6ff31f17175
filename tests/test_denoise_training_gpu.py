"""The denoiser at its full size, trained on pairs the engine makes, and what the
committed network gains over the Monte Carlo it starts from.

* Pairs come from the engine on seeded CT-like phantoms (128x128x96 at 2.5 mm, 5x5 mm
  beamlets from 100 cm SAD): each column at 1e4 histories (input) and 1e6 (target),
  in its 64x64 beam-aligned window.
* The network is the (16, 32, 64, 128) encoder-decoder on 64^3 crops. Adam steps on a
  fixed batch lower the sigma-weighted loss, and the training loop (random crops,
  flips, cosine schedule) runs.
* Every layer's forward, data gradient, weight and bias gradients at that width, from
  an engine crop and the loss's own gradient, agree with the float64 oracle
  (oracle/conv_ref.py) fed the layer's bf16 input and output gradient; one Adam step
  agrees with the update written out in float64; the whole bf16 network follows the
  oracle's float64 network.
* The committed model (pyradmc_amd/models/denoiser_ct.npz), on columns of a phantom it
  never saw, has a dose-weighted RMS error against the 1e6-history target well below
  that of its 1e4-history input: the gain over plain Monte Carlo at the same histories.
* Plain Monte Carlo judges the denoised build: its audit accepts the trained network on
  an unseen phantom and rejects the same network with a bias added to its prediction,
  whereupon the result is the plain Dij bit for bit.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import conv_ref as cr

pytestmark = pytest.mark.gpu

WIDTHS = (16, 32, 64, 128)
PAIR_SEED = 20261203      # phantoms used by no training run of the committed model
HELD_OUT_SEED = 20261207
CROP = 64
MODEL = Path(__file__).resolve().parent.parent / "pyradmc_amd" / "models" / "denoiser_ct.npz"


@pytest.fixture(scope="module")
def dn():
    import torch

    from pyradmc_amd import denoise

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    denoise._lib()
    return denoise


@pytest.fixture(scope="module")
def pairs(dn):
    # 2 phantoms x 2 beamlets: 4 columns at 1e4 and 1e6 histories
    out = dn.generate_pairs(2, 2, low=10_000, high=1_000_000, seed=PAIR_SEED, n_batches=10)
    assert len(out) == 4
    return out


def _np(t) -> np.ndarray:
    import torch

    return t.to(dtype=torch.float64).cpu().numpy()


def _batch(dn, pairs, idx, z0: int):
    """Fixed 64^3 crops (depth from z0) of the chosen pairs, packed as the training loop
    packs them: bf16 input and its channel-major copy, target, normalised weight."""
    import torch

    noisy, dens, target, weight = dn._stack_device([pairs[i] for i in idx], "cuda:0", "sigma")
    a, d, t, w = (v[:, :CROP, :CROP, z0:z0 + CROP].permute(0, 3, 1, 2).contiguous()
                  for v in (noisy, dens, target, weight))
    w = w / w.mean()
    x, x_cm = dn._pack_device(a, d)
    torch.cuda.synchronize()
    return x, x_cm, t, w


def _loss(dn, model, x, x_cm, t, w, keep: bool):
    out, cache = model.forward(x, x_cm, keep=keep)
    partial, dout, dout_cm = model.loss_and_grad(out, x, t, w, 1.0 / t.numel())
    return float(partial.double().sum().item()) / t.numel(), out, cache, dout, dout_cm


def test_engine_pairs_are_windows_of_two_statistics(dn, pairs):
    for p in pairs:
        assert p.noisy.shape == (64, 64, 96) and p.target.shape == (64, 64, 96)
        assert p.scale > 0 and np.isfinite(p.noisy).all() and np.isfinite(p.target).all()
        assert (p.target_sigma >= 0).all()
        # the input is 100x noisier: its error against the target dominates the target's sigma
        err = dn.dose_weighted_rms(p.noisy, p.target)
        assert 0.02 < err < 0.5, err
        assert 0.05 <= p.density.min() and p.density.max() <= 1.9


def test_adam_steps_lower_the_loss_at_full_width(dn, pairs):
    import torch

    model = dn.Conv3dDenoiser(widths=WIDTHS, seed=1)
    assert 1.0e6 < model.n_parameters < 2.0e6
    x, x_cm, t, w = _batch(dn, pairs, [0, 1, 2, 3], z0=8)
    # 16 Adam steps at 1e-4 on one fixed batch of four crops; the loss starts at the
    # noisy input's own error (the residual head starts near zero) and falls to about a
    # third of it on the MI355X.
    losses = []
    for _ in range(16):
        loss, out, cache, dout, dout_cm = _loss(dn, model, x, x_cm, t, w, keep=True)
        losses.append(loss)
        model.backward(cache, dout, dout_cm)
        model.adam_step(1e-4)
    losses.append(_loss(dn, model, x, x_cm, t, w, keep=False)[0])
    torch.cuda.synchronize()
    print("loss per step:", " ".join(f"{v:.5g}" for v in losses), flush=True)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.6 * losses[0], losses
    # the training loop itself: random crops and flips, cosine decay, two epochs
    hist = dn.train(model, pairs, epochs=2, batch=2, crop=CROP, lr=1e-4, lr_final=1e-5, seed=3)
    assert len(hist) == 2 and np.isfinite(hist).all()
    assert model.step_count == 16 + 2 * 2


def test_every_layer_matches_the_oracle_at_full_width(dn, pairs):
    # One engine crop through the (16, 32, 64, 128) network and back from the loss's
    # own gradient; each kernel is checked against the oracle fed that layer's actual
    # bf16 input and output gradient (tolerances as in the kernel tests: 2^-8 relative
    # for bf16 outputs, 1e-3 for float32 gradient sums).
    import torch

    model = dn.Conv3dDenoiser(widths=WIDTHS, seed=2)
    model.keep_gradients = True
    x, x_cm, t, w = _batch(dn, pairs, [1], z0=16)
    loss, out, cache, dout, dout_cm = _loss(dn, model, x, x_cm, t, w, keep=True)
    model.backward(cache, dout, dout_cm)
    torch.cuda.synchronize()

    rtol = 2.0 ** -8
    for name, lay in model.layers.items():
        inp, _, y, _, _ = cache[name]
        dz = _np(cache[name + "_dz"])
        xin = _np(inp)
        wq = _np(lay.wf)
        z = cr.conv_fwd(xin, wq, _np(lay.b))
        ref = z if name == "head" else cr.leaky(z, dn.SLOPE)
        np.testing.assert_allclose(_np(y), ref, rtol=rtol, atol=rtol * 0.05 * np.abs(ref).max(), err_msg=name)
        ref_w = cr.conv_wgrad(xin, dz)
        np.testing.assert_allclose(_np(lay.gw), ref_w, rtol=1e-3, atol=1e-4 * np.abs(ref_w).max(), err_msg=name)
        ref_b = dz.sum(axis=(0, 1, 2, 3))
        np.testing.assert_allclose(_np(lay.gb), ref_b, rtol=1e-3, atol=1e-4 * np.abs(ref_b).max() + 1e-12,
                                   err_msg=name)
        if name != "e0a":
            ref_dx = cr.conv_dgrad(dz, wq)
            np.testing.assert_allclose(_np(cache[name + "_dx"]), ref_dx, rtol=rtol,
                                       atol=rtol * 0.05 * np.abs(ref_dx).max(), err_msg=name)
        assert np.abs(ref_w).max() > 0, name   # the loss reaches every layer

    # one Adam step of a layer against the update written out in float64
    lay = model.layers["e2b"]
    w0, g, m0, v0 = (_np(a) for a in (lay.w, lay.gw, lay.mw, lay.vw))
    step = model.step_count + 1
    model.adam_step(1e-3)
    torch.cuda.synchronize()
    # the kernel's constants are float32: 0.999f is 0.99900001..., so (1 - b2) differs
    # from 0.001 by 1.3e-5 relative
    b1, b2, eps = float(np.float32(0.9)), float(np.float32(0.999)), float(np.float32(1e-8))
    m1 = b1 * m0 + (1 - b1) * g
    v1 = b2 * v0 + (1 - b2) * g * g
    w1 = w0 - 1e-3 * (m1 / (1 - b1 ** step)) / (np.sqrt(v1 / (1 - b2 ** step)) + eps)
    np.testing.assert_allclose(_np(lay.mw), m1, rtol=1e-5, atol=1e-6 * np.abs(m1).max())
    np.testing.assert_allclose(_np(lay.vw), v1, rtol=1e-5, atol=1e-6 * np.abs(v1).max())
    np.testing.assert_allclose(_np(lay.w), w1, rtol=1e-6, atol=1e-6 * np.abs(w0).max())
    assert not np.array_equal(_np(lay.w), w0)


def test_whole_network_follows_the_float64_oracle_at_full_width(dn, pairs):
    # End to end: the bf16 network's head output and weight gradients against the
    # oracle's float64 network on the same weights, input and output gradient, within
    # what bf16 activations through 15 layers allow.
    import torch

    model = dn.Conv3dDenoiser(widths=WIDTHS, seed=4)
    x, x_cm, t, w = _batch(dn, pairs, [2], z0=24)
    loss, out, cache, dout, dout_cm = _loss(dn, model, x, x_cm, t, w, keep=True)
    model.backward(cache, dout, dout_cm)
    torch.cuda.synchronize()
    params = {name: (_np(lay.wf), _np(lay.b)) for name, lay in model.layers.items()}
    ref_out, ref_cache = cr.network_forward(params, _np(x), dn.SLOPE)
    rel = np.linalg.norm(_np(out) - ref_out) / np.linalg.norm(ref_out)
    assert rel < 0.05, rel
    # the loss gradient the kernel formed (bf16) drives the float64 backward too
    grads = cr.network_backward(params, ref_cache, _np(dout), dn.SLOPE)
    for name, lay in model.layers.items():
        got, ref = _np(lay.gw), grads[name][0]
        nr = np.linalg.norm(ref)
        cos = float(np.sum(got * ref) / (np.linalg.norm(got) * nr))
        err = float(np.linalg.norm(got - ref) / nr)
        assert cos > 0.99 and err < 0.15, (name, cos, err)


def test_committed_model_beats_its_noisy_input_on_a_held_out_phantom(dn):
    # Fresh engine columns on a phantom seed no training run used: 8 beamlets at 1e4
    # histories (the network's input, i.e. plain Monte Carlo without the network) and
    # 1e6 (the reference). The trained network must cut the dose-weighted RMS error of
    # every column, and the median at least threefold.
    model = dn.load(MODEL)
    assert tuple(model.widths) == WIDTHS
    held = dn.generate_pairs(1, 8, low=10_000, high=1_000_000, seed=HELD_OUT_SEED, n_batches=10)
    ev = dn.evaluate(model, held)
    noisy, den = ev["noisy"], ev["denoised"]
    print(f"held-out dose-weighted RMS: noisy median {np.median(noisy):.4f}, "
          f"denoised median {np.median(den):.4f} (per column {np.round(den, 4).tolist()})", flush=True)
    assert (den < noisy).all(), (noisy, den)
    assert np.median(den) < np.median(noisy) / 3.0
    assert np.median(den) < 0.03


def _ct_plan(n_beamlets: int = 10):
    from pyradmc_amd import HipEngine, SpectralBeamletSource, workloads
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    grid = workloads.ct_like_grid(seed=HELD_OUT_SEED + 1)
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    full = workloads.ct_like_beamlets()
    pick = np.linspace(0, full.n_beamlets - 1, n_beamlets).round().astype(int).tolist()
    src = SpectralBeamletSource(full.spectrum, full.focal_point, [full.centers[j] for j in pick],
                                full.width_u, full.width_v)
    return HipEngine(grid, xs), src


def test_plain_monte_carlo_judge_accepts_the_trained_network(dn):
    # On a phantom no training run used, the audit (plain Monte Carlo on 3 of 10 columns
    # at 250,000 histories) estimates the trained network's error below the 2% budget:
    # the audited columns come back as plain Monte Carlo, the others denoised.
    eng, src = _ct_plan()
    model = dn.load(MODEL)
    builder = dn.DenoisedDijBuilder(eng, model, low_histories=20_000, tolerance=0.02,
                                    fallback_histories=250_000, n_batches=10, audit_fraction=0.3)
    dij, report = builder.build(src, seed=5, truncation=1e-4)
    eng.close()
    print(f"audit error {report.audit_error:.4f}: {np.round(report.estimated_error[report.audited], 4).tolist()}")
    assert report.audited.sum() == 3 and np.array_equal(report.fallback, report.audited)
    assert 0.0 < report.audit_error < 0.02
    assert dij.indptr[-1] > 0 and np.all(np.diff(dij.indptr) > 0)


def test_plain_monte_carlo_judge_rejects_a_biased_network(dn):
    # The same network with a bias added to its prediction (+5% of each column's scale
    # everywhere): two denoised runs of it would agree with each other, but the plain
    # Monte Carlo audit sees the bias, and every column is re-simulated; the result is
    # the plain Dij bit for bit.
    eng, src = _ct_plan()
    model = dn.load(MODEL)
    head = model.layers["head"]
    head.b[0] += 0.05
    model._prepare(head)
    builder = dn.DenoisedDijBuilder(eng, model, low_histories=20_000, tolerance=0.02,
                                    fallback_histories=250_000, n_batches=10, audit_fraction=0.3)
    dij, report = builder.build(src, seed=5, truncation=1e-4)
    plain = eng.run_dij(src, 250_000, 10, seed=7, truncation=1e-4)
    eng.close()
    print(f"audit error {report.audit_error:.4f}")
    assert report.audit_error > 0.02 and report.fallback.all()
    np.testing.assert_array_equal(dij.indptr, plain.indptr)
    np.testing.assert_array_equal(dij.indices, plain.indices)
    np.testing.assert_array_equal(dij.dose, plain.dose)
    np.testing.assert_array_equal(dij.sigma, plain.sigma)
