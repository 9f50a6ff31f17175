"""The denoiser's HIP kernels against the float64 oracle (oracle/conv_ref.py).

Inputs and weights are rounded to bfloat16 first, so the oracle sees exactly the
values the kernels multiply; the kernels accumulate in float32 and round their
outputs to bfloat16, hence the tolerances: 2^-8 relative for a bf16 output, 1e-4 for
float32 weight-gradient sums, and per-layer gradient agreement through the whole
bf16 network (cosine > 0.995, relative error < 5 percent).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import conv_ref as cr

pytestmark = pytest.mark.gpu

BF16_RTOL = 2.0 ** -8


@pytest.fixture(scope="module")
def dn():
    import torch

    from pyradmc_amd import denoise

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    denoise._lib()
    return denoise


def _bf(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)


def _rounded(a) -> np.ndarray:
    return _bf(a).to(dtype=__import__("torch").float64).numpy()


def _dev(t):
    return t.to("cuda:0")


def _np(t) -> np.ndarray:
    import torch

    return t.to(dtype=torch.float64).cpu().numpy()


def _cm(t, C):
    return t.reshape(-1, C).t().contiguous()


def _close_bf16(got, ref):
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=BF16_RTOL, atol=BF16_RTOL * 0.05 * scale)


# Shapes with w = 64 (h % 4 == 0) run the LDS-tiled kernel of the full-resolution
# layers, forward and data gradient; w = 32 (h % 8 == 0, cout 32) its half-resolution
# form in forward.
@pytest.mark.parametrize("shape,cin,cout", [((1, 8, 8, 16), 16, 16), ((1, 4, 8, 8), 32, 48),
                                            ((2, 4, 4, 16), 16, 32), ((1, 4, 4, 8), 32, 64),
                                            ((1, 4, 4, 8), 16, 128), ((1, 3, 4, 64), 16, 16),
                                            ((2, 2, 8, 64), 16, 32), ((1, 3, 4, 64), 32, 16),
                                            ((1, 3, 8, 32), 16, 32), ((2, 2, 8, 32), 32, 32),
                                            ((1, 2, 16, 32), 64, 32)])
def test_conv_forward_data_and_weight_gradients(dn, shape, cin, cout):
    import torch

    rng = np.random.default_rng(cin * 1000 + cout)
    n, d, h, w = shape
    P = n * d * h * w
    x = _rounded(rng.standard_normal(shape + (cin,)))
    master = rng.standard_normal((27, cout, cin)).astype(np.float32) * 0.2
    wq = _rounded(master)
    bias = rng.standard_normal(cout).astype(np.float32)
    dz = _rounded(rng.standard_normal(shape + (cout,)))

    xd = _dev(_bf(x))
    wf = torch.empty((27, cout, cin), dtype=torch.bfloat16, device="cuda:0")
    wd = torch.empty((27, cin, cout), dtype=torch.bfloat16, device="cuda:0")
    md = _dev(torch.from_numpy(master))
    dn._call("hdn_prepare_weights", md.data_ptr(), wf.data_ptr(), wd.data_ptr(), cin, cout, None)
    np.testing.assert_array_equal(_np(wf), wq)
    bd = _dev(torch.from_numpy(bias))
    y = torch.empty(shape + (cout,), dtype=torch.bfloat16, device="cuda:0")
    y_cm = torch.empty((cout, P), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_conv3d_fwd", xd.data_ptr(), wf.data_ptr(), bd.data_ptr(), y.data_ptr(), y_cm.data_ptr(),
             n, d, h, w, cin, cout, 0.1, None)
    torch.cuda.synchronize()
    ref = cr.leaky(cr.conv_fwd(x, wq, bias.astype(np.float64)), 0.1)
    _close_bf16(_np(y), ref)
    np.testing.assert_array_equal(_np(y_cm), _np(y).reshape(-1, cout).T)

    # data gradient: the forward kernel on the flipped, transposed copy
    dzd = _dev(_bf(dz))
    dx = torch.empty(shape + (cin,), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_conv3d_fwd", dzd.data_ptr(), wd.data_ptr(), None, dx.data_ptr(), None,
             n, d, h, w, cout, cin, 1.0, None)
    torch.cuda.synchronize()
    _close_bf16(_np(dx), cr.conv_dgrad(dz, wq))

    # weight gradient from channel-major copies, split partials reduced in order
    if w % 8 == 0:
        splits = 3 if P >= 256 else 2
        part = torch.empty((splits, 27, cout, cin), dtype=torch.float32, device="cuda:0")
        gw = torch.empty((27, cout, cin), dtype=torch.float32, device="cuda:0")
        x_cm, dz_cm = _cm(xd, cin), _cm(dzd, cout)  # held: a freed temporary could be reused
        dn._call("hdn_conv3d_wgrad", x_cm.data_ptr(), dz_cm.data_ptr(), part.data_ptr(),
                 splits, n, d, h, w, cin, cout, None)
        dn._call("hdn_reduce_splits", part.data_ptr(), gw.data_ptr(), gw.numel(), splits, None)
        torch.cuda.synchronize()
        ref_w = cr.conv_wgrad(x, dz)
        np.testing.assert_allclose(_np(gw), ref_w, rtol=1e-4, atol=1e-4 * np.abs(ref_w).max())


def test_elementwise_kernels(dn):
    import torch

    rng = np.random.default_rng(5)
    n, d, h, w, C = 1, 4, 4, 8, 32
    P = n * d * h * w
    x = _rounded(rng.standard_normal((n, d, h, w, C)))
    xd = _dev(_bf(x))
    # pooling
    yp = torch.empty((n, d // 2, h // 2, w // 2, C), dtype=torch.bfloat16, device="cuda:0")
    yp_cm = torch.empty((C, P // 8), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_pool_fwd", xd.data_ptr(), yp.data_ptr(), yp_cm.data_ptr(), n, d // 2, h // 2, w // 2, C, None)
    dxp = torch.empty_like(xd)
    dn._call("hdn_pool_bwd", yp.data_ptr(), dxp.data_ptr(), n, d, h, w, C, None)
    # upsampling of the pooled tensor and its gradient
    yu = torch.empty_like(xd)
    yu_cm = torch.empty((C, P), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_up_fwd", yp.data_ptr(), yu.data_ptr(), yu_cm.data_ptr(), n, d, h, w, C, None)
    dxu = torch.empty_like(yp)
    dn._call("hdn_up_bwd", xd.data_ptr(), dxu.data_ptr(), n, d // 2, h // 2, w // 2, C, None)
    # add and activation gradient
    s = torch.empty_like(xd)
    s_cm = torch.empty((C, P), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_add", xd.data_ptr(), yu.data_ptr(), s.data_ptr(), s_cm.data_ptr(), P, C, None)
    dy = _rounded(rng.standard_normal(x.shape))
    dyd = _dev(_bf(dy))
    dz = torch.empty_like(xd)
    dz_cm = torch.empty((C, P), dtype=torch.bfloat16, device="cuda:0")
    dn._call("hdn_act_bwd", xd.data_ptr(), dyd.data_ptr(), dz.data_ptr(), dz_cm.data_ptr(), P, C, 0.1, None)
    # channel sums
    part = torch.empty((3, C), dtype=torch.float32, device="cuda:0")
    cs = torch.empty(C, dtype=torch.float32, device="cuda:0")
    dn._call("hdn_channel_sum", dz_cm.data_ptr(), part.data_ptr(), P, C, 3, None)
    dn._call("hdn_reduce_splits", part.data_ptr(), cs.data_ptr(), C, 3, None)
    torch.cuda.synchronize()

    pooled = cr.pool_fwd(x)
    _close_bf16(_np(yp), pooled)
    np.testing.assert_array_equal(_np(yp_cm), _np(yp).reshape(-1, C).T)
    _close_bf16(_np(dxp), cr.pool_bwd(_np(yp)))
    np.testing.assert_array_equal(_np(yu), cr.up_fwd(_np(yp)))
    _close_bf16(_np(dxu), cr.up_bwd(x))
    _close_bf16(_np(s), x + _np(yu))
    np.testing.assert_array_equal(_np(s_cm), _np(s).reshape(-1, C).T)
    _close_bf16(_np(dz), cr.leaky_bwd(x, dy, 0.1))
    np.testing.assert_allclose(_np(cs), _np(dz).reshape(-1, C).sum(axis=0), rtol=1e-5, atol=1e-5)


def test_loss_kernel(dn):
    import torch

    rng = np.random.default_rng(6)
    P = 1000
    out = _rounded(rng.standard_normal((P, 16)))
    inp = _rounded(rng.standard_normal((P, 16)))
    target = rng.standard_normal(P).astype(np.float32)
    weight = rng.uniform(0.5, 2.0, P).astype(np.float32)
    od, idd = _dev(_bf(out)), _dev(_bf(inp))
    td, wd = _dev(torch.from_numpy(target)), _dev(torch.from_numpy(weight))
    dout = torch.empty_like(od)
    dout_cm = torch.empty((16, P), dtype=torch.bfloat16, device="cuda:0")
    pred = torch.empty(P, dtype=torch.float32, device="cuda:0")
    partial = torch.empty((P + 255) // 256, dtype=torch.float32, device="cuda:0")
    dn._call("hdn_loss", od.data_ptr(), idd.data_ptr(), 16, 16, td.data_ptr(), wd.data_ptr(), dout.data_ptr(),
             dout_cm.data_ptr(), pred.data_ptr(), partial.data_ptr(), P, 0.5, None)
    torch.cuda.synchronize()
    loss, pref, grad = cr.loss(out[:, 0], inp[:, 0], target.astype(np.float64), weight.astype(np.float64))
    np.testing.assert_allclose(_np(pred), pref, rtol=1e-6, atol=1e-6)
    assert float(partial.sum()) == pytest.approx(loss, rel=1e-5)
    g = _np(dout)
    _close_bf16(g[:, 0], 0.5 * grad)
    assert not g[:, 1:].any()
    np.testing.assert_array_equal(_np(dout_cm), g.T)


def test_network_gradients_match_oracle(dn):
    # The whole encoder-decoder, forward and backward on the GPU in bf16. Layer by
    # layer, each kernel's output is checked against the oracle fed the GPU's own bf16
    # input and output gradient (tight tolerances); end to end, the GPU's head output
    # and weight gradients follow the oracle's float64 network on the same weights and
    # input within what bf16 activations through 15 layers allow.
    import torch

    widths = (16, 16, 32, 32)
    model = dn.Conv3dDenoiser(widths=widths, seed=3)
    model.keep_gradients = True
    rng = np.random.default_rng(7)
    shape = (1, 32, 32, 64)
    noisy = rng.random(shape).astype(np.float32)
    dens = rng.uniform(0.2, 1.8, shape).astype(np.float32)
    x, x_cm = dn.pack_input(noisy, dens * 0.5)
    out, cache = model.forward(x, x_cm, keep=True)
    r = rng.standard_normal(shape + (16,)).astype(np.float32)
    r[..., 1:] = 0.0
    rq = _rounded(r)
    rd = _dev(_bf(r))
    rd_cm = _cm(rd, 16)
    model.backward(cache, rd, rd_cm)
    torch.cuda.synchronize()

    for name, lay in model.layers.items():
        inp, _, y, _, _ = cache[name]
        dz = _np(cache[name + "_dz"])
        xin = _np(inp)
        wq = _np(lay.wf)
        # forward of this layer from its actual input
        z = cr.conv_fwd(xin, wq, _np(lay.b))
        _close_bf16(_np(y), z if name == "head" else cr.leaky(z, dn.SLOPE))
        # weight and bias gradients from the actual input and output gradient
        ref_w = cr.conv_wgrad(xin, dz)
        np.testing.assert_allclose(_np(lay.gw), ref_w, rtol=1e-3, atol=1e-4 * np.abs(ref_w).max())
        ref_b = dz.sum(axis=(0, 1, 2, 3))
        np.testing.assert_allclose(_np(lay.gb), ref_b, rtol=1e-3, atol=1e-4 * np.abs(ref_b).max() + 1e-12)
        if name != "e0a":
            _close_bf16(_np(cache[name + "_dx"]), cr.conv_dgrad(dz, wq))

    params = {name: (_np(lay.wf), _np(lay.b)) for name, lay in model.layers.items()}
    ref_out, ref_cache = cr.network_forward(params, _np(x), dn.SLOPE)
    rel = np.linalg.norm(_np(out) - ref_out) / np.linalg.norm(ref_out)
    assert rel < 0.05, rel
    grads = cr.network_backward(params, ref_cache, rq, dn.SLOPE)
    for name, lay in model.layers.items():
        got, ref = _np(lay.gw), grads[name][0]
        nr = np.linalg.norm(ref)
        cos = float(np.sum(got * ref) / (np.linalg.norm(got) * nr))
        err = float(np.linalg.norm(got - ref) / nr)
        assert cos > 0.99 and err < 0.15, (name, cos, err)


def _small_ct(dn):
    from pyradmc_amd import SpectralBeamletSource, VoxelGrid, bremsstrahlung_6mv_like
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    rng = np.random.default_rng(12)
    shape = (72, 70, 24)
    density = np.where(rng.random(shape) < 0.1, 0.3, 1.0)
    grid = VoxelGrid(shape=shape, spacing=(0.25, 0.25, 0.25), density=density,
                     material=np.zeros(shape, dtype=np.int32))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    centers = [(8.0, 8.0, 3.0), (9.5, 8.5, 3.0), (7.0, 9.0, 3.0)]
    src = SpectralBeamletSource(bremsstrahlung_6mv_like(), (8.5, 8.5, -97.0), centers, 0.5, 0.5)
    return grid, xs, src


def test_window_gather_scatter_round_trip(dn):
    import ctypes

    import torch

    from pyradmc_amd import HipEngine, _native

    grid, xs, src = _small_ct(dn)
    eng = HipEngine(grid, xs)
    ref = eng.run_dij(src, 20_000, 4, seed=3, truncation=0.0)
    windows = dn._windows_for(grid, src, range(3))
    origins = np.ascontiguousarray([[wx.start, wy.start] for wx, wy in windows], dtype=np.int32)
    counts = np.full(3, 20_000, dtype=np.int64)
    books = {}
    eng._dij_group(src, range(3), counts, 4, 3, True, eng._default_scoring(), books)
    out = torch.empty((3, 24, 64, 64), dtype=torch.float32, device="cuda:0")
    eng._ctx.call("hmc_gather_windows", _native.ptr(origins, ctypes.c_int32), 3, 64, 64, out.data_ptr())
    got = out.cpu().numpy()
    for j, (wx, wy) in enumerate(windows):
        expect = np.transpose(ref.column_dense(j)[wx, wy, :], (2, 0, 1)).astype(np.float32)
        np.testing.assert_array_equal(got[j], expect)  # same seed, same device arithmetic
    # scatter the windows back (sigma = the mean, to see it land in the sigma output too)
    eng._ctx.call("hmc_scatter_windows", _native.ptr(origins, ctypes.c_int32), 3, 64, 64, out.data_ptr(),
                  out.data_ptr())
    kept = np.zeros(3, dtype=np.int64)
    total = _native.c_int64()
    eng._ctx.call("hmc_truncate", 0.0, _native.ptr(kept, _native.c_int64), ctypes.byref(total))
    n = int(total.value)
    idx, dose, sig = np.empty(n, np.int64), np.empty(n), np.empty(n)
    eng._ctx.call("hmc_fetch_csc", _native.ptr(idx, _native.c_int64), _native.ptr(dose, ctypes.c_double),
                  _native.ptr(sig, ctypes.c_double))
    eng.close()
    lo = 0
    for j, (wx, wy) in enumerate(windows):
        dense = np.zeros(grid.shape)
        dense.reshape(-1)[idx[lo:lo + kept[j]]] = dose[lo:lo + kept[j]]
        lo += int(kept[j])
        expect = np.zeros(grid.shape)
        expect[wx, wy, :] = ref.column_dense(j)[wx, wy, :].astype(np.float32)
        np.testing.assert_array_equal(dense, expect)
    np.testing.assert_array_equal(sig, dose)


def test_denoised_builder_falls_back_to_plain_monte_carlo(dn):
    # An untrained network and a zero error budget: every column is re-simulated, and
    # the result is the plain Monte Carlo Dij bit for bit.
    from pyradmc_amd import HipEngine

    grid, xs, src = _small_ct(dn)
    eng = HipEngine(grid, xs)
    model = dn.Conv3dDenoiser(seed=5)
    builder = dn.DenoisedDijBuilder(eng, model, low_histories=4_000, tolerance=0.0,
                                    fallback_histories=40_000, n_batches=4)
    dij, report = builder.build(src, seed=21, truncation=1e-4)
    plain = eng.run_dij(src, 40_000, 4, seed=23, truncation=1e-4)
    eng.close()
    assert report.fallback.all()
    np.testing.assert_array_equal(dij.indptr, plain.indptr)
    np.testing.assert_array_equal(dij.indices, plain.indices)
    np.testing.assert_array_equal(dij.dose, plain.dose)
    np.testing.assert_array_equal(dij.sigma, plain.sigma)
    assert abs(dij.energy_emitted - dij.energy_deposited - dij.energy_escaped - dij.energy_unscored) \
        < 1e-6 * dij.energy_emitted


def test_denoised_builder_float32_values_are_the_float64_values_rounded(dn):
    # No fallbacks (an unbounded error budget): the same seed gives the same denoised
    # columns, and a float32 CSC holds the float64 one rounded to nearest.
    from pyradmc_amd import HipEngine

    grid, xs, src = _small_ct(dn)
    eng = HipEngine(grid, xs)
    builder = dn.DenoisedDijBuilder(eng, dn.Conv3dDenoiser(seed=5), low_histories=4_000, tolerance=1e9,
                                    fallback_histories=40_000, n_batches=4)
    wide, report = builder.build(src, seed=31, truncation=1e-4)
    narrow, _ = builder.build(src, seed=31, truncation=1e-4, value_dtype=np.float32)
    with pytest.raises(ValueError):
        builder.build(src, seed=31, value_dtype=np.float16)
    eng.close()
    # only the judge's sample is plain Monte Carlo
    assert report.audited.sum() == 1 and np.array_equal(report.fallback, report.audited) and wide.indptr[-1] > 0
    assert wide.dose.dtype == np.float64 and narrow.dose.dtype == np.float32
    np.testing.assert_array_equal(narrow.indptr, wide.indptr)
    np.testing.assert_array_equal(narrow.indices, wide.indices)
    np.testing.assert_array_equal(narrow.dose, wide.dose.astype(np.float32))
    np.testing.assert_array_equal(narrow.sigma, wide.sigma.astype(np.float32))


@pytest.mark.parametrize("widths", [(16, 32, 64, 128), (16, 16, 32, 32)])
def test_inference_with_fused_inputs_equals_the_stored_forward(dn, widths):
    # Inference reads each upsampling and skip addition inside the next convolution
    # (full-resolution tiled and general kernels alike); the values are those the
    # separate kernels store, so the head output is the same bit for bit.
    import torch

    model = dn.Conv3dDenoiser(widths=widths, seed=11)
    rng = np.random.default_rng(12)
    shape = (2, 16, 32, 64)
    x, x_cm = dn.pack_input(rng.random(shape).astype(np.float32),
                            rng.uniform(0.1, 0.9, shape).astype(np.float32))
    stored, _ = model.forward(x, x_cm, keep=True)
    fused, _ = model.forward(x, None, keep=False)
    torch.cuda.synchronize()
    assert stored.dtype == fused.dtype and stored.shape == fused.shape
    assert torch.equal(stored.view(torch.int16), fused.view(torch.int16))
    assert torch.count_nonzero(stored.float()) > 0


@pytest.mark.parametrize("shape", [(1, 4, 8, 64), (2, 4, 4, 16)])
def test_first_layer_skips_the_zero_channel_group(dn, shape):
    # src 3 reads only the first 8 input channels; with the rest zero (the network's
    # padded input) the outputs equal the full reduction's (tiled and general kernels).
    import torch

    model = dn.Conv3dDenoiser(seed=4)
    lay = model.layers["e0a"]
    rng = np.random.default_rng(9)
    x = rng.standard_normal(shape + (16,)).astype(np.float32)
    x[..., 8:] = 0.0
    xd = torch.from_numpy(x).to(torch.bfloat16).to("cuda:0")
    full, _ = model._conv(lay, xd, shape, True, False, src=0)
    low, _ = model._conv(lay, xd, shape, True, False, src=3)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(low.float().cpu().numpy(), full.float().cpu().numpy())


def test_fused_convolution_arguments_rejected(dn):
    # Checked on the host before any launch: unknown source mode, an upsampled source
    # with odd output dims, an added source without its second input.
    import torch

    model = dn.Conv3dDenoiser(seed=4)
    lay = model.layers["e0b"]
    x = torch.zeros((1, 4, 4, 16, 16), dtype=torch.bfloat16, device="cuda:0")
    with pytest.raises(ValueError):
        model._conv(lay, x, (1, 4, 4, 16), True, False, src=7)
    with pytest.raises(ValueError):
        model._conv(lay, x, (1, 3, 4, 16), True, False, src=1)
    with pytest.raises(ValueError):
        model._conv(lay, x, (1, 4, 4, 16), True, False, x2=None, src=2)
