"""Pins of the denoiser oracle (oracle/conv_ref.py): forward against a library
cross-correlation, every backward against central finite differences of its
forward (float64), including the whole encoder-decoder."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import signal

from oracle import conv_ref as cr


def _rand(rng, *shape):
    return rng.standard_normal(shape)


def test_conv_forward_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = _rand(rng, 2, 5, 6, 7, 3)
    w = _rand(rng, 27, 4, 3)
    b = _rand(rng, 4)
    y = cr.conv_fwd(x, w, b)
    for n in range(2):
        for co in range(4):
            ref = np.full(x.shape[1:4], b[co])
            for ci in range(3):
                k = w[:, co, ci].reshape(3, 3, 3)  # tap = (kz*3 + ky)*3 + kx
                ref += signal.correlate(x[n, ..., ci], k, mode="same", method="direct")
            np.testing.assert_allclose(y[n, ..., co], ref, rtol=1e-12, atol=1e-12)


def _fd_check(f, arg, grad, rng, n_probe=12, eps=1e-6):
    """Central differences of scalar f along a few random coordinates of arg."""
    flat = arg.reshape(-1)
    g = grad.reshape(-1)
    for idx in rng.choice(flat.size, size=min(n_probe, flat.size), replace=False):
        old = flat[idx]
        flat[idx] = old + eps
        up = f()
        flat[idx] = old - eps
        dn = f()
        flat[idx] = old
        num = (up - dn) / (2 * eps)
        assert num == pytest.approx(g[idx], rel=1e-5, abs=1e-7), (idx, num, g[idx])


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(2)
    x = _rand(rng, 1, 4, 5, 6, 3)
    w = _rand(rng, 27, 2, 3)
    r = _rand(rng, 1, 4, 5, 6, 2)  # loss = sum(r * conv(x, w))
    f = lambda: float(np.sum(r * cr.conv_fwd(x, w)))
    _fd_check(f, x, cr.conv_dgrad(r, w), rng)
    _fd_check(f, w, cr.conv_wgrad(x, r), rng)


def test_pool_up_act_loss_backward_finite_differences():
    rng = np.random.default_rng(3)
    x = _rand(rng, 1, 4, 4, 6, 2)
    r = _rand(rng, 1, 2, 2, 3, 2)
    _fd_check(lambda: float(np.sum(r * cr.pool_fwd(x))), x, cr.pool_bwd(r), rng)
    xc = _rand(rng, 1, 2, 2, 3, 2)
    rf = _rand(rng, 1, 4, 4, 6, 2)
    _fd_check(lambda: float(np.sum(rf * cr.up_fwd(xc))), xc, cr.up_bwd(rf), rng)
    z = _rand(rng, 3, 4, 5)
    rz = _rand(rng, 3, 4, 5)
    _fd_check(lambda: float(np.sum(rz * cr.leaky(z, 0.1))), z, cr.leaky_bwd(cr.leaky(z, 0.1), rz, 0.1), rng)
    out0, inp0, t, wt = (_rand(rng, 20) for _ in range(4))
    wt = np.abs(wt)
    _fd_check(lambda: cr.loss(out0, inp0, t, wt)[0], out0, cr.loss(out0, inp0, t, wt)[2], rng)


def test_network_backward_finite_differences():
    rng = np.random.default_rng(4)
    specs = cr.layer_specs(2, (3, 4, 5, 6))
    params = {name: (0.3 * _rand(rng, 27, cout, cin), 0.1 * _rand(rng, cout)) for name, cin, cout in specs}
    x = _rand(rng, 1, 8, 8, 8, 2)
    r = _rand(rng, 1, 8, 8, 8, 16)

    def f():
        out, _ = cr.network_forward(params, x, 0.1)
        return float(np.sum(r * out))

    _, cache = cr.network_forward(params, x, 0.1)
    grads = cr.network_backward(params, cache, r, 0.1)
    for name in ("e0a", "e2b", "e3a", "d1a", "d0b", "head"):
        w, b = params[name]
        gw, gb = grads[name]
        _fd_check(f, w, gw, rng, n_probe=4)
        _fd_check(f, b, gb, rng, n_probe=2)
