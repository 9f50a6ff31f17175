"""ORACLE (test infrastructure only): CPU restatement of the denoiser network's ops.

Only tests/ load this module, as the checker of the HIP kernels in
pyradmc_amd/csrc/denoise_kernels.hip. The reference has no learned component, so
nothing here restates reference code; the ops are the textbook definitions, in
float64 NumPy, on the same layouts as the kernels:

* activations position-major ``[N][D][H][W][C]``;
* weights ``[27][Cout][Cin]`` with tap ``(kz * 3 + ky) * 3 + kx`` at offset
  ``(kz - 1, ky - 1, kx - 1)``: ``y[p][co] = b[co] + sum x[p + off][ci] w[tap][co][ci]``
  with zero padding (a cross-correlation, as in every deep-learning framework).

Pins (tests/test_denoise_oracle.py): the forward against scipy.signal.correlate per
channel pair; every backward against central finite differences of its forward.
"""

from __future__ import annotations

import numpy as np

OFFSETS = [(kz - 1, ky - 1, kx - 1) for kz in range(3) for ky in range(3) for kx in range(3)]


def _shifted(x: np.ndarray, off) -> np.ndarray:
    """x[n, z + dz, y + dy, x + dx, c] with zeros outside."""
    dz, dy, dx = off
    out = np.zeros_like(x)
    n, d, h, w, _ = x.shape

    def rng(delta, size):
        lo_out = max(0, -delta)
        hi_out = min(size, size - delta)
        return slice(lo_out, hi_out), slice(lo_out + delta, hi_out + delta)

    (zo, zi), (yo, yi), (xo, xi) = rng(dz, d), rng(dy, h), rng(dx, w)
    out[:, zo, yo, xo, :] = x[:, zi, yi, xi, :]
    return out


def conv_fwd(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """Pre-activation z[p][co] = b[co] + sum_{tap, ci} x[p + off(tap)][ci] w[tap][co][ci]."""
    y = np.zeros(x.shape[:-1] + (w.shape[1],))
    for tap, off in enumerate(OFFSETS):
        y += _shifted(x, off) @ w[tap].T
    if b is not None:
        y += b
    return y


def conv_dgrad(dz: np.ndarray, w: np.ndarray) -> np.ndarray:
    """dx[p][ci] = sum_{tap, co} dz[p - off(tap)][co] w[tap][co][ci]."""
    dx = np.zeros(dz.shape[:-1] + (w.shape[2],))
    for tap, (oz, oy, ox) in enumerate(OFFSETS):
        dx += _shifted(dz, (-oz, -oy, -ox)) @ w[tap]
    return dx


def conv_wgrad(x: np.ndarray, dz: np.ndarray) -> np.ndarray:
    """dw[tap][co][ci] = sum_p dz[p][co] x[p + off(tap)][ci]."""
    cin, cout = x.shape[-1], dz.shape[-1]
    dzf = dz.reshape(-1, cout)
    out = np.zeros((27, cout, cin))
    for tap, off in enumerate(OFFSETS):
        out[tap] = dzf.T @ _shifted(x, off).reshape(-1, cin)
    return out


def leaky(z: np.ndarray, slope: float) -> np.ndarray:
    return np.where(z >= 0.0, z, slope * z)


def leaky_bwd(y: np.ndarray, dy: np.ndarray, slope: float) -> np.ndarray:
    """From the stored output: act'(z) = 1 where y > 0, slope elsewhere."""
    return np.where(y > 0.0, dy, slope * dy)


def pool_fwd(x: np.ndarray) -> np.ndarray:
    n, d, h, w, c = x.shape
    return x.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4, 6))


def pool_bwd(dy: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2), 2, axis=3) / 8.0


def up_fwd(x: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3)


def up_bwd(dy: np.ndarray) -> np.ndarray:
    n, d2, h2, w2, c = dy.shape
    return dy.reshape(n, d2 // 2, 2, h2 // 2, 2, w2 // 2, 2, c).sum(axis=(2, 4, 6))


def loss(out0: np.ndarray, inp0: np.ndarray, target: np.ndarray, weight: np.ndarray):
    """Residual head pred = inp0 + out0 and sum weight (pred - target)^2; returns
    (loss, pred, d loss / d out0)."""
    pred = inp0 + out0
    diff = pred - target
    return float(np.sum(weight * diff * diff)), pred, 2.0 * weight * diff


# ------------------------------------------------------------------ the network
#
# Encoder-decoder with additive skips (pyradmc_amd/denoise.py builds the same graph):
#   level 0: conv(cin0 -> c0), conv(c0 -> c0)        -> skip0      (full size)
#   pool; level 1: conv(c0 -> c1), conv(c1 -> c1)    -> skip1
#   pool; level 2: conv(c1 -> c2), conv(c2 -> c2)    -> skip2
#   pool; level 3: conv(c2 -> c3), conv(c3 -> c3)
#   up; conv(c3 -> c2) + skip2; conv(c2 -> c2)
#   up; conv(c2 -> c1) + skip1; conv(c1 -> c1)
#   up; conv(c1 -> c0) + skip0; conv(c0 -> c0)
#   head conv(c0 -> 16) without activation; prediction = input channel 0 + head channel 0


def layer_specs(cin0: int, widths: tuple[int, int, int, int]):
    c0, c1, c2, c3 = widths
    return [
        ("e0a", cin0, c0), ("e0b", c0, c0),
        ("e1a", c0, c1), ("e1b", c1, c1),
        ("e2a", c1, c2), ("e2b", c2, c2),
        ("e3a", c2, c3), ("e3b", c3, c3),
        ("d2a", c3, c2), ("d2b", c2, c2),
        ("d1a", c2, c1), ("d1b", c1, c1),
        ("d0a", c1, c0), ("d0b", c0, c0),
        ("head", c0, 16),
    ]


def network_forward(params: dict, x: np.ndarray, slope: float):
    """Returns (head output, cache of every activation for the backward)."""
    cache = {"x": x}

    def conv(name, inp, activate=True):
        w, b = params[name]
        z = conv_fwd(inp, w, b)
        y = leaky(z, slope) if activate else z
        cache[name + "_in"] = inp
        cache[name] = y
        return y

    h = conv("e0a", x)
    s0 = conv("e0b", h)
    h = conv("e1a", pool_fwd(s0))
    s1 = conv("e1b", h)
    h = conv("e2a", pool_fwd(s1))
    s2 = conv("e2b", h)
    h = conv("e3a", pool_fwd(s2))
    h = conv("e3b", h)
    h = conv("d2a", up_fwd(h)) + s2
    h = conv("d2b", h)
    h = conv("d1a", up_fwd(h)) + s1
    h = conv("d1b", h)
    h = conv("d0a", up_fwd(h)) + s0
    h = conv("d0b", h)
    out = conv("head", h, activate=False)
    return out, cache


def network_backward(params: dict, cache: dict, dout: np.ndarray, slope: float) -> dict:
    """Gradients {name: (dw, db)} of sum(dout * out)."""
    grads = {}

    def back(name, dy, activate=True):
        w, _ = params[name]
        dz = leaky_bwd(cache[name], dy, slope) if activate else dy
        grads[name] = (conv_wgrad(cache[name + "_in"], dz), dz.sum(axis=(0, 1, 2, 3)))
        return conv_dgrad(dz, w)

    dh = back("head", dout, activate=False)
    dh = back("d0b", dh)
    ds0 = dh
    dh = up_bwd(back("d0a", dh))
    dh = back("d1b", dh)
    ds1 = dh
    dh = up_bwd(back("d1a", dh))
    dh = back("d2b", dh)
    ds2 = dh
    dh = up_bwd(back("d2a", dh))
    dh = back("e3b", dh)
    dh = pool_bwd(back("e3a", dh))
    dh = back("e2b", dh + ds2)
    dh = pool_bwd(back("e2a", dh))
    dh = back("e1b", dh + ds1)
    dh = pool_bwd(back("e1a", dh))
    dh = back("e0b", dh + ds0)
    back("e0a", dh)
    return grads
