"""Learned Dij-column denoiser: a 3D convolutional encoder-decoder on hand-written
BF16 MFMA kernels, trained on column pairs the engine itself generates.

A beamlet column computed with few histories is denoised inside a beam-aligned
window (64 x 64 voxels around the beamlet's central axis, full depth) from two
input channels - the low-history dose and the density - into a prediction of the
full-statistics column. Plain Monte Carlo stays the fallback and the judge: every
column is also scored by a second, independent low-history run; when the two
denoised estimates disagree by more than the error budget, the column is
re-simulated with full statistics (``DenoisedDijBuilder``).

The reference has no learned component; this module extends its Dij builder. The
network's forward and backward passes run entirely in this build's HIP kernels
(``include/hipdn.h``); PyTorch only allocates the device buffers and supplies the
stream.

Network: four levels of two 3x3x3 convolutions (widths 16, 32, 64, 128 by default,
about 1.3 M parameters), 2x2x2 average pooling down, nearest upsampling up,
additive skips, leaky ReLU (slope 0.1), and a residual head: prediction = noisy
input + head output. Loss: squared error weighted by 1 / (sigma_target^2 + floor^2)
(the target's own Monte Carlo variance, floored), averaged over voxels.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from pyradmc_amd import _native

SLOPE = 0.1
CIN_PAD = 16          # input channels padded to the MFMA k-group multiple
HEAD_OUT = 16         # head output channels (channel 0 is the prediction)
WINDOW_XY = 64        # beam-aligned window, voxels
SIGMA_FLOOR = 0.02    # loss weight floor, in units of the column scale

_c_void = ctypes.c_void_p
_c_ll = ctypes.c_longlong


def _lib():
    h = _native.lib()
    if getattr(h, "_hdn_bound", False):
        return h
    i, f = ctypes.c_int, ctypes.c_float
    sig = {
        "hdn_conv3d_fwd": [_c_void, _c_void, _c_void, _c_void, _c_void, i, i, i, i, i, i, f, _c_void],
        "hdn_conv3d_wgrad": [_c_void, _c_void, _c_void, i, i, i, i, i, i, i, _c_void],
        "hdn_reduce_splits": [_c_void, _c_void, _c_ll, i, _c_void],
        "hdn_channel_sum": [_c_void, _c_void, _c_ll, i, i, _c_void],
        "hdn_act_bwd": [_c_void, _c_void, _c_void, _c_void, _c_ll, i, f, _c_void],
        "hdn_add": [_c_void, _c_void, _c_void, _c_void, _c_ll, i, _c_void],
        "hdn_pool_fwd": [_c_void, _c_void, _c_void, i, i, i, i, i, _c_void],
        "hdn_pool_bwd": [_c_void, _c_void, i, i, i, i, i, _c_void],
        "hdn_up_fwd": [_c_void, _c_void, _c_void, i, i, i, i, i, _c_void],
        "hdn_up_bwd": [_c_void, _c_void, i, i, i, i, i, _c_void],
        "hdn_loss": [_c_void, _c_void, i, i, _c_void, _c_void, _c_void, _c_void, _c_void, _c_void,
                     _c_ll, f, _c_void],
        "hdn_prepare_weights": [_c_void, _c_void, _c_void, i, i, _c_void],
        "hdn_adam": [_c_void, _c_void, _c_void, _c_void, _c_ll, f, f, f, f, i, f, _c_void],
    }
    for name, argtypes in sig.items():
        fn = getattr(h, name)
        fn.restype = ctypes.c_int
        fn.argtypes = argtypes
    h._hdn_bound = True
    return h


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the denoiser needs a HIP device (torch.cuda is unavailable)")
    return torch


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def _call(name: str, *args) -> None:
    rc = getattr(_lib(), name)(*args)
    if rc != 0:
        raise (ValueError if rc == -1 else RuntimeError)(f"{name} failed with status {rc}")


def layer_specs(cin0: int, widths: tuple[int, int, int, int]):
    """(name, cin, cout) of every convolution, in forward order (the oracle builds the
    same list)."""
    c0, c1, c2, c3 = widths
    return [
        ("e0a", cin0, c0), ("e0b", c0, c0),
        ("e1a", c0, c1), ("e1b", c1, c1),
        ("e2a", c1, c2), ("e2b", c2, c2),
        ("e3a", c2, c3), ("e3b", c3, c3),
        ("d2a", c3, c2), ("d2b", c2, c2),
        ("d1a", c2, c1), ("d1b", c1, c1),
        ("d0a", c1, c0), ("d0b", c0, c0),
        ("head", c0, HEAD_OUT),
    ]


@dataclass
class _Layer:
    name: str
    cin: int
    cout: int
    w: object = None      # fp32 master [27][cout][cin]
    b: object = None      # fp32 [cout]
    wf: object = None     # bf16 forward copy
    wd: object = None     # bf16 data-gradient copy
    gw: object = None
    gb: object = None
    mw: object = None
    vw: object = None
    mb: object = None
    vb: object = None


@dataclass
class Conv3dDenoiser:
    """The encoder-decoder on one HIP device; parameters live on the device."""

    widths: tuple[int, int, int, int] = (16, 32, 64, 128)
    seed: int = 0
    device: str = "cuda:0"
    layers: dict = field(default_factory=dict, init=False, repr=False)
    step_count: int = field(default=0, init=False)
    # Keep every layer's output gradient and input gradient in the forward cache (a
    # testing aid: cache[name + "_dz"], cache[name + "_dx"]).
    keep_gradients: bool = field(default=False, init=False)

    def __post_init__(self) -> None:
        torch = _torch()
        rng = np.random.default_rng(self.seed)
        dev = torch.device(self.device)
        for name, cin, cout in layer_specs(CIN_PAD, self.widths):
            real_in = 2 if name == "e0a" else cin
            std = math.sqrt(2.0 / (27.0 * real_in)) if name != "head" else 1e-3
            w = np.zeros((27, cout, cin), dtype=np.float32)
            real_out = 1 if name == "head" else cout
            w[:, :real_out, :real_in] = rng.normal(0.0, std, size=(27, real_out, real_in))
            lay = _Layer(name, cin, cout)
            lay.w = torch.from_numpy(w).to(dev)
            lay.b = torch.zeros(cout, dtype=torch.float32, device=dev)
            lay.wf = torch.empty((27, cout, cin), dtype=torch.bfloat16, device=dev)
            lay.wd = torch.empty((27, cin, cout), dtype=torch.bfloat16, device=dev)
            lay.gw = torch.zeros_like(lay.w)
            lay.gb = torch.zeros_like(lay.b)
            lay.mw, lay.vw = torch.zeros_like(lay.w), torch.zeros_like(lay.w)
            lay.mb, lay.vb = torch.zeros_like(lay.b), torch.zeros_like(lay.b)
            self.layers[name] = lay
            self._prepare(lay)

    # ---------------------------------------------------------------- plumbing

    @property
    def n_parameters(self) -> int:
        """Trainable (unpadded) parameters."""
        total = 0
        for name, cin, cout in layer_specs(CIN_PAD, self.widths):
            real_in = 2 if name == "e0a" else cin
            real_out = 1 if name == "head" else cout
            total += 27 * real_in * real_out + real_out
        return total

    def _stream(self):
        return _torch().cuda.current_stream().cuda_stream

    def _prepare(self, lay: _Layer) -> None:
        _call("hdn_prepare_weights", _ptr(lay.w), _ptr(lay.wf), _ptr(lay.wd), lay.cin, lay.cout,
              self._stream())

    def state_dict(self) -> dict:
        out = {}
        for name, lay in self.layers.items():
            out[name + ".w"] = lay.w.detach().cpu().numpy().copy()
            out[name + ".b"] = lay.b.detach().cpu().numpy().copy()
        return out

    def load_state_dict(self, state: dict) -> None:
        torch = _torch()
        for name, lay in self.layers.items():
            lay.w.copy_(torch.from_numpy(np.asarray(state[name + ".w"], dtype=np.float32)))
            lay.b.copy_(torch.from_numpy(np.asarray(state[name + ".b"], dtype=np.float32)))
            self._prepare(lay)

    # ---------------------------------------------------------------- forward

    def _conv(self, lay: _Layer, x, dims, activate: bool, want_cm: bool):
        torch = _torch()
        n, d, h, w = dims
        y = torch.empty((n, d, h, w, lay.cout), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((lay.cout, n * d * h * w), dtype=torch.bfloat16, device=x.device) if want_cm else None
        _call("hdn_conv3d_fwd", _ptr(x), _ptr(lay.wf), _ptr(lay.b), _ptr(y), _ptr(y_cm), n, d, h, w,
              lay.cin, lay.cout, SLOPE if activate else 1.0, self._stream())
        return y, y_cm

    def _pool(self, x, dims, C):
        torch = _torch()
        n, d, h, w = dims
        y = torch.empty((n, d // 2, h // 2, w // 2, C), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((C, y.numel() // C), dtype=torch.bfloat16, device=x.device)
        _call("hdn_pool_fwd", _ptr(x), _ptr(y), _ptr(y_cm), n, d // 2, h // 2, w // 2, C, self._stream())
        return y, y_cm

    def _up(self, x, dims_fine, C):
        torch = _torch()
        n, d, h, w = dims_fine
        y = torch.empty((n, d, h, w, C), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((C, y.numel() // C), dtype=torch.bfloat16, device=x.device)
        _call("hdn_up_fwd", _ptr(x), _ptr(y), _ptr(y_cm), n, d, h, w, C, self._stream())
        return y, y_cm

    def _add(self, a, b, C):
        torch = _torch()
        y = torch.empty_like(a)
        y_cm = torch.empty((C, a.numel() // C), dtype=torch.bfloat16, device=a.device)
        _call("hdn_add", _ptr(a), _ptr(b), _ptr(y), _ptr(y_cm), a.numel() // C, C, self._stream())
        return y, y_cm

    def forward(self, x, x_cm, keep: bool = True):
        """x: [N][D][H][W][16] bf16 (channel 0 noisy dose, 1 density, rest 0) and its
        channel-major copy (only needed for training). D, H, W divisible by 8 with
        W % 64 == 0 at full size. Returns (head output [N][D][H][W][16], cache)."""
        n, d, h, w, _ = x.shape
        L = self.layers
        c0, c1, c2, c3 = self.widths
        dims = [(n, d, h, w), (n, d // 2, h // 2, w // 2), (n, d // 4, h // 4, w // 4),
                (n, d // 8, h // 8, w // 8)]
        cache = {"x": (x, x_cm), "dims": dims}

        def conv(name, inp, inp_cm, lvl, activate=True):
            y, y_cm = self._conv(L[name], inp, dims[lvl], activate, want_cm=keep)
            if keep:
                cache[name] = (inp, inp_cm, y, y_cm, lvl)
            return y, y_cm

        h0 = conv("e0a", x, x_cm, 0)
        s0 = conv("e0b", *h0, 0)
        t = conv("e1a", *self._pool(s0[0], dims[0], c0), 1)
        s1 = conv("e1b", *t, 1)
        t = conv("e2a", *self._pool(s1[0], dims[1], c1), 2)
        s2 = conv("e2b", *t, 2)
        t = conv("e3a", *self._pool(s2[0], dims[2], c2), 3)
        t = conv("e3b", *t, 3)
        t = conv("d2a", *self._up(t[0], dims[2], c3), 2)
        t = conv("d2b", *self._add(t[0], s2[0], c2), 2)
        t = conv("d1a", *self._up(t[0], dims[1], c2), 1)
        t = conv("d1b", *self._add(t[0], s1[0], c1), 1)
        t = conv("d0a", *self._up(t[0], dims[0], c1), 0)
        t = conv("d0b", *self._add(t[0], s0[0], c0), 0)
        out, _ = conv("head", *t, 0, activate=False)
        return out, cache

    # ---------------------------------------------------------------- backward

    def _back(self, name, cache, dy, activate=True, dz_nd=None, dz_cm=None):
        """Gradients of one convolution; returns the gradient of its input (nd)."""
        torch = _torch()
        lay = self.layers[name]
        inp, inp_cm, y, _, lvl = cache[name]
        n, d, h, w = cache["dims"][lvl]
        P = n * d * h * w
        if activate:
            dz_nd = torch.empty_like(y)
            dz_cm = torch.empty((lay.cout, P), dtype=torch.bfloat16, device=y.device)
            _call("hdn_act_bwd", _ptr(y), _ptr(dy), _ptr(dz_nd), _ptr(dz_cm), P, lay.cout, SLOPE,
                  self._stream())
        # weight and bias gradients (deterministic split reductions)
        splits = max(1, min(P // 512, 2048 // max(1, (lay.cin // 16) * (lay.cout // 16))))
        part = torch.empty((splits, 27, lay.cout, lay.cin), dtype=torch.float32, device=y.device)
        _call("hdn_conv3d_wgrad", _ptr(inp_cm), _ptr(dz_cm), _ptr(part), splits, n, d, h, w,
              lay.cin, lay.cout, self._stream())
        _call("hdn_reduce_splits", _ptr(part), _ptr(lay.gw), lay.gw.numel(), splits, self._stream())
        bsplits = max(1, min(P // 4096, 256))
        bpart = torch.empty((bsplits, lay.cout), dtype=torch.float32, device=y.device)
        _call("hdn_channel_sum", _ptr(dz_cm), _ptr(bpart), P, lay.cout, bsplits, self._stream())
        _call("hdn_reduce_splits", _ptr(bpart), _ptr(lay.gb), lay.cout, bsplits, self._stream())
        if self.keep_gradients:
            cache[name + "_dz"] = dz_nd
        if name == "e0a":
            return None
        dx = torch.empty((n, d, h, w, lay.cin), dtype=torch.bfloat16, device=y.device)
        _call("hdn_conv3d_fwd", _ptr(dz_nd), _ptr(lay.wd), None, _ptr(dx), None, n, d, h, w,
              lay.cout, lay.cin, 1.0, self._stream())
        if self.keep_gradients:
            cache[name + "_dx"] = dx
        return dx

    def _pool_bwd(self, dy, dims_fine, C):
        torch = _torch()
        n, d, h, w = dims_fine
        dx = torch.empty((n, d, h, w, C), dtype=torch.bfloat16, device=dy.device)
        _call("hdn_pool_bwd", _ptr(dy), _ptr(dx), n, d, h, w, C, self._stream())
        return dx

    def _up_bwd(self, dy, dims_coarse, C):
        torch = _torch()
        n, d, h, w = dims_coarse
        dx = torch.empty((n, d, h, w, C), dtype=torch.bfloat16, device=dy.device)
        _call("hdn_up_bwd", _ptr(dy), _ptr(dx), n, d, h, w, C, self._stream())
        return dx

    def _sum(self, a, b, C):
        torch = _torch()
        y = torch.empty_like(a)
        _call("hdn_add", _ptr(a), _ptr(b), _ptr(y), None, a.numel() // C, C, self._stream())
        return y

    def backward(self, cache, dout_nd, dout_cm) -> None:
        """Fills every layer's gw, gb from the gradient of the head output."""
        c0, c1, c2, c3 = self.widths
        dims = cache["dims"]
        dh = self._back("head", cache, dout_nd, activate=False, dz_nd=dout_nd, dz_cm=dout_cm)
        da0 = self._back("d0b", cache, dh)
        dh1 = self._up_bwd(self._back("d0a", cache, da0), dims[1], c1)
        da1 = self._back("d1b", cache, dh1)
        dh2 = self._up_bwd(self._back("d1a", cache, da1), dims[2], c2)
        da2 = self._back("d2b", cache, dh2)
        dh3 = self._up_bwd(self._back("d2a", cache, da2), dims[3], c3)
        t = self._back("e3b", cache, dh3)
        t = self._pool_bwd(self._back("e3a", cache, t), dims[2], c2)
        t = self._back("e2b", cache, self._sum(t, da2, c2))
        t = self._pool_bwd(self._back("e2a", cache, t), dims[1], c1)
        t = self._back("e1b", cache, self._sum(t, da1, c1))
        t = self._pool_bwd(self._back("e1a", cache, t), dims[0], c0)
        t = self._back("e0b", cache, self._sum(t, da0, c0))
        self._back("e0a", cache, t)

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
                  weight_decay: float = 0.0) -> None:
        self.step_count += 1
        s = self._stream()
        for lay in self.layers.values():
            _call("hdn_adam", _ptr(lay.w), _ptr(lay.gw), _ptr(lay.mw), _ptr(lay.vw), lay.w.numel(), lr,
                  beta1, beta2, eps, self.step_count, weight_decay, s)
            _call("hdn_adam", _ptr(lay.b), _ptr(lay.gb), _ptr(lay.mb), _ptr(lay.vb), lay.b.numel(), lr,
                  beta1, beta2, eps, self.step_count, 0.0, s)
            self._prepare(lay)

    # ------------------------------------------------------------------ loss

    def loss_and_grad(self, out, x, target, weight, grad_scale: float):
        """(loss partial sums, dout nd, dout cm) for prediction = x[..., 0] + out[..., 0]."""
        torch = _torch()
        P = out.numel() // HEAD_OUT
        dout = torch.empty_like(out)
        dout_cm = torch.empty((HEAD_OUT, P), dtype=torch.bfloat16, device=out.device)
        partial = torch.empty((P + 255) // 256, dtype=torch.float32, device=out.device)
        _call("hdn_loss", _ptr(out), _ptr(x), HEAD_OUT, CIN_PAD, _ptr(target), _ptr(weight),
              _ptr(dout), _ptr(dout_cm), None, _ptr(partial), P, grad_scale, self._stream())
        return partial, dout, dout_cm

    def predict_tensor(self, x):
        """Prediction [N][D][H][W] (float32) for a packed input batch."""
        torch = _torch()
        out, _ = self.forward(x, None, keep=False)
        P = out.numel() // HEAD_OUT
        pred = torch.empty(P, dtype=torch.float32, device=x.device)
        _call("hdn_loss", _ptr(out), _ptr(x), HEAD_OUT, CIN_PAD, None, None, None, None, _ptr(pred),
              None, P, 1.0, self._stream())
        return pred.reshape(x.shape[:4])


# ------------------------------------------------------------------------ data


def to_network_axes(a: np.ndarray) -> np.ndarray:
    """[N][x][y][z] windows -> [N][z][x][y]: the network's innermost axis is a 64-wide
    lateral one (the convolutions are cubic, so this only relabels axes)."""
    return np.ascontiguousarray(np.transpose(a, (0, 3, 1, 2)))


def from_network_axes(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.transpose(a, (0, 2, 3, 1)))


def pack_input(noisy: np.ndarray, density: np.ndarray, device: str = "cuda:0"):
    """[N][D][H][W] scaled noisy dose and density (network axes) -> padded bf16 nd
    tensor and its channel-major copy."""
    torch = _torch()
    x = np.zeros(noisy.shape + (CIN_PAD,), dtype=np.float32)
    x[..., 0] = noisy
    x[..., 1] = density
    t = torch.from_numpy(x).to(device=device, dtype=torch.bfloat16)
    t_cm = t.reshape(-1, CIN_PAD).t().contiguous()
    return t, t_cm


def column_scale(noisy_window: np.ndarray) -> float:
    """A column's normalisation: the mean of its 64 largest window voxels (robust to
    a single noisy spike), or 1 for an empty column."""
    flat = np.asarray(noisy_window, dtype=np.float64).ravel()
    k = min(64, flat.size)
    top = np.partition(flat, flat.size - k)[flat.size - k:]
    s = float(top.mean())
    return s if s > 0.0 else 1.0


def beam_window(center_xy_vox: tuple[int, int], shape: tuple[int, int, int]) -> tuple[slice, slice]:
    """x/y slices of the WINDOW_XY-wide window around a beamlet axis, kept inside the
    grid."""
    nx, ny, _ = shape
    half = WINDOW_XY // 2
    x0 = min(max(int(center_xy_vox[0]) - half, 0), nx - WINDOW_XY)
    y0 = min(max(int(center_xy_vox[1]) - half, 0), ny - WINDOW_XY)
    return slice(x0, x0 + WINDOW_XY), slice(y0, y0 + WINDOW_XY)


# ------------------------------------------------------------- training pairs


@dataclass
class ColumnPair:
    """One beamlet column in its beam-aligned window ([64][64][depth], MeV/g per
    history): the low-history input, the full-statistics target and its sigma, the
    density, and the normalisation scale taken from the input."""

    noisy: np.ndarray
    target: np.ndarray
    target_sigma: np.ndarray
    density: np.ndarray
    scale: float
    phantom: int
    beamlet: int


def _windows_for(grid, source, beamlets):
    spacing = float(grid.spacing[0])
    out = []
    for j in beamlets:
        cx, cy, _ = source.centers[j]
        out.append(beam_window((int(cx / spacing), int(cy / spacing)), tuple(grid.shape)))
    return out


def generate_pairs(n_phantoms: int, beamlets_per_phantom: int, low: int = 10_000,
                   high: int = 1_000_000, seed: int = 20260801, n_batches: int = 10,
                   engine_factory=None, log=None) -> list[ColumnPair]:
    """Engine-generated training pairs on seeded CT-like phantoms (128x128x96 at
    2.5 mm, 100-beamlet fan of 5x5 mm from 100 cm SAD): for each phantom a random
    subset of beamlets at ``low`` and ``high`` histories (independent seeds)."""
    from pyradmc_amd import HipEngine, SpectralBeamletSource, workloads
    from pyradmc_amd.data.analytic import AnalyticCrossSections

    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n_phantoms):
        grid = workloads.ct_like_grid(seed=seed + 7919 * (i + 1))
        xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
        full = workloads.ct_like_beamlets()
        chosen = sorted(rng.choice(full.n_beamlets, size=beamlets_per_phantom, replace=False).tolist())
        src = SpectralBeamletSource(full.spectrum, full.focal_point, [full.centers[j] for j in chosen],
                                    full.width_u, full.width_v)
        eng = engine_factory(grid, xs) if engine_factory else HipEngine(grid, xs)
        lo = eng.run_dij(src, low, n_batches, seed=seed + 2 * i, truncation=0.0)
        hi = eng.run_dij(src, high, n_batches, seed=seed + 2 * i + 1, truncation=0.0)
        eng.close()
        dens = np.asarray(grid.density, dtype=np.float32)
        for k, (wx, wy) in enumerate(_windows_for(grid, src, range(len(chosen)))):
            noisy = lo.column_dense(k)[wx, wy, :].astype(np.float32)
            s = column_scale(noisy)
            pairs.append(ColumnPair(
                noisy=noisy / s, target=hi.column_dense(k)[wx, wy, :].astype(np.float32) / s,
                target_sigma=hi.sigma_dense(k)[wx, wy, :].astype(np.float32) / s,
                density=dens[wx, wy, :], scale=s, phantom=i, beamlet=chosen[k]))
        if log and (i + 1) % 10 == 0:
            log(f"pairs: {i + 1}/{n_phantoms} phantoms")
    return pairs


# ------------------------------------------------------------------ training


def train(model: Conv3dDenoiser, pairs: list[ColumnPair], epochs: int = 10, batch: int = 4,
          lr: float = 1e-3, crop: int = 64, seed: int = 0, log=None) -> list[float]:
    """Adam on random crop x crop x crop sub-windows (random x/y flips); returns the
    mean loss of each epoch."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    history = []
    n = len(pairs)
    depth = pairs[0].noisy.shape[2]
    for epoch in range(epochs):
        order = rng.permutation(n)
        total, count = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = order[b0:b0 + batch]
            xs, ds, ts, ws = [], [], [], []
            for j in idx:
                p = pairs[j]
                z0 = int(rng.integers(0, depth - crop + 1))
                sl = (slice(0, crop), slice(0, crop), slice(z0, z0 + crop))
                a, dd, t, s = p.noisy[sl], p.density[sl], p.target[sl], p.target_sigma[sl]
                if rng.random() < 0.5:
                    a, dd, t, s = a[::-1], dd[::-1], t[::-1], s[::-1]
                if rng.random() < 0.5:
                    a, dd, t, s = a[:, ::-1], dd[:, ::-1], t[:, ::-1], s[:, ::-1]
                xs.append(a)
                ds.append(dd)
                ts.append(t)
                ws.append(1.0 / (s * s + SIGMA_FLOOR * SIGMA_FLOOR))
            x, x_cm = pack_input(to_network_axes(np.stack(xs)), to_network_axes(np.stack(ds)) * 0.5,
                                 model.device)
            target = torch.from_numpy(to_network_axes(np.stack(ts).astype(np.float32))).to(model.device)
            weight = to_network_axes(np.stack(ws).astype(np.float32))
            weight /= weight.mean()
            weight = torch.from_numpy(np.ascontiguousarray(weight)).to(model.device)
            out, cache = model.forward(x, x_cm, keep=True)
            partial, dout, dout_cm = model.loss_and_grad(out, x, target, weight, 1.0 / target.numel())
            model.backward(cache, dout, dout_cm)
            model.adam_step(lr)
            total += float(partial.sum().item()) / target.numel()
            count += 1
        history.append(total / max(count, 1))
        if log:
            log(f"epoch {epoch}: loss {history[-1]:.6g}")
    return history


def predict(model: Conv3dDenoiser, noisy: np.ndarray, density: np.ndarray) -> np.ndarray:
    """Denoised window(s), same scaling as the input: [N][64][64][depth] or a single
    [64][64][depth] window."""
    single = noisy.ndim == 3
    a = noisy[None] if single else noisy
    dd = density[None] if single else density
    x, _ = pack_input(to_network_axes(a.astype(np.float32)), to_network_axes(dd.astype(np.float32)) * 0.5,
                      model.device)
    pred = from_network_axes(model.predict_tensor(x).cpu().numpy())
    return pred[0] if single else pred


def dose_weighted_rms(dose: np.ndarray, reference: np.ndarray) -> float:
    """sqrt(sum R (D - R)^2 / sum R) / max R over R > 0.1 max R."""
    r_max = float(reference.max())
    if r_max <= 0.0:
        return 0.0
    mask = reference > 0.1 * r_max
    r = reference[mask].astype(np.float64)
    d = dose[mask].astype(np.float64)
    return math.sqrt(float(np.sum(r * (d - r) ** 2) / np.sum(r))) / r_max


def evaluate(model: Conv3dDenoiser, pairs: list[ColumnPair], batch: int = 4) -> dict:
    """Dose-weighted RMS error against the full-statistics target, per column, of the
    low-history input and of its denoised prediction."""
    noisy_err, den_err = [], []
    for b0 in range(0, len(pairs), batch):
        chunk = pairs[b0:b0 + batch]
        pred = predict(model, np.stack([p.noisy for p in chunk]), np.stack([p.density for p in chunk]))
        for p, d in zip(chunk, pred):
            noisy_err.append(dose_weighted_rms(p.noisy, p.target))
            den_err.append(dose_weighted_rms(d, p.target))
    return {"noisy": np.array(noisy_err), "denoised": np.array(den_err)}


# ------------------------------------------------------------ denoised Dij


@dataclass
class DenoiseReport:
    """Per-column outcome of a denoised Dij build."""

    disagreement: np.ndarray   # dose-weighted RMS of (estimate A - estimate B) / 2
    fallback: np.ndarray       # bool: column re-simulated with plain Monte Carlo
    seconds: dict


class DenoisedDijBuilder:
    """Dij from low-history Monte Carlo plus the denoiser, with plain Monte Carlo as the
    fallback and the judge.

    Every beamlet is simulated twice with ``low_histories`` (independent seeds); both
    runs are denoised in their beam-aligned windows and averaged. A column whose two
    denoised estimates disagree by more than ``tolerance`` (half the dose-weighted RMS
    of their difference relative to the column maximum: the error of their average,
    were the two errors independent) is re-simulated with ``fallback_histories``. Outside the window the
    column keeps the averaged low-history dose. Sigma of a denoised entry is half the
    difference of the two denoised estimates.
    """

    def __init__(self, engine, model: Conv3dDenoiser, low_histories: int = 10_000,
                 tolerance: float = 0.02, fallback_histories: int = 1_000_000, n_batches: int = 10):
        self.engine = engine
        self.model = model
        self.low = int(low_histories)
        self.tolerance = float(tolerance)
        self.fallback = int(fallback_histories)
        self.n_batches = int(n_batches)

    def build(self, source, seed: int, truncation: float = 1e-4, batch: int = 8):
        import time

        from pyradmc_amd.scoring import DijResult

        t0 = time.perf_counter()
        eng = self.engine
        grid = eng.grid
        shape = tuple(grid.shape)
        a = eng.run_dij(source, self.low, self.n_batches, seed=seed, truncation=0.0)
        b = eng.run_dij(source, self.low, self.n_batches, seed=seed + 1, truncation=0.0)
        t1 = time.perf_counter()
        dens = np.asarray(grid.density, dtype=np.float32)
        n_cols = int(source.n_beamlets)
        windows = _windows_for(grid, source, range(n_cols))
        cols = []
        disagreement = np.zeros(n_cols)
        for c0 in range(0, n_cols, batch):
            idx = list(range(c0, min(c0 + batch, n_cols)))
            ins, scales, dd = [], [], []
            for j in idx:
                wx, wy = windows[j]
                for run in (a, b):
                    win = run.column_dense(j)[wx, wy, :].astype(np.float32)
                    s = column_scale(win)
                    ins.append(win / s)
                    scales.append(s)
                    dd.append(dens[wx, wy, :])
            pred = predict(self.model, np.stack(ins), np.stack(dd))
            for k, j in enumerate(idx):
                pa = np.maximum(pred[2 * k], 0.0) * scales[2 * k]
                pb = np.maximum(pred[2 * k + 1], 0.0) * scales[2 * k + 1]
                mean = 0.5 * (pa + pb)
                # rms(pa - mean) = rms(pa - pb) / 2: the error of the average, were the
                # two estimates' errors independent
                disagreement[j] = dose_weighted_rms(pa, mean) if mean.max() > 0 else 0.0
                full = 0.5 * (a.column_dense(j) + b.column_dense(j))
                sig = 0.5 * np.sqrt(a.sigma_dense(j) ** 2 + b.sigma_dense(j) ** 2)
                wx, wy = windows[j]
                full[wx, wy, :] = mean
                sig[wx, wy, :] = 0.5 * np.abs(pa - pb)
                cols.append((full.ravel(), sig.ravel()))
        t2 = time.perf_counter()
        fallback = disagreement > self.tolerance
        if fallback.any():
            from pyradmc_amd import SpectralBeamletSource

            redo = np.flatnonzero(fallback).tolist()
            sub = SpectralBeamletSource(source.spectrum, source.focal_point, [source.centers[j] for j in redo],
                                        source.width_u, source.width_v, source.u_axis, source.v_axis)
            mc = eng.run_dij(sub, self.fallback, self.n_batches, seed=seed + 2, truncation=0.0)
            for k, j in enumerate(redo):
                cols[j] = (mc.column_dense(k).ravel(), mc.sigma_dense(k).ravel())
        t3 = time.perf_counter()
        # truncation and CSC, as the engine's device compaction does
        indptr = np.zeros(n_cols + 1, dtype=np.int64)
        idx_l, dose_l, sig_l = [], [], []
        for j, (d, s) in enumerate(cols):
            cmax = float(d.max(initial=0.0))
            keep = np.flatnonzero((d > 0.0) & (d >= truncation * cmax))
            idx_l.append(keep.astype(np.int64))
            dose_l.append(d[keep])
            sig_l.append(s[keep])
            indptr[j + 1] = indptr[j] + keep.size
        t4 = time.perf_counter()
        result = DijResult(
            grid_shape=shape, n_beamlets=n_cols, n_histories_per_beamlet=2 * self.low, n_batches=self.n_batches,
            truncation=float(truncation), indptr=indptr, indices=np.concatenate(idx_l),
            dose=np.concatenate(dose_l), sigma=np.concatenate(sig_l),
            energy_emitted=a.energy_emitted + b.energy_emitted,
            energy_deposited=a.energy_deposited + b.energy_deposited,
            energy_escaped=a.energy_escaped + b.energy_escaped,
            energy_unscored=a.energy_unscored + b.energy_unscored,
            correlated=True, scoring_mode="dose_to_medium", provenance=a.provenance, counters={})
        report = DenoiseReport(disagreement=disagreement, fallback=fallback,
                               seconds={"low_mc": t1 - t0, "denoise": t2 - t1, "fallback_mc": t3 - t2,
                                        "assemble": t4 - t3, "total": t4 - t0})
        return result, report


def _main(argv=None) -> int:
    """``python -m pyradmc_amd.denoise train --out FILE [--phantoms N] ...``: generate
    pairs with the engine, train on the GPU, save the weights (NumPy .npz)."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(prog="python -m pyradmc_amd.denoise")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", required=True)
    tr.add_argument("--phantoms", type=int, default=200)
    tr.add_argument("--beamlets", type=int, default=8)
    tr.add_argument("--held-out", type=int, default=20)
    tr.add_argument("--low", type=int, default=10_000)
    tr.add_argument("--high", type=int, default=1_000_000)
    tr.add_argument("--epochs", type=int, default=10)
    tr.add_argument("--batch", type=int, default=4)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--seed", type=int, default=20260801)
    args = ap.parse_args(argv)
    started = time.perf_counter()
    say = lambda m: print(m, flush=True)  # noqa: E731
    pairs = generate_pairs(args.phantoms, args.beamlets, args.low, args.high, seed=args.seed, log=say)
    held = generate_pairs(args.held_out, args.beamlets, args.low, args.high, seed=args.seed + 1_000_003,
                          log=say)
    t_data = time.perf_counter() - started
    model = Conv3dDenoiser(seed=args.seed)
    t0 = time.perf_counter()
    hist = train(model, pairs, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=say)
    _torch().cuda.synchronize()
    t_train = time.perf_counter() - t0
    ev = evaluate(model, held)
    state = model.state_dict()
    np.savez_compressed(args.out, **state, widths=np.array(model.widths))
    print(json.dumps({
        "pairs": len(pairs), "held_out": len(held), "parameters": model.n_parameters,
        "data_s": round(t_data, 1), "train_s": round(t_train, 1), "loss": [round(v, 6) for v in hist],
        "held_out_noisy_rms_median": float(np.median(ev["noisy"])),
        "held_out_denoised_rms_median": float(np.median(ev["denoised"])),
        "held_out_denoised_rms_p90": float(np.percentile(ev["denoised"], 90)),
    }), flush=True)
    return 0


def load(path, device: str = "cuda:0") -> Conv3dDenoiser:
    """A trained denoiser from ``python -m pyradmc_amd.denoise train`` output."""
    with np.load(path) as f:
        widths = tuple(int(v) for v in f["widths"])
        state = {k: f[k] for k in f.files if k != "widths"}
    model = Conv3dDenoiser(widths=widths, device=device)
    model.load_state_dict(state)
    return model


if __name__ == "__main__":
    import sys

    sys.exit(_main())
