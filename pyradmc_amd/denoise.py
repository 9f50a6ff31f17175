"""Learned Dij-column denoiser: a 3D convolutional encoder-decoder on hand-written
BF16 MFMA kernels, trained on column pairs the engine itself generates.

A beamlet column computed with few histories is denoised inside a beam-aligned
window (64 x 64 voxels around the beamlet's central axis, full depth) from two
input channels - the low-history dose and the density - into a prediction of the
full-statistics column. Plain Monte Carlo stays the fallback and the judge: a random
sample of the plan's columns is also simulated with full statistics, the denoised
error is estimated on it without bias, and if that estimate is not within the error
budget every column is simulated with full statistics instead
(``DenoisedDijBuilder``).

The reference has no learned component; this module extends its Dij builder. The
network's forward and backward passes run entirely in this build's HIP kernels
(``include/hipdn.h``); PyTorch only allocates the device buffers and supplies the
stream.

Network: four levels of two 3x3x3 convolutions (widths 16, 32, 64, 128 by default,
about 1.3 M parameters), 2x2x2 average pooling down, nearest upsampling up,
additive skips, leaky ReLU (slope 0.1), and a residual head: prediction = noisy
input + head output. Loss: squared error weighted by 1 / (sigma_target^2 + floor^2)
(the target's own Monte Carlo variance, floored), or by the target dose (the weight
of the dose-weighted error that judges a column), averaged over voxels.
"""

from __future__ import annotations

import ctypes
import math
import sys
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

# The denoiser's buffers are torch tensors and its kernels run on torch's stream, so
# the process must use one HIP runtime: torch's. torch is therefore imported here, at
# module import, before the native library can be loaded by anything else.
try:
    import torch as _torch_module
except ImportError:  # pragma: no cover - the denoiser then refuses to run
    _torch_module = None

from pyradmc_amd import _native

SLOPE = 0.1
CIN_PAD = 16          # input channels padded to the MFMA k-group multiple
HEAD_OUT = 16         # head output channels (channel 0 is the prediction)
WINDOW_XY = 64        # beam-aligned window, voxels
SIGMA_FLOOR = 0.02    # loss weight floor, in units of the column scale
SCRATCH_FLOATS = 16 * 1024 * 1024  # HDN_SCRATCH_FLOATS: fixed array of hdn_conv3d_wgrad_noalloc

_c_void = ctypes.c_void_p
_c_ll = ctypes.c_longlong


def _lib():
    h = _native.lib()
    if getattr(h, "_hdn_bound", False):
        return h
    i, f = ctypes.c_int, ctypes.c_float
    sig = {
        "hdn_conv3d_fwd": [_c_void, _c_void, _c_void, _c_void, _c_void, i, i, i, i, i, i, f, _c_void],
        "hdn_conv3d_fwd_src": [_c_void, _c_void, i, _c_void, _c_void, _c_void, _c_void, i, i, i, i, i, i, f,
                               _c_void],
        "hdn_conv3d_wgrad": [_c_void, _c_void, _c_void, i, i, i, i, i, i, i, _c_void],
        "hdn_conv3d_wgrad_noalloc": [_c_void, _c_void, _c_void, i, i, i, i, i, i, i, _c_void],
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
    torch = _torch_module
    if torch is None:
        raise RuntimeError("the denoiser needs PyTorch (ROCm) for its device buffers")
    if not torch.cuda.is_available():
        hint = ""
        if _native._lib is not None and "pyradmc_amd.denoise" in sys.modules:
            hint = (" - the native library was loaded before torch initialised its HIP runtime;"
                    " import pyradmc_amd.denoise (or torch) before creating a HipEngine")
        raise RuntimeError("the denoiser needs a HIP device (torch.cuda is unavailable)" + hint)
    return torch


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def _call(name: str, *args) -> None:
    rc = getattr(_lib(), name)(*args)
    if rc == -3:
        raise ValueError(f"{name}: the call needs more than the library's fixed device array "
                         f"of {SCRATCH_FLOATS} floats (status {rc})")
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


def forward_flops(shape: tuple[int, int, int, int], widths: tuple[int, int, int, int] = (16, 32, 64, 128)) -> int:
    """Algorithmic FLOPs of one forward pass over an [N][D][H][W] batch: 2 * 27 * C_in *
    C_out per output voxel of every convolution, with the real input (2) and output (1)
    channel counts rather than their MFMA padding."""
    n, d, h, w = shape
    level = {"0": 0, "1": 1, "2": 2, "3": 3}
    total = 0
    for name, cin, cout in layer_specs(2, widths):
        lvl = 0 if name == "head" else level[name[1]]
        voxels = n * (d >> lvl) * (h >> lvl) * (w >> lvl)
        total += 2 * 27 * cin * (1 if name == "head" else cout) * voxels
    return total


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

    def _conv(self, lay: _Layer, x, dims, activate: bool, want_cm: bool, x2=None, src: int = 0):
        """One convolution at output dims ``dims``; ``src`` 1 reads the nearest 2x
        upsampling of ``x`` (half the dims), 2 reads ``x + x2``, 3 reads ``x`` whose
        channels beyond the first 8 are zero (include/hipdn.h)."""
        torch = _torch()
        n, d, h, w = dims
        y = torch.empty((n, d, h, w, lay.cout), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((lay.cout, n * d * h * w), dtype=torch.bfloat16, device=x.device) if want_cm else None
        _call("hdn_conv3d_fwd_src", _ptr(x), _ptr(x2), int(src), _ptr(lay.wf), _ptr(lay.b), _ptr(y), _ptr(y_cm),
              n, d, h, w, lay.cin, lay.cout, SLOPE if activate else 1.0, self._stream())
        return y, y_cm

    def _pool(self, x, dims, C, want_cm: bool = True):
        torch = _torch()
        n, d, h, w = dims
        y = torch.empty((n, d // 2, h // 2, w // 2, C), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((C, y.numel() // C), dtype=torch.bfloat16, device=x.device) if want_cm else None
        _call("hdn_pool_fwd", _ptr(x), _ptr(y), _ptr(y_cm), n, d // 2, h // 2, w // 2, C, self._stream())
        return y, y_cm

    def _up(self, x, dims_fine, C, want_cm: bool = True):
        torch = _torch()
        n, d, h, w = dims_fine
        y = torch.empty((n, d, h, w, C), dtype=torch.bfloat16, device=x.device)
        y_cm = torch.empty((C, y.numel() // C), dtype=torch.bfloat16, device=x.device) if want_cm else None
        _call("hdn_up_fwd", _ptr(x), _ptr(y), _ptr(y_cm), n, d, h, w, C, self._stream())
        return y, y_cm

    def _add(self, a, b, C, want_cm: bool = True):
        torch = _torch()
        y = torch.empty_like(a)
        y_cm = torch.empty((C, a.numel() // C), dtype=torch.bfloat16, device=a.device) if want_cm else None
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
            # The first layer's input has 2 real channels of 16: its reduction skips the
            # zero channel group (src 3).
            y, y_cm = self._conv(L[name], inp, dims[lvl], activate, want_cm=keep, src=3 if name == "e0a" else 0)
            if keep:
                cache[name] = (inp, inp_cm, y, y_cm, lvl)
            return y, y_cm

        if not keep:
            return self._infer(x, dims), cache
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

    def _infer(self, x, dims):
        """The forward pass without anything kept for the backward: no channel-major
        copies, and each upsampling or skip addition read by the next convolution
        instead of stored (identical values, so identical outputs)."""
        L = self.layers
        c0, c1, c2, _ = self.widths

        def conv(name, inp, lvl, activate=True, x2=None, src=0):
            return self._conv(L[name], inp, dims[lvl], activate, False, x2=x2, src=src)[0]

        s0 = conv("e0b", conv("e0a", x, 0, src=3), 0)
        s1 = conv("e1b", conv("e1a", self._pool(s0, dims[0], c0, False)[0], 1), 1)
        s2 = conv("e2b", conv("e2a", self._pool(s1, dims[1], c1, False)[0], 2), 2)
        t = conv("e3b", conv("e3a", self._pool(s2, dims[2], c2, False)[0], 3), 3)
        t = conv("d2b", conv("d2a", t, 2, src=1), 2, x2=s2, src=2)
        t = conv("d1b", conv("d1a", t, 1, src=1), 1, x2=s1, src=2)
        t = conv("d0b", conv("d0a", t, 0, src=1), 0, x2=s0, src=2)
        return conv("head", t, 0, activate=False)

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
    """[N][D][H][W] scaled noisy dose and halved density (network axes, host arrays)
    -> padded bf16 nd tensor and its channel-major copy."""
    torch = _torch()
    a = torch.from_numpy(np.ascontiguousarray(noisy, dtype=np.float32)).to(device)
    d = torch.from_numpy(np.ascontiguousarray(density, dtype=np.float32)).to(device)
    return _pack_device(a, 2.0 * d)


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


def _pack_device(noisy, density):
    """[N][D][H][W] float32 device tensors (network axes) -> padded bf16 nd input and
    its channel-major copy (device-side plumbing; the density channel is halved)."""
    torch = _torch()
    x = torch.zeros(tuple(noisy.shape) + (CIN_PAD,), dtype=torch.bfloat16, device=noisy.device)
    x[..., 0] = noisy.to(torch.bfloat16)
    x[..., 1] = (density * 0.5).to(torch.bfloat16)
    return x, x.reshape(-1, CIN_PAD).t().contiguous()


def _stack_device(pairs: list[ColumnPair], device: str, weighting: str = "sigma"):
    """Device tensors of the pairs and the per-voxel loss weight: ``"sigma"`` weighs
    by 1/(sigma_target^2 + floor^2); ``"dose"`` by the target dose plus 1 percent of
    the column scale (the dose-weighted error that judges a column, with a floor that
    keeps the low-dose tail in the loss)."""
    torch = _torch()
    if weighting not in ("sigma", "dose"):
        raise ValueError(f"weighting must be 'sigma' or 'dose', got {weighting!r}")

    def up(name):
        return torch.from_numpy(np.ascontiguousarray(np.stack([getattr(p, name) for p in pairs]),
                                                     dtype=np.float32)).to(device)

    noisy, dens, target = up("noisy"), up("density"), up("target")
    if weighting == "dose":
        weight = target.clamp(min=0.0) + 0.01
    else:
        sig = up("target_sigma")
        weight = 1.0 / (sig * sig + SIGMA_FLOOR * SIGMA_FLOOR)
        del sig
    return noisy, dens, target, weight


def train(model: Conv3dDenoiser, pairs: list[ColumnPair], epochs: int = 10, batch: int = 8,
          lr: float = 1e-3, lr_final: float = 3e-5, crop: int = 64, seed: int = 0, log=None,
          weighting: str = "sigma", on_epoch=None) -> list[float]:
    """Adam with cosine learning-rate decay on random crop^3 sub-windows (random
    x/y flips); the pairs are held on the device for the whole run. Returns the mean
    loss of each epoch; ``on_epoch(epoch)`` runs after each epoch."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    noisy, dens, target, weight = _stack_device(pairs, model.device, weighting)
    n = len(pairs)
    depth = noisy.shape[3]
    steps_per_epoch = max(n // batch, 1)
    total = epochs * steps_per_epoch
    history = []
    step = 0
    for epoch in range(epochs):
        order = rng.permutation(n)
        loss_sum = torch.zeros((), dtype=torch.float64, device=model.device)
        for b0 in range(0, steps_per_epoch * batch, batch):
            idx = order[b0:b0 + batch]
            crops = []
            for j in idx:
                z0 = int(rng.integers(0, depth - crop + 1))
                fx, fy = rng.random() < 0.5, rng.random() < 0.5
                parts = []
                for arr in (noisy, dens, target, weight):
                    c = arr[int(j), :crop, :crop, z0:z0 + crop]
                    if fx:
                        c = torch.flip(c, (0,))
                    if fy:
                        c = torch.flip(c, (1,))
                    parts.append(c)
                crops.append(parts)
            # [N][x][y][z] -> network axes [N][z][x][y]
            a, d, t, w = (torch.stack([c[k] for c in crops]).permute(0, 3, 1, 2).contiguous() for k in range(4))
            w = w / w.mean()
            x, x_cm = _pack_device(a, d)
            out, cache = model.forward(x, x_cm, keep=True)
            partial, dout, dout_cm = model.loss_and_grad(out, x, t, w, 1.0 / t.numel())
            model.backward(cache, dout, dout_cm)
            rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / max(total, 1)))
            model.adam_step(rate)
            loss_sum += partial.sum().double() / t.numel()
            step += 1
        history.append(float(loss_sum.item()) / steps_per_epoch)
        if log:
            log(f"epoch {epoch}: loss {history[-1]:.6g} lr {rate:.3g}")
        if on_epoch:
            on_epoch(epoch)
    return history


def predict(model: Conv3dDenoiser, noisy: np.ndarray, density: np.ndarray) -> np.ndarray:
    """Denoised window(s), same scaling as the input: [N][64][64][depth] or a single
    [64][64][depth] window."""
    torch = _torch()
    single = noisy.ndim == 3
    a = noisy[None] if single else noisy
    dd = density[None] if single else density
    at = torch.from_numpy(to_network_axes(np.asarray(a, dtype=np.float32))).to(model.device)
    dt = torch.from_numpy(to_network_axes(np.asarray(dd, dtype=np.float32))).to(model.device)
    x, _ = _pack_device(at, dt)
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


def evaluate(model: Conv3dDenoiser, pairs: list[ColumnPair], batch: int = 32) -> dict:
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
    """Outcome of a denoised Dij build."""

    audited: np.ndarray          # bool per column: re-simulated with plain Monte Carlo as the judge's sample
    estimated_error: np.ndarray  # per audited column, the denoised column's estimated error (NaN elsewhere)
    audit_error: float           # RMS of the audited estimates: the judge's verdict on the network here
    fallback: np.ndarray         # bool per column: the column in the result is plain Monte Carlo
    seconds: dict


def _window_entries(dij, column: int, window, shape):
    """The entries of one CSC column inside a beam-aligned window: local (x, y, z)
    coordinates and their dose and sigma (float64), read from the column's entries
    only."""
    wx, wy = window
    lo, hi = int(dij.indptr[column]), int(dij.indptr[column + 1])
    idx = np.asarray(dij.indices[lo:hi], dtype=np.int64)
    _, ny, nz = shape
    x, rest = np.divmod(idx, ny * nz)
    y, z = np.divmod(rest, nz)
    inside = (x >= wx.start) & (x < wx.stop) & (y >= wy.start) & (y < wy.stop)
    return (x[inside] - wx.start, y[inside] - wy.start, z[inside],
            np.asarray(dij.dose[lo:hi], dtype=np.float64)[inside],
            np.asarray(dij.sigma[lo:hi], dtype=np.float64)[inside])


class DenoisedDijBuilder:
    """Dij from low-history Monte Carlo plus the denoiser, with plain Monte Carlo as the
    fallback and the judge.

    Every beamlet is simulated once with ``low_histories`` (one batch: the network reads
    only the mean) and denoised in its beam-aligned window. Plain Monte Carlo judges the
    network on this plan first: a seeded random sample of ``audit_fraction`` of the
    columns (at least one) is simulated both ways, plain at ``fallback_histories``, and
    for each sampled column, with D the denoised window and M, s_M the plain run's mean
    and batch sigma,

        e^2 = sum_v D_v ((D_v - M_v)^2 - s_M,v^2) / sum_v D_v / max(D)^2

    over voxels above 10 percent of max(D) estimates the squared dose-weighted RMS error
    of D against the truth without bias (M is independent of D). A network bias shows in
    it like noise does. If the RMS of the sample's estimates is not below ``tolerance``
    the network is not trusted for this plan and the result is the plain Monte Carlo
    Dij of every column (bit for bit ``run_dij(source, fallback_histories, n_batches,
    seed + 2)``). Otherwise the sampled columns keep their plain Monte Carlo entries and
    every other column its denoised window; dose outside a window is dropped, and a
    denoised entry's sigma is the audit's RMS estimate times its column maximum.
    """

    def __init__(self, engine, model: Conv3dDenoiser, low_histories: int = 20_000,
                 tolerance: float = 0.02, fallback_histories: int = 1_000_000, n_batches: int = 10,
                 audit_fraction: float = 0.1):
        self.engine = engine
        self.model = model
        self.low = int(low_histories)
        self.tolerance = float(tolerance)
        self.fallback = int(fallback_histories)
        self.n_batches = int(n_batches)
        self.audit_fraction = float(audit_fraction)
        if self.low < 1:
            raise ValueError("the low-history run needs at least one history per beamlet")
        if not 0.0 < self.audit_fraction <= 1.0:
            raise ValueError(f"audit_fraction must lie in (0, 1], got {audit_fraction}")

    def _denoise_group(self, source, group: range, windows, seed: int, books: dict, dens, batch: int):
        """Low-history means of the group's columns (on the device, one batch), their
        windows gathered and denoised: (denoised windows [n][z][x][y], origins, the
        fold's deposited quanta, seconds of Monte Carlo, seconds of inference)."""
        import time

        torch = _torch()
        t0 = time.perf_counter()
        eng, model = self.engine, self.model
        depth = eng.grid.shape[2]
        ng = len(group)
        org = np.ascontiguousarray([[windows[j][0].start, windows[j][1].start] for j in group], dtype=np.int32)
        counts = np.full(int(source.n_beamlets), self.low, dtype=np.int64)
        _, _, quanta = eng._dij_group(source, group, counts, 1, seed, True, eng._default_scoring(), books)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        wins = torch.empty((ng, depth, WINDOW_XY, WINDOW_XY), dtype=torch.float32, device=model.device)
        eng._ctx.call("hmc_gather_windows", _native.ptr(org, ctypes.c_int32), ng, WINDOW_XY, WINDOW_XY,
                      wins.data_ptr())
        scale = torch.topk(wins.reshape(ng, -1), 64, dim=1).values.mean(dim=1).clamp_min(1e-30)
        noisy = wins / scale[:, None, None, None]
        dwin = torch.stack([dens[:, windows[j][0], windows[j][1]] for j in group])  # [n][z][x][y]
        preds = []
        for b0 in range(0, ng, batch):
            x, _ = _pack_device(noisy[b0:b0 + batch].contiguous(), dwin[b0:b0 + batch].contiguous())
            preds.append(model.predict_tensor(x))
        pred = torch.cat(preds).reshape(ng, depth, WINDOW_XY, WINDOW_XY).clamp_min(0.0) * scale[:, None, None, None]
        torch.cuda.synchronize()
        return pred, org, quanta, t1 - t0, time.perf_counter() - t1

    def build(self, source, seed: int, truncation: float = 1e-4, batch: int = 64,
              group_size: int | None = None, value_dtype=np.float64):
        """(DijResult, DenoiseReport) for the source's beamlets. The judge's sample runs
        first; then columns go through the engine in groups: each group's means stay on
        the device, the windows are gathered straight into the network's input buffers,
        and the denoised windows (and the sampled columns' plain entries) are written
        back for the engine's own truncation and CSC compaction, which return the
        columns in order. ``value_dtype`` is the CSC dose and sigma type, float64 or
        float32, as in ``HipEngine.run_dij``."""
        import time

        from pyradmc_amd import SpectralBeamletSource
        from pyradmc_amd.engine import _joined, host_array
        from pyradmc_amd.scoring import DijResult

        torch = _torch()
        value_dtype = np.dtype(value_dtype)
        if value_dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError(f"value_dtype must be float64 or float32, got {value_dtype}")
        t0 = time.perf_counter()
        eng, model = self.engine, self.model
        grid = eng.grid
        shape = tuple(grid.shape)
        n_cols = int(source.n_beamlets)
        nb = self.n_batches
        windows = _windows_for(grid, source, range(n_cols))
        scoring = eng._default_scoring()
        eng._prepare(source)  # default cuts and physics, as run_dij's defaults
        dens = torch.from_numpy(np.ascontiguousarray(np.transpose(np.asarray(grid.density, np.float32),
                                                                  (2, 0, 1)))).to(model.device)  # [z][x][y]
        seconds = {"audit_mc": 0.0, "audit": 0.0, "low_mc": 0.0, "denoise": 0.0, "fallback_mc": 0.0,
                   "assemble": 0.0}

        def subset(cols):
            return SpectralBeamletSource(source.spectrum, source.focal_point, [source.centers[j] for j in cols],
                                         source.width_u, source.width_v, source.u_axis, source.v_axis)

        # The judge. Correlated keys: a column's bits do not depend on which other columns
        # run with it, so the sampled columns' plain runs are those of the full plain Dij.
        rng = np.random.default_rng([int(seed), 3])
        n_audit = max(1, int(math.ceil(self.audit_fraction * n_cols)))
        audit_cols = sorted(rng.choice(n_cols, size=n_audit, replace=False).tolist())
        audited = np.zeros(n_cols, dtype=bool)
        audited[audit_cols] = True
        audit_books: dict = {}
        plain = eng.run_dij(subset(audit_cols), self.fallback, nb, seed=seed + 2, truncation=truncation)
        self._merge_books(audit_books, plain)
        seconds["audit_mc"] = time.perf_counter() - t0
        sample = subset(audit_cols)
        sample_windows = [windows[j] for j in audit_cols]
        pred, _, _, _, _ = self._denoise_group(sample, range(n_audit), sample_windows, seed, {}, dens, batch)
        # sum_v w (D - M)^2 - w s^2 = sum w D^2 - 2 sum w D M + sum w (M^2 - s^2): the first
        # sum over the dense window (on the device), the others over M's entries only
        flat = pred.reshape(n_audit, -1).double()
        dmax_t = flat.max(dim=1).values
        w_t = torch.where(flat > 0.1 * dmax_t[:, None], flat, torch.zeros_like(flat))
        s_w = w_t.sum(dim=1).cpu().numpy()
        s_wdd = (w_t * flat * flat).sum(dim=1).cpu().numpy()
        dmax = dmax_t.cpu().numpy()
        sample_pred = pred.cpu().numpy()  # [n][z][x][y], float32
        del pred, flat, w_t
        estimate = np.full(n_cols, np.nan)
        for k, j in enumerate(audit_cols):
            if not dmax[k] > 0.0:
                estimate[j] = np.inf  # nothing to judge: the network is not trusted with it
                continue
            x, y, z, m, s_m = _window_entries(plain, k, windows[j], shape)
            d = sample_pred[k][z, x, y].astype(np.float64)
            w = np.where(d > 0.1 * dmax[k], d, 0.0)
            e2 = (s_wdd[k] - 2.0 * float(np.sum(w * d * m)) + float(np.sum(w * (m * m - s_m * s_m)))) / s_w[k]
            estimate[j] = math.sqrt(max(e2, 0.0)) / dmax[k]
        audit_error = float(np.sqrt(np.mean(estimate[audited] ** 2)))
        seconds["audit"] = time.perf_counter() - t0

        if not audit_error < self.tolerance:
            # not trusted here: plain Monte Carlo for every column
            te = time.perf_counter()
            result = eng.run_dij(source, self.fallback, nb, seed=seed + 2, truncation=truncation,
                                 value_dtype=value_dtype)
            seconds["fallback_mc"] = time.perf_counter() - te
            seconds["total"] = time.perf_counter() - t0
            return result, DenoiseReport(audited=audited, estimated_error=estimate, audit_error=audit_error,
                                         fallback=np.ones(n_cols, dtype=bool), seconds=seconds)

        if group_size is None:
            free = _native.device_free_bytes(eng._index)
            per_col = 8 * scoring.n_voxels * 4  # one map slot, the sums, and the window buffers
            group_size = max(1, min(n_cols, int(0.25 * free) // max(per_col, 1)))
        books: dict = {}
        deposited_quanta = 0
        col_counts, col_idx, col_dose, col_sigma = [], [], [], []
        for g0 in range(0, n_cols, group_size):
            group = range(g0, min(g0 + group_size, n_cols))
            ng = len(group)
            pred, org, quanta, t_mc, t_net = self._denoise_group(source, group, windows, seed, books, dens, batch)
            seconds["low_mc"] += t_mc
            seconds["denoise"] += t_net
            deposited_quanta += quanta
            local_audited = [j - g0 for j in group if audited[j]]
            if local_audited:
                pred[local_audited] = 0.0
            col_max = pred.reshape(ng, -1).max(dim=1).values
            sigma = (audit_error * col_max)[:, None, None, None].expand_as(pred).contiguous()
            torch.cuda.synchronize()
            tc = time.perf_counter()
            eng._ctx.call("hmc_scatter_windows", _native.ptr(org, ctypes.c_int32), ng, WINDOW_XY, WINDOW_XY,
                          pred.contiguous().data_ptr(), sigma.data_ptr())
            for j in group:
                if audited[j]:
                    k = audit_cols.index(j)
                    lo, hi = int(plain.indptr[k]), int(plain.indptr[k + 1])
                    eng._ctx.call("hmc_set_column_entries", j - g0, hi - lo,
                                  _native.ptr(np.ascontiguousarray(plain.indices[lo:hi], dtype=np.int64),
                                              _native.c_int64),
                                  _native.ptr(np.ascontiguousarray(plain.dose[lo:hi], dtype=np.float64),
                                              ctypes.c_double),
                                  _native.ptr(np.ascontiguousarray(plain.sigma[lo:hi], dtype=np.float64),
                                              ctypes.c_double))
            kept = np.zeros(ng, dtype=np.int64)
            total = _native.c_int64()
            eng._ctx.call("hmc_truncate", float(truncation), _native.ptr(kept, _native.c_int64),
                          ctypes.byref(total))
            mtot = int(total.value)
            idx = host_array(mtot, np.int64)
            dose = host_array(mtot, value_dtype)
            sig = host_array(mtot, value_dtype)
            if mtot:
                narrow = value_dtype == np.dtype(np.float32)
                vt = ctypes.c_float if narrow else ctypes.c_double
                eng._ctx.call("hmc_fetch_csc_f32" if narrow else "hmc_fetch_csc", _native.ptr(idx, _native.c_int64),
                              _native.ptr(dose, vt), _native.ptr(sig, vt))
            col_counts.append(kept)
            col_idx.append(idx)
            col_dose.append(dose)
            col_sigma.append(sig)
            seconds["assemble"] += time.perf_counter() - tc
        kept_all = np.concatenate(col_counts)
        indptr = np.zeros(n_cols + 1, dtype=np.int64)
        np.cumsum(kept_all, out=indptr[1:])
        indices, dose, sigma = _joined(col_idx, np.int64), _joined(col_dose, value_dtype), _joined(col_sigma, value_dtype)
        for name in ("emitted", "escaped", "unscored"):
            books[name] = books.get(name, 0) + audit_books.get(name, 0)
        seconds["total"] = time.perf_counter() - t0
        q = 1.0e-9  # ledger quanta -> MeV
        result = DijResult(
            grid_shape=shape, n_beamlets=n_cols, n_histories_per_beamlet=self.low, n_batches=1,
            truncation=float(truncation), indptr=indptr, indices=indices, dose=dose, sigma=sigma,
            energy_emitted=float(books.get("emitted", 0)) * q,
            energy_deposited=float(deposited_quanta + audit_books.get("fallback_deposited", 0)) * q,
            energy_escaped=float(books.get("escaped", 0)) * q, energy_unscored=float(books.get("unscored", 0)) * q,
            correlated=True, scoring_mode="dose_to_medium", provenance=None, counters=books)
        return result, DenoiseReport(audited=audited, estimated_error=estimate, audit_error=audit_error,
                                     fallback=audited.copy(), seconds=seconds)

    @staticmethod
    def _merge_books(books: dict, dij) -> None:
        for name in ("emitted", "escaped", "unscored"):
            books[name] = books.get(name, 0) + int(round(getattr(dij, "energy_" + name) / 1.0e-9))
        books["fallback_deposited"] = books.get("fallback_deposited", 0) + int(round(dij.energy_deposited / 1.0e-9))


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
    tr.add_argument("--batch", type=int, default=8)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--seed", type=int, default=20260801)
    tr.add_argument("--weighting", default="sigma",
                    help="loss weighting: sigma, dose, or both (two models from one set of pairs; "
                         "the second is saved with '_dose' before the suffix)")
    tr.add_argument("--eval-every", type=int, default=5, help="held-out evaluation every N epochs")
    args = ap.parse_args(argv)
    started = time.perf_counter()
    say = lambda m: print(m, flush=True)  # noqa: E731
    pairs = generate_pairs(args.phantoms, args.beamlets, args.low, args.high, seed=args.seed, log=say)
    held = generate_pairs(args.held_out, args.beamlets, args.low, args.high, seed=args.seed + 1_000_003,
                          log=say)
    t_data = time.perf_counter() - started
    weightings = ("sigma", "dose") if args.weighting == "both" else (args.weighting,)
    for k, weighting in enumerate(weightings):
        out = Path(args.out)
        if k:
            out = out.with_name(out.stem + "_" + weighting + out.suffix)
        model = Conv3dDenoiser(seed=args.seed)

        def progress(epoch, model=model):
            if args.eval_every and (epoch + 1) % args.eval_every == 0:
                ev = evaluate(model, held)
                say(f"epoch {epoch}: held-out denoised rms median {np.median(ev['denoised']):.5f} "
                    f"p90 {np.percentile(ev['denoised'], 90):.5f} (noisy {np.median(ev['noisy']):.5f})")

        t0 = time.perf_counter()
        hist = train(model, pairs, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=say,
                     weighting=weighting, on_epoch=progress)
        _torch().cuda.synchronize()
        t_train = time.perf_counter() - t0
        ev = evaluate(model, held)
        np.savez_compressed(out, **model.state_dict(), widths=np.array(model.widths))
        print(json.dumps({
            "out": str(out), "weighting": weighting,
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
