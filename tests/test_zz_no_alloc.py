"""The allocation-free weight gradient (hdn_conv3d_wgrad_noalloc) against the allocating
pair it stands for: hdn_conv3d_wgrad into caller partials, then hdn_reduce_splits.

Both run the same kernels in the same split order, so the gradients must agree bit for
bit. Two calls in a row, the second on a smaller input with fewer splits, check that the
shared device array carries nothing from one call into the next; a call beyond the
array's size must be refused before anything runs.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import conv_ref as cr

pytestmark = pytest.mark.gpu


def _bf16_input(rng, shape):
    import torch

    return torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(torch.bfloat16)


def _pair(dn, rng, shape, cin, cout, splits):
    """Channel-major bf16 input and output gradient on the device, and the gradient the
    allocating entry points give."""
    import torch

    n, d, h, w = shape
    P = n * d * h * w
    x = _bf16_input(rng, (cin, P)).to("cuda:0")
    dz = _bf16_input(rng, (cout, P)).to("cuda:0")
    part = torch.empty((splits, 27, cout, cin), dtype=torch.float32, device="cuda:0")
    want = torch.empty((27, cout, cin), dtype=torch.float32, device="cuda:0")
    dn._call("hdn_conv3d_wgrad", x.data_ptr(), dz.data_ptr(), part.data_ptr(), splits, n, d, h, w,
             cin, cout, None)
    dn._call("hdn_reduce_splits", part.data_ptr(), want.data_ptr(), want.numel(), splits, None)
    return x, dz, want


def test_noalloc_weight_gradient_matches_allocating_entry_points():
    import torch

    from pyradmc_amd import denoise as dn

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    dn._lib()
    rng = np.random.default_rng(20261203)
    cases = [((2, 8, 8, 64), 32, 48, 12),   # first call: larger input, more splits
             ((1, 4, 4, 16), 16, 16, 3)]    # second: smaller, fewer splits
    for shape, cin, cout, splits in cases:
        n, d, h, w = shape
        print(f"noalloc wgrad case shape={shape} cin={cin} cout={cout} splits={splits}", flush=True)
        x, dz, want = _pair(dn, rng, shape, cin, cout, splits)
        got = torch.full((27, cout, cin), float("nan"), dtype=torch.float32, device="cuda:0")
        dn._call("hdn_conv3d_wgrad_noalloc", x.data_ptr(), dz.data_ptr(), got.data_ptr(), splits,
                 n, d, h, w, cin, cout, None)
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), want.cpu().numpy()
        assert np.isfinite(g).all()
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32))
        # and both are the convolution's weight gradient (oracle in float64)
        xs = x.to(torch.float64).cpu().numpy().T.reshape(shape + (cin,))
        dzs = dz.to(torch.float64).cpu().numpy().T.reshape(shape + (cout,))
        ref = cr.conv_wgrad(xs, dzs)
        np.testing.assert_allclose(g, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


def test_noalloc_weight_gradient_refuses_calls_beyond_its_array():
    import torch

    from pyradmc_amd import denoise as dn

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    dn._lib()
    n, d, h, w, cin, cout = 1, 32, 32, 64, 128, 128
    splits = dn.SCRATCH_FLOATS // (27 * cin * cout) + 1   # one split too many
    x = torch.zeros((cin, n * d * h * w), dtype=torch.bfloat16, device="cuda:0")
    dz = torch.zeros((cout, n * d * h * w), dtype=torch.bfloat16, device="cuda:0")
    grad = torch.full((27, cout, cin), 7.0, dtype=torch.float32, device="cuda:0")
    with pytest.raises(ValueError, match="fixed device array"):
        dn._call("hdn_conv3d_wgrad_noalloc", x.data_ptr(), dz.data_ptr(), grad.data_ptr(), splits,
                 n, d, h, w, cin, cout, None)
    torch.cuda.synchronize()
    assert bool((grad == 7.0).all())
