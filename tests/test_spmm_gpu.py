"""The device planning product D = Dij @ W (include/hipsp.h) against the oracle's
dose_for_weights and the reference's golden products.

In float64 the product is the reference's sum bit for bit (ascending beamlet order,
product rounded before the add), so those comparisons are exact. In float32 each
output carries at most (n + 3) * 2^-24 of the sum of |entry * weight| over its n
entries (weights rounded to float32, one rounding per fused multiply-add).
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import mc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spmm():
    import torch  # one HIP runtime for the process: torch's, loaded before the library

    from pyradmc_amd import _native, spmm

    if not torch.cuda.is_available():
        pytest.fail("no HIP device")
    _native.lib()
    return spmm


def _matrix(indptr, indices, dose, n_rows, shape=None, dtype=None):
    dose = np.asarray(dose, dtype=dtype or np.asarray(dose).dtype if len(dose) else np.float64)
    return SimpleNamespace(indptr=np.asarray(indptr, dtype=np.int64), indices=np.asarray(indices, dtype=np.int64),
                           dose=dose, n_beamlets=len(indptr) - 1, n_voxels=n_rows, grid_shape=shape or (n_rows,))


def _beamlet_like(rng, shape=(20, 22, 36), n_cols=150, value_dtype=np.float32):
    """Columns over box-shaped voxel sets (as beamlets cover the grid), random
    positive values over four decades, plus a few scattered entries and empty
    columns."""
    nx, ny, nz = shape
    indptr, indices, values = [0], [], []
    for c in range(n_cols):
        if c % 37 == 5:
            rows = np.zeros(0, dtype=np.int64)
        else:
            x0, y0 = rng.integers(0, nx - 4), rng.integers(0, ny - 4)
            w = int(rng.integers(2, 6))
            xs, ys, zs = np.meshgrid(np.arange(x0, min(nx, x0 + w)), np.arange(y0, min(ny, y0 + w)),
                                     np.arange(int(rng.integers(0, 10)), nz), indexing="ij")
            rows = ((xs * ny + ys) * nz + zs).ravel()
            extra = rng.choice(nx * ny * nz, size=int(rng.integers(0, 20)), replace=False)
            rows = np.unique(np.concatenate([rows, extra]))
            rows = rng.permutation(rows)
        indices.append(rows)
        values.append(rng.random(rows.size) * 10.0 ** rng.uniform(-2, 2, rows.size))
        indptr.append(indptr[-1] + rows.size)
    n_rows = nx * ny * nz
    return _matrix(indptr, np.concatenate(indices), np.concatenate(values).astype(value_dtype), n_rows, shape)


def _expected(m, w):
    return mc.dose_for_weights(m.indptr, m.indices, m.dose, w, m.n_voxels)


def test_golden_products_bitwise(spmm, golden_dir):
    g = np.load(golden_dir / "ref_dose_for_weights.npz")
    fan = np.load(golden_dir / "ref_dij_fan.npz")
    dev = spmm.DeviceDij(_matrix(fan["indptr"], fan["indices"], fan["dose"], 16 * 16 * 24, (16, 16, 24)))
    assert np.array_equal(dev.apply(g["fan_weights"].T), g["fan_dose"])
    for k in range(g["fan_weights"].shape[0]):
        grid = dev.dose_for_weights(g["fan_weights"][k])
        assert grid.shape == (16, 16, 24)
        assert np.array_equal(grid.ravel(), g["fan_dose"][:, k])
    n = int(g["syn_n_rows"])
    syn = spmm.DeviceDij(_matrix(g["syn_indptr"], g["syn_indices"], g["syn_dose"], n))
    w = g["syn_weights"].T  # [beamlets][4]
    assert np.array_equal(syn.apply(w), g["syn_out"])
    for n_w in (9, 70, 300):  # the wide kernels (64, 128 and 256 vectors per pass)
        reps = -(-n_w // 4)
        got = syn.apply(np.tile(w, (1, reps))[:, :n_w])
        assert np.array_equal(got, np.tile(g["syn_out"], (1, reps))[:, :n_w]), n_w


def test_csr_form_is_sorted_with_repeats_in_csc_order(spmm, golden_dir):
    g = np.load(golden_dir / "ref_dose_for_weights.npz")
    n = int(g["syn_n_rows"])
    dev = spmm.DeviceDij(_matrix(g["syn_indptr"], g["syn_indices"], g["syn_dose"], n))
    row_ptr, col, val = dev.csr()
    cols = np.repeat(np.arange(len(g["syn_indptr"]) - 1), np.diff(g["syn_indptr"]))
    order = np.lexsort((np.arange(len(cols)), cols, g["syn_indices"]))
    assert np.array_equal(row_ptr, np.concatenate([[0], np.cumsum(np.bincount(g["syn_indices"], minlength=n))]))
    assert np.array_equal(col, cols[order])
    assert np.array_equal(val, g["syn_dose"][order])
    blocks = 0
    for b in range(0, n, 8):
        pairs = [(c, k) for r in range(b, min(b + 8, n))
                 for c, k in zip(*np.unique(col[row_ptr[r]:row_ptr[r + 1]], return_counts=True))]
        per_col = {}
        for c, k in pairs:
            per_col[c] = max(per_col.get(c, 0), k)
        blocks += sum(per_col.values())
    assert dev.block_entries == blocks


@pytest.mark.parametrize("n_w", [1, 3, 8, 9, 64, 65, 128, 129, 300, 600])
def test_random_products_bitwise(spmm, n_w):
    rng = np.random.default_rng(1000 + n_w)
    m = _beamlet_like(rng, value_dtype=np.float64 if n_w % 2 else np.float32)
    w = rng.random((m.n_beamlets, n_w))
    w[rng.random(w.shape) < 0.1] = 0.0
    ref = _expected(m, w)
    dev = spmm.DeviceDij(m)
    assert np.array_equal(dev.apply(w), ref)
    for r0, r1 in ((5, 1234), (8, 16), (m.n_voxels - 3, m.n_voxels), (4099, 9000)):
        assert np.array_equal(dev.apply(w, rows=(r0, r1)), ref[r0:r1]), (r0, r1)


@pytest.mark.parametrize("n_w", [1, 5, 40, 200, 300, 1000])
def test_float32_products_within_rounding(spmm, n_w):
    rng = np.random.default_rng(2000 + n_w)
    m = _beamlet_like(rng, value_dtype=np.float64 if n_w % 2 else np.float32)
    w = rng.random((m.n_beamlets, n_w))
    ref = _expected(m, w)
    got = spmm.DeviceDij(m).apply(w, precision="f32")
    assert got.dtype == np.float32
    absum = mc.dose_for_weights(m.indptr, m.indices, np.abs(m.dose), np.abs(w), m.n_voxels)
    terms = np.bincount(m.indices, minlength=m.n_voxels)[:, None]
    bound = (terms + 3) * 2.0 ** -24 * absum * 1.0001
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound)


def test_engine_dij_product_equals_host_container(spmm):
    from pyradmc_amd import AnalyticCrossSections, BeamletGridSource, HipEngine, VoxelGrid

    grid = VoxelGrid.uniform_water((12, 12, 20), (0.5, 0.5, 0.5))
    xs = AnalyticCrossSections(geometry_densities=grid.max_density_by_material())
    dij = HipEngine(grid, xs).run_dij(BeamletGridSource(4.0, -0.1, (1.0, 5.0), (1.0, 5.0), 3, 3),
                                      4000, 4, seed=7, truncation=1e-3)
    rng = np.random.default_rng(3)
    dev = dij.to_device()
    for _ in range(3):
        w = rng.random(dij.n_beamlets)
        assert np.array_equal(dev.dose_for_weights(w), dij.dose_for_weights(w))
    w = rng.random((dij.n_beamlets, 100))
    assert np.array_equal(dev.apply(w), _expected(dij, w))


def test_device_resident_weights(spmm):
    import torch

    rng = np.random.default_rng(4)
    m = _beamlet_like(rng, n_cols=60)
    dev = spmm.DeviceDij(m)
    w = rng.random((m.n_beamlets, 77))
    for dtype, precision in ((torch.float64, "f64"), (torch.float32, "f32")):
        wt = torch.from_numpy(w).to(dtype).to("cuda:0")
        out = torch.empty((m.n_voxels - 10, 77), dtype=dtype, device="cuda:0")
        ms = dev.apply_device(wt, out, rows=(10, m.n_voxels))
        assert ms >= 0.0
        assert np.array_equal(out.cpu().numpy(), dev.apply(w, rows=(10, m.n_voxels), precision=precision))


def test_buffers_resize_between_products(spmm):
    rng = np.random.default_rng(5)
    m = _beamlet_like(rng, n_cols=40)
    dev = spmm.DeviceDij(m)
    for n_w in (300, 2, 300, 17):
        w = rng.random((m.n_beamlets, n_w))
        assert np.array_equal(dev.apply(w), _expected(m, w)), n_w


def test_empty_matrix_and_argument_errors(spmm):
    empty = spmm.DeviceDij(_matrix(np.zeros(4, dtype=np.int64), [], [], 50))
    assert np.array_equal(empty.apply(np.ones((3, 2))), np.zeros((50, 2)))
    assert empty.block_entries == 0
    with pytest.raises(ValueError):
        spmm.DeviceDij(_matrix([0, 2, 1, 2], [0, 1], [1.0, 1.0], 5))  # decreasing indptr
    with pytest.raises(ValueError):
        spmm.DeviceDij(_matrix([0, 2], [0, 5], [1.0, 1.0], 5))  # row index out of range
    dev = spmm.DeviceDij(_matrix([0, 2], [0, 4], [1.0, 2.0], 5))
    with pytest.raises(ValueError):
        dev.apply(np.ones((2, 3)))
    with pytest.raises(ValueError):
        dev.apply(np.ones(1), rows=(3, 3))
    with pytest.raises(ValueError):
        dev.apply(np.ones(1), rows=(0, 6))
    assert np.array_equal(dev.apply(np.array([0.5])), [0.5, 0, 0, 0, 1.0])


@pytest.mark.parametrize("precision,n_cols", [("f64", 280_000), ("f32", 540_000)])
def test_weights_beyond_two_gib(spmm, precision, n_cols):
    """Weights of more than 2 GiB take the kernel with 64-bit weight addressing; its
    products equal those of the same matrix restricted to its non-empty columns,
    whose weights take the buffer-resource kernel."""
    rng = np.random.default_rng(6)
    small = _beamlet_like(rng, shape=(8, 10, 50), n_cols=30)
    n_w = 1000
    cols = np.concatenate([np.arange(20), n_cols - 10 + np.arange(10)])  # where the 30 columns go
    counts = np.zeros(n_cols, dtype=np.int64)
    counts[cols] = np.diff(small.indptr)
    big = _matrix(np.concatenate([[0], np.cumsum(counts)]), small.indices, small.dose, small.n_voxels)
    w_small = rng.random((30, n_w))
    w_big = np.zeros((n_cols, n_w))
    w_big[cols] = w_small
    assert n_cols * n_w * (8 if precision == "f64" else 4) > 2 ** 31
    got = spmm.DeviceDij(big).apply(w_big, precision=precision)
    del w_big
    want = spmm.DeviceDij(small).apply(w_small, precision=precision)
    assert np.array_equal(got, want)
    if precision == "f64":
        assert np.array_equal(want, _expected(small, w_small))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_many_repeated_entries_and_wide_column_sets(spmm, precision):
    """Row blocks with more entries in one group of 16 distinct columns than a chunk
    of 32 holds (each (row, column) pair repeated three times), and a 128-row group
    that touches more than 1024 distinct columns. Float64 equals the oracle bit for
    bit; float32 stays within its rounding bound."""
    rng = np.random.default_rng(7)
    n_rows, n_cols = 700, 1500
    indptr, indices, values = [0], [], []
    for c in range(n_cols):
        rows = [c % 128]  # rows 0..127 see every column: 1500 distinct columns in that group
        if c < 40:
            rows += [1, 1, 1, 2, 2, 2]  # rows 1 and 2: three entries per column
        rows += list(rng.choice(np.arange(128, n_rows), size=3, replace=False))
        indices += rows
        values += list(rng.random(len(rows)) * 10.0 ** rng.uniform(-3, 1, len(rows)))
        indptr.append(len(indices))
    m = _matrix(indptr, indices, np.asarray(values), n_rows)
    dev = spmm.DeviceDij(m)
    for n_w in (100, 300, 1000):
        w = rng.random((n_cols, n_w))
        ref = _expected(m, w)
        got = dev.apply(w, precision=precision)
        if precision == "f64":
            assert np.array_equal(got, ref), n_w
            assert np.array_equal(dev.apply(w, rows=(3, 650)), ref[3:650]), n_w
        else:
            absum = mc.dose_for_weights(m.indptr, m.indices, np.abs(m.dose), np.abs(w), m.n_voxels)
            terms = np.bincount(m.indices, minlength=m.n_voxels)[:, None]
            bound = (terms + 3) * 2.0 ** -24 * absum * 1.0001
            assert np.all(np.abs(got.astype(np.float64) - ref) <= bound), n_w
            part = dev.apply(w, rows=(3, 650), precision=precision).astype(np.float64)
            assert np.all(np.abs(part - ref[3:650]) <= bound[3:650]), n_w
