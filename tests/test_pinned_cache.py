"""The cache of page-locked result blocks (``pyradmc_amd._native``), driven through a
stand-in allocator so it runs without a device: blocks come back for reuse, the size
window and the byte cap hold, and a release frees exactly what is cached."""

import ctypes

import numpy as np
import pytest

from pyradmc_amd import _native as nat


class _FakeLib:
    def __init__(self):
        self.live = {}
        self.allocs = 0
        self.frees = 0

    def hmc_host_alloc(self, nbytes, out):
        buf = ctypes.create_string_buffer(int(nbytes))
        address = ctypes.addressof(buf)
        self.live[address] = buf
        self.allocs += 1
        ctypes.cast(out, ctypes.POINTER(ctypes.c_void_p))[0] = address
        return 0

    def hmc_host_free(self, address):
        del self.live[int(address)]
        self.frees += 1


@pytest.fixture
def fake(monkeypatch):
    f = _FakeLib()
    monkeypatch.setattr(nat, "_lib", f)
    monkeypatch.setattr(nat, "_pinned_cache", [])
    monkeypatch.setattr(nat, "_pinned_cached_bytes", 0)
    yield f
    nat.release_pinned_cache()


def _address(a):
    return a.__array_interface__["data"][0]


def test_released_block_is_reused_for_the_next_result(fake):
    a = nat.pinned_empty(1 << 18, np.float64)
    a[:] = 1.0
    first = _address(a)
    del a
    assert fake.frees == 0 and nat._pinned_cached_bytes == 8 << 18
    b = nat.pinned_empty(1 << 18, np.float64)
    assert _address(b) == first and fake.allocs == 1
    c = nat.pinned_empty(1 << 18, np.float64)  # b still alive: a second block
    assert _address(c) != first and fake.allocs == 2
    view = c[10:20]
    del c
    assert nat._pinned_cached_bytes == 0  # a view keeps the block
    view[:] = 3.0
    del view, b
    assert nat._pinned_cached_bytes == 2 * (8 << 18)


def test_size_window_and_cap(fake, monkeypatch):
    a = nat.pinned_empty(1 << 20, np.uint8)
    del a
    small = nat.pinned_empty(1 << 18, np.uint8)  # a 1 MiB block is over twice 256 KiB
    assert fake.allocs == 2
    half = nat.pinned_empty(600_000, np.uint8)  # within twice: takes the cached MiB
    assert fake.allocs == 2 and half.size == 600_000
    del small, half
    cached = nat._pinned_cached_bytes
    monkeypatch.setattr(nat, "PINNED_CACHE_MAX_BYTES", cached + (1 << 20))
    big = nat.pinned_empty(3 << 20, np.uint8)
    del big  # would pass the cap: freed, not cached
    assert fake.frees == 1 and nat._pinned_cached_bytes == cached


def test_release_frees_every_cached_block(fake):
    arrays = [nat.pinned_empty(5000 * (k + 1), np.int64) for k in range(4)]
    del arrays
    cached = nat._pinned_cached_bytes
    assert cached > 0
    assert nat.release_pinned_cache() == cached
    assert fake.frees == 4 and not fake.live and nat._pinned_cached_bytes == 0
