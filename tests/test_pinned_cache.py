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
    huge = nat.pinned_empty((cached >> 20) + (8 << 20), np.uint8)
    del huge  # larger than the cap itself: freed, not cached
    assert fake.frees == 1 and nat._pinned_cached_bytes == cached
    big = nat.pinned_empty(2 << 20, np.uint8)  # within the cap, not beside what is cached
    big_address = _address(big)
    del big  # would pass the cap: the oldest cached blocks go to make room
    assert fake.frees >= 2 and nat._pinned_cached_bytes <= nat.PINNED_CACHE_MAX_BYTES
    assert nat._pinned_cache[-1] == (2 << 20, big_address)
    again = nat.pinned_empty(2 << 20, np.uint8)
    assert _address(again) == big_address  # the newest block serves the next run


def test_release_frees_every_cached_block(fake):
    arrays = [nat.pinned_empty(5000 * (k + 1), np.int64) for k in range(4)]
    del arrays
    cached = nat._pinned_cached_bytes
    assert cached > 0
    assert nat.release_pinned_cache() == cached
    assert fake.frees == 4 and not fake.live and nat._pinned_cached_bytes == 0


class _Handle:
    def __init__(self, log):
        self.log = log

    def close(self):
        self.log.append(self)


def test_shutdown_closes_handles_then_frees_every_block(fake, monkeypatch):
    monkeypatch.setattr(nat, "_exiting", False)
    monkeypatch.setattr(nat, "_open_handles", nat.weakref.WeakSet())
    closed = []
    handles = [_Handle(closed), _Handle(closed)]
    for h in handles:
        nat.register_handle(h)
    cached = nat.pinned_empty(1 << 16, np.float64)
    held = nat.pinned_empty(1 << 17, np.float64)  # still viewed when the process exits
    del cached
    assert fake.frees == 0 and nat._pinned_cached_bytes > 0
    nat.shutdown()
    assert sorted(map(id, closed)) == sorted(map(id, handles))
    assert fake.frees == 1 and nat._pinned_cached_bytes == 0
    held[:] = 2.0  # the block a live array views stays valid
    del held  # ... and is freed, not cached, once the array goes
    assert fake.frees == 2 and not fake.live and nat._pinned_cached_bytes == 0
    nat.shutdown()  # a second call finds nothing left
    assert len(closed) == 2 and fake.frees == 2


def test_library_load_registers_shutdown_at_exit():
    import subprocess
    import sys
    from pathlib import Path

    # A child process loads the library, registers a handle and exits without closing
    # it: the exit handler must close it (the library loads without a device).
    code = (
        "from pyradmc_amd import _native as nat\n"
        "nat.lib()\n"
        "class H:\n"
        "    def close(self):\n"
        "        print('closed at exit', flush=True)\n"
        "h = H()\n"
        "nat.register_handle(h)\n"
    )
    root = Path(__file__).resolve().parent.parent
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert "closed at exit" in out.stdout
