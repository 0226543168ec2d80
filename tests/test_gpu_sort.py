"""The binning primitives on the GPU, called through the library's exported C functions:
`va_sort_pairs` (stable LSD radix sort by the low `bits` key bits; bins keep API primitive order
only because it is stable) against numpy's stable argsort, and `va_scan_u32` (exclusive scan of
the per-primitive tile counts) against numpy's cumulative sum. Sizes cover a partial workgroup,
exact workgroup multiples, one and two radix passes, keys with bits above `bits`, equal keys, a
capacity above the pair count (idle workgroups) and a device pair count above the capacity (only
the first `cap` pairs are sorted); the scan covers one block, block edges and the size at which
the block totals get their own scan kernel."""
import ctypes as C
import os

import numpy as np
import pytest

from visor_amd import vk

pytestmark = pytest.mark.gpu

vp = C.c_void_p


class Hip:
    """Device memory through the HIP runtime the library itself links (torch bundles another
    runtime, which cannot share the process with it)."""

    def __init__(self):
        path = "/opt/rocm/lib/libamdhip64.so.7"
        self.h = C.CDLL(path if os.path.exists(path) else "libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        self.h.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [vp]
        self.live = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        p = vp()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.zeros(n, np.uint32)
        if n:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def kit(ctx):
    lib = vk.load().lib
    lib.va_sort_temp_words.restype = C.c_size_t
    lib.va_sort_temp_words.argtypes = [C.c_uint32]
    lib.va_sort_pairs.restype = C.c_int
    lib.va_sort_pairs.argtypes = [vp, vp, vp, vp, vp, vp, C.c_uint32, C.c_int, C.POINTER(C.c_int), vp]
    lib.va_scan_temp_words.restype = C.c_size_t
    lib.va_scan_temp_words.argtypes = [C.c_uint32]
    lib.va_scan_u32.restype = C.c_int
    lib.va_scan_u32.argtypes = [vp, vp, vp, C.c_uint32, vp]
    hip = Hip()
    yield lib, hip
    hip.free_all()


def sort_on_gpu(kit, keys, vals, bits, cap=None, n_dev=None):
    lib, hip = kit
    n = len(keys)
    cap = n if cap is None else cap
    n_dev = n if n_dev is None else n_dev
    size = max(cap, 1)
    k = np.zeros(size, np.uint32)
    v = np.zeros(size, np.uint32)
    k[:n], v[:n] = keys, vals
    dk, dv = hip.upload(k), hip.upload(v)
    dk2, dv2 = hip.upload(np.zeros(size, np.uint32)), hip.upload(np.zeros(size, np.uint32))
    dn = hip.upload(np.array([n_dev], np.uint32))
    temp = hip.upload(np.zeros(max(int(lib.va_sort_temp_words(cap)), 1), np.uint32))
    alt = C.c_int(-1)
    rc = lib.va_sort_pairs(temp, dk, dv, dk2, dv2, dn, cap, bits, C.byref(alt), None)
    assert rc == 0
    hip.sync()
    ok, ov = (dk2, dv2) if alt.value else (dk, dv)
    out = hip.download(ok, size), hip.download(ov, size), alt.value
    hip.free_all()
    return out


def expect_sorted(keys, vals, bits):
    order = np.argsort(keys & np.uint32((1 << bits) - 1), kind="stable")
    return keys[order], vals[order]


@pytest.mark.parametrize("n", [1, 255, 4095, 4096, 4097, 3 * 4096, 100_003])
@pytest.mark.parametrize("bits", [8, 15, 16])
def test_sort_is_numpy_stable_argsort(kit, n, bits):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)  # primitive order
    got_k, got_v, alt = sort_on_gpu(kit, keys, vals, bits)
    want_k, want_v = expect_sorted(keys, vals, bits)
    assert alt == ((bits + 7) // 8) % 2
    assert np.array_equal(got_k[:n], want_k)
    assert np.array_equal(got_v[:n], want_v)


def test_sort_ignores_key_bits_above_bits(kit):
    """Only the low `bits` bits order the pairs; the high bits travel with their key."""
    rng = np.random.default_rng(7)
    n = 20_000
    keys = (rng.integers(0, 1 << 16, n, dtype=np.uint32) | (rng.integers(0, 1 << 16, n, dtype=np.uint32) << 16))
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint32)
    got_k, got_v, _ = sort_on_gpu(kit, keys, vals, 16)
    want_k, want_v = expect_sorted(keys, vals, 16)
    assert np.array_equal(got_k[:n], want_k) and np.array_equal(got_v[:n], want_v)


def test_sort_of_equal_keys_keeps_order(kit):
    n = 9000
    keys = np.full(n, 12345, np.uint32)
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    got_k, got_v, _ = sort_on_gpu(kit, keys, vals, 16)
    assert np.array_equal(got_k[:n], keys) and np.array_equal(got_v[:n], vals)


def test_sort_with_capacity_above_the_count(kit):
    """The grid is sized by the capacity; workgroups past the device count do nothing."""
    rng = np.random.default_rng(11)
    n = 5000
    keys = rng.integers(0, 1 << 15, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    got_k, got_v, _ = sort_on_gpu(kit, keys, vals, 15, cap=40_000)
    want_k, want_v = expect_sorted(keys, vals, 15)
    assert np.array_equal(got_k[:n], want_k) and np.array_equal(got_v[:n], want_v)


def test_sort_with_count_above_the_capacity(kit):
    """A pair count above the capacity (bins overflowed) sorts the first `cap` pairs only."""
    rng = np.random.default_rng(13)
    cap = 6000
    keys = rng.integers(0, 1 << 15, cap, dtype=np.uint32)
    vals = np.arange(cap, dtype=np.uint32)
    got_k, got_v, _ = sort_on_gpu(kit, keys, vals, 15, cap=cap, n_dev=cap + 777)
    want_k, want_v = expect_sorted(keys, vals, 15)
    assert np.array_equal(got_k, want_k) and np.array_equal(got_v, want_v)


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 1_000_001, 4_194_305])
def test_scan_is_numpy_exclusive_cumsum(kit, n):
    """4,194,305 items make 2049 blocks of 2048: the block totals then get their own scan."""
    lib, hip = kit
    rng = np.random.default_rng(n)
    x = rng.integers(0, 64, n, dtype=np.uint32)
    dx = hip.upload(x)
    out = hip.upload(np.zeros(n, np.uint32))
    temp = hip.upload(np.zeros(int(lib.va_scan_temp_words(n)), np.uint32))
    rc = lib.va_scan_u32(temp, dx, out, n, None)
    assert rc == 0
    hip.sync()
    got = hip.download(out, n)
    hip.free_all()
    want = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(got, want)
