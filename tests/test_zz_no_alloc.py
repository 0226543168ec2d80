"""`va_sort_pairs_noalloc`, the binning sort without a workspace argument, against `va_sort_pairs`
with its own workspace: the same keys, values and result buffer on two calls in a row (the second
with fewer pairs, so stale histogram words of the first call sit in the shared device array), at
the largest capacity the array is sized for, and the refusal of a capacity beyond it. Both are
also checked against numpy's stable argsort."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_sort import Hip, expect_sorted
from visor_amd import vk

pytestmark = pytest.mark.gpu

vp = C.c_void_p
MAX_PAIRS = 1 << 24


@pytest.fixture(scope="module")
def kit(ctx):
    lib = vk.load().lib
    lib.va_sort_temp_words.restype = C.c_size_t
    lib.va_sort_temp_words.argtypes = [C.c_uint32]
    lib.va_sort_pairs.restype = C.c_int
    lib.va_sort_pairs.argtypes = [vp, vp, vp, vp, vp, vp, C.c_uint32, C.c_int, C.POINTER(C.c_int), vp]
    lib.va_sort_pairs_noalloc.restype = C.c_int
    lib.va_sort_pairs_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, C.c_int, C.POINTER(C.c_int), vp]
    hip = Hip()
    yield lib, hip
    hip.free_all()


def run(kit, keys, vals, bits, cap, noalloc):
    lib, hip = kit
    n = len(keys)
    k = np.zeros(cap, np.uint32)
    v = np.zeros(cap, np.uint32)
    k[:n], v[:n] = keys, vals
    dk, dv = hip.upload(k), hip.upload(v)
    dk2, dv2 = hip.upload(np.zeros(cap, np.uint32)), hip.upload(np.zeros(cap, np.uint32))
    dn = hip.upload(np.array([n], np.uint32))
    alt = C.c_int(-1)
    if noalloc:
        rc = lib.va_sort_pairs_noalloc(dk, dv, dk2, dv2, dn, cap, bits, C.byref(alt), None)
    else:
        temp = hip.upload(np.zeros(int(lib.va_sort_temp_words(cap)), np.uint32))
        rc = lib.va_sort_pairs(temp, dk, dv, dk2, dv2, dn, cap, bits, C.byref(alt), None)
    assert rc == 0
    hip.sync()
    ok, ov = (dk2, dv2) if alt.value else (dk, dv)
    out = hip.download(ok, n), hip.download(ov, n), alt.value
    hip.free_all()
    return out


@pytest.mark.parametrize("n,cap,bits", [(1_000_003, 1_000_003, 16), (5_001, 5_001, 24), (70_000, MAX_PAIRS, 16)])
def test_noalloc_sort_equals_the_workspace_sort(kit, n, cap, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    want = run(kit, keys, vals, bits, cap, noalloc=False)
    got = run(kit, keys, vals, bits, cap, noalloc=True)
    assert got[2] == want[2]
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    wk, wv = expect_sorted(keys, vals, bits)
    assert np.array_equal(got[0], wk) and np.array_equal(got[1], wv)


def test_noalloc_sort_twice_in_a_row_second_smaller(kit):
    rng = np.random.default_rng(5)
    for n in (300_000, 777):
        keys = rng.integers(0, 1 << 16, n, dtype=np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint32)
        got = run(kit, keys, vals, 16, n, noalloc=True)
        want = run(kit, keys, vals, 16, n, noalloc=False)
        assert got[2] == want[2]
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_noalloc_sort_refuses_a_capacity_beyond_its_array(kit):
    lib, _ = kit
    alt = C.c_int(-1)
    # refused on the host before any device work: the buffers are never touched
    rc = lib.va_sort_pairs_noalloc(None, None, None, None, None, MAX_PAIRS + 1, 16, C.byref(alt), None)
    assert rc != 0 and alt.value == 0
