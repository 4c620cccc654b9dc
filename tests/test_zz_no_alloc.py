"""GPU: the pair seam's allocation-free entry point (b2g_pairs_noalloc) against b2g_pairs.

b2g_pairs_noalloc takes device inputs and outputs and keeps its scratch in a
file-scope device array; the test allocates the device buffers itself through
the HIP runtime that libb2g.so links (never torch's), calls the entry point
twice in a row (the second call with a smaller input, so stale scratch from
the first call would show), and compares each result with b2g_pairs on the
same input: the same pairs in the same canonical order, bit for bit.  Sizes
beyond the scratch array are rejected before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from pypybox2d_amd import _native as N

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip(object):
    """hipMalloc / hipMemcpy / hipFree of the runtime libb2g.so is linked against."""

    def __init__(self):
        N.load()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(16, nbytes)) == 0
        self.live.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def get(self, p, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def bind():
    lib = N.load()
    lib.b2g_pairs_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return lib


def boxes(rng, n, grid):
    lo = rng.uniform(-20, 20, (n, 2))
    if grid:  # boxes on a half-metre grid touch exactly (the overlap test is inclusive)
        lo = np.round(lo * 2) / 2
    size = np.full((n, 2), 0.5) if grid else rng.uniform(0.2, 2.0, (n, 2))
    return np.concatenate([lo, lo + size], axis=1)


def noalloc_pairs(hip, lib, fat, move, cap):
    dfat, dmove = hip.put(fat.astype(np.float64)), hip.put(move.astype(np.int32))
    dout, dcnt = hip.alloc(8 * max(cap, 1)), hip.alloc(4)
    rc = lib.b2g_pairs_noalloc(fat.shape[0], dfat, move.shape[0], dmove, dout, cap, dcnt, None)
    assert rc == N.B2G_OK, rc
    hip.sync()
    count = int(hip.get(dcnt, np.int32, 1)[0])
    pairs = hip.get(dout, np.int32, 2 * min(max(count, 0), cap)).reshape(-1, 2)
    return count, pairs


def test_noalloc_pair_seam_matches_the_allocating_one_on_two_calls_in_a_row():
    hip, lib = Hip(), bind()
    rng = np.random.default_rng(21)
    try:
        for trial, (n, nmove, grid) in enumerate(((900, 700, False), (300, 120, True))):  # second call smaller
            fat = boxes(rng, n, grid)
            move = rng.permutation(n)[:nmove].astype(np.int32)
            want = N.pairs(fat, move)
            count, got = noalloc_pairs(hip, lib, fat, move, cap=want.shape[0] + 64)
            print("call", trial, "proxies", n, "moved", nmove, "pairs", want.shape[0], "noalloc count", count,
                  flush=True)
            assert count == want.shape[0]
            assert np.array_equal(got, want)
        # a buffer too small for the pairs: the count is still the total, the prefix is exact
        count, got = noalloc_pairs(hip, lib, fat, move, cap=5)
        assert count == want.shape[0] and np.array_equal(got, want[:5])
        # an invalid move buffer (a duplicate) is reported as count -1 and nothing is written
        count, _ = noalloc_pairs(hip, lib, fat, np.array([3, 7, 3], np.int32), cap=16)
        assert count == -1
    finally:
        hip.free()


def test_noalloc_pair_seam_rejects_sizes_beyond_its_scratch():
    lib = bind()
    big = (1 << 20) + 1
    assert lib.b2g_pairs_noalloc(big, None, 0, None, None, 0, None, None) == N.B2G_E_INVALID
    fake = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    assert lib.b2g_pairs_noalloc(16, fake, big, fake, fake, 4, fake, None) == N.B2G_E_INVALID
    assert lib.b2g_pairs_noalloc(-1, fake, 1, fake, fake, 4, fake, None) == N.B2G_E_INVALID
