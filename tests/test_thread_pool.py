"""The host thread pool (include/thread_pool/thread_pool.h): the reference's
parallel-for and map-reduce over index ranges (thread_pool/thread_pool.c:
296-401), checked against serial loops.  Host code only (no device)."""
import ctypes as C
import threading

import numpy as np
import pytest

from acdalm import capi

RANGE_FN = C.CFUNCTYPE(None, C.c_size_t, C.c_size_t, C.c_void_p, C.c_uint)
MAP_FN = C.CFUNCTYPE(C.c_double, C.c_size_t, C.c_size_t, C.c_void_p, C.c_uint)
SUM, MAX, MIN = 0, 1, 2


@pytest.fixture(scope="module")
def lib():
    L = capi.CLib(capi.LIB_PATH, init=False).lib
    L.thread_pool_init.restype = C.c_void_p
    L.thread_pool_init.argtypes = [C.c_uint]
    L.thread_pool_free.argtypes = [C.c_void_p]
    L.thread_pool_count.restype = C.c_size_t
    L.thread_pool_count.argtypes = [C.c_void_p]
    L.thread_pool_for.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t,
                                  C.c_size_t, RANGE_FN, C.c_void_p]
    L.thread_pool_map_reduce.restype = C.c_double
    L.thread_pool_map_reduce.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t,
                                         C.c_size_t, MAP_FN, C.c_void_p,
                                         C.c_int, C.c_double]
    L.thread_pool_worker_storage.restype = C.POINTER(C.c_void_p)
    L.thread_pool_worker_storage.argtypes = [C.c_void_p, C.c_size_t]
    L.thread_pool_worker_storage_flat.restype = C.c_void_p
    L.thread_pool_worker_storage_flat.argtypes = [
        C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    return L


@pytest.fixture(params=[0, 1, 4])
def pool(lib, request):
    """NULL (serial), one worker, four workers."""
    p = lib.thread_pool_init(request.param) if request.param else None
    yield p
    if p:
        lib.thread_pool_free(p)


@pytest.mark.parametrize("lo, hi, gran", [(0, 1000, 1), (7, 12345, 16),
                                          (5, 5, 4), (3, 4, 100),
                                          (0, 100000, 64)])
def test_for_covers_the_range_once(lib, pool, lo, hi, gran):
    ranges, lock = [], threading.Lock()
    nworkers = lib.thread_pool_count(pool) if pool else 1

    def body(b, e, info, worker):
        with lock:
            ranges.append((b, e, worker))

    cb = RANGE_FN(body)
    lib.thread_pool_for(pool, lo, hi, gran, cb, None)
    ranges.sort()
    if lo == hi:
        # an empty range is handed over once, as one empty call
        # (thread_pool/thread_pool.c:304-313)
        assert ranges == [(lo, hi, 0)]
        return
    covered = lo
    for b, e, w in ranges:
        assert b == covered and b < e, (b, e, covered)
        assert w < max(1, nworkers)
        covered = e
    assert covered == hi
    # every chunk but the last is a whole number of granules
    for b, e, _ in ranges[:-1]:
        assert (e - b) % gran == 0, (b, e, gran)


def test_map_reduce_matches_serial_loops(lib, pool):
    n = 200000
    rng = np.random.default_rng(3)
    vals = rng.integers(-10**6, 10**6, size=n).astype(np.float64)
    buf = vals.ctypes.data_as(C.POINTER(C.c_double))

    def chunk(kind):
        def f(b, e, info, worker):
            part = vals[b:e]
            if kind == SUM:
                return float(part.sum())
            return float(part.max() if kind == MAX else part.min())
        return MAP_FN(f)

    for kind, init, want in [(SUM, 0.0, vals.sum()),
                             (MAX, -np.inf, vals.max()),
                             (MIN, np.inf, vals.min())]:
        cb = chunk(kind)
        got = lib.thread_pool_map_reduce(pool, 0, n, 32, cb, buf, kind, init)
        # integer-valued terms: every summation order gives the exact sum
        assert got == want, (kind, got, want)
    empty = lib.thread_pool_map_reduce(pool, 9, 9, 4, chunk(SUM), None, SUM,
                                       0.0)
    assert empty == 0.0


def test_worker_storage_is_zeroed_and_separate(lib):
    p = lib.thread_pool_init(3)
    try:
        assert lib.thread_pool_count(p) == 3
        ptrs = lib.thread_pool_worker_storage(p, 100)
        addrs = [ptrs[i] for i in range(3)]
        assert len(set(addrs)) == 3
        for a in addrs:
            assert a % 64 == 0
            assert bytes((C.c_char * 100).from_address(a)) == bytes(100)
        srt = sorted(addrs)
        assert all(b - a >= 100 for a, b in zip(srt, srt[1:]))
        stride = C.c_size_t()
        flat = lib.thread_pool_worker_storage_flat(p, 100, C.byref(stride))
        assert flat and stride.value >= 100 and stride.value % 64 == 0
    finally:
        lib.thread_pool_free(p)
