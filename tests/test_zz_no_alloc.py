"""The allocation-free product acd_sparse_matrix_multiply_noalloc, called
through the C ABI: bit for bit the result of sparse_matrix_multiply (and
of the oracle) on two calls in a row, the second with a smaller input, and
clear errors for lengths beyond the staging maximum or not matching the
operator."""
import ctypes as C

import numpy as np
import pytest

from acdalm import instances as I

pytestmark = pytest.mark.gpu
MAX = 4194304
dp = C.POINTER(C.c_double)


def noalloc(gpu, m, x, transpose, ny=None):
    L = gpu.lib
    fn = L.acd_sparse_matrix_multiply_noalloc
    fn.restype = C.c_int
    fn.argtypes = [dp, C.c_size_t, C.c_void_p, dp, C.c_size_t, C.c_int]
    nr, nc = gpu.shape(m)
    if ny is None:
        ny = nc if transpose else nr
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.full(max(ny, 1), np.nan)
    rc = fn(y.ctypes.data_as(dp), ny, m, x.ctypes.data_as(dp), x.shape[0],
            transpose)
    return rc, y[:ny]


def test_two_calls_match_the_allocating_entry_point(gpu, oracle):
    big = I.random_lp(3000, 7000, 6, 31)
    small = I.random_lp(50, 120, 3, 32)
    rng = np.random.default_rng(5)
    for lp in (big, small):          # second call: smaller input
        print(f"[no_alloc] {lp.nrows}x{lp.ncols}", flush=True)
        m = gpu.make_matrix(lp, 1)
        mo = oracle.make_matrix(lp, 1)
        for tr in (0, 1):
            nx = lp.nrows if tr else lp.ncols
            x = rng.standard_normal(nx)
            rc, y = noalloc(gpu, m, x, tr)
            assert rc == 0, gpu.last_error()
            assert np.array_equal(y, gpu.multiply(m, x, tr))
            assert np.array_equal(y, oracle.multiply(mo, x, tr))
        gpu.free_matrix(m)
        oracle.free_matrix(mo)


def test_empty_rows_are_written(gpu):
    """Rows without entries come out as +0.0 even though the staging memory
    keeps the previous call's values."""
    lp = I.random_lp(40, 90, 3, 33)
    keep = lp.rows != 7
    m = gpu.make_matrix(lp, 0, rows=lp.rows[keep], cols=lp.cols[keep],
                        vals=lp.vals[keep])
    print("[no_alloc] empty row", flush=True)
    rc, y = noalloc(gpu, m, np.ones(lp.ncols), 0)
    assert rc == 0 and y[7] == 0.0 and np.all(np.isfinite(y))
    gpu.free_matrix(m)


def test_rejects_lengths_beyond_the_maximum_and_mismatches(gpu):
    n = MAX + 5
    rows = np.array([0, 1, 2], dtype=np.uint32)
    cols = np.array([0, MAX, MAX + 4], dtype=np.uint32)
    vals = np.array([1.0, 2.0, 3.0])
    lp = I.random_lp(3, 8, 1, 34)
    m = gpu.make_matrix(lp, 0, rows=rows, cols=cols, vals=vals, nrows=3,
                        ncols=n)
    print("[no_alloc] limits", flush=True)
    gpu.last_error()
    rc, _ = noalloc(gpu, m, np.ones(n), 0)
    assert rc == -1 and "maximum" in gpu.last_error()[1]
    gpu.free_matrix(m)
    small = gpu.make_matrix(lp, 0)
    rc, _ = noalloc(gpu, small, np.ones(lp.ncols + 1), 0)
    assert rc == -1 and "mismatch" in gpu.last_error()[1]
    rc, y = noalloc(gpu, small, np.ones(lp.ncols), 0)
    assert rc == 0 and np.array_equal(y, gpu.multiply(small, np.ones(lp.ncols)))
    gpu.free_matrix(small)
