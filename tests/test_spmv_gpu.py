"""GPU parity of the sparse-matrix layer against the oracle.

Covers spmv.h: sparse_matrix_make/copy (device permutation, CSR order and
Z-order, which must be bit-identical), sparse_matrix_multiply(_2) on A and
A' (bit-identical per-row sums), sparse_matrix_row/col_permute, and the
accessors.  The case list follows the reference's TEST_SPMV
(spmv/spmv.c:341-385: every shape below 20x20, dense +-1 at density 0.5)
and adds empty rows/columns, a single entry, long rows and the workload
shapes.
"""
import numpy as np
import pytest

from acdalm import instances as I
from helpers import bit_equal, spmv_scaled_error

pytestmark = pytest.mark.gpu


def dense_pm1(nrows, ncols, seed):
    rng = np.random.default_rng(seed)
    d = np.where(rng.random((nrows, ncols)) < 0.5,
                 np.where(rng.random((nrows, ncols)) < 0.5, -1.0, 1.0), 0.0)
    r, c = np.nonzero(d)
    return I.LP(nrows, ncols, r, c, d[r, c], np.zeros(nrows), np.zeros(ncols),
                np.zeros(ncols), np.zeros(ncols), name=f"dense{nrows}x{ncols}")


def special_cases():
    yield I.tiny_lp()
    # one entry, empty rows and columns around it
    yield I.LP(5, 7, [3], [6], [2.5], np.zeros(5), np.zeros(7), np.zeros(7),
               np.zeros(7), name="single")
    # a long row (dense row over 3000 columns) plus sparse rows
    n = 3000
    r = np.concatenate([np.zeros(n, np.int64), np.arange(1, 200)])
    c = np.concatenate([np.arange(n), np.arange(1, 200) * 7 % n])
    v = np.linspace(-1, 1, r.shape[0]) + 0.25
    yield I.LP(200, n, r, c, v, np.zeros(200), np.zeros(n), np.zeros(n),
               np.zeros(n), name="long_row")
    yield I.random_lp(301, 555, 7, 12)
    yield I.config0(1)
    yield I.set_cover(3000, 15000, 5)
    yield I.multicommodity(12, 5, 3)


CASES = [dense_pm1(r, c, 100 * r + c) for r in range(0, 20, 3)
         for c in range(0, 20, 4)] + list(special_cases())


@pytest.mark.parametrize("lp", CASES, ids=lambda lp: lp.name)
@pytest.mark.parametrize("permute", [0, 1])
def test_build_and_multiply_bit_exact(gpu, oracle, lp, permute):
    print(f"[spmv] {lp.name} permute={permute}", flush=True)
    mg = gpu.make_matrix(lp, permute)
    mo = oracle.make_matrix(lp, permute)
    try:
        assert gpu.shape(mg) == oracle.shape(mo)
        for a, b in zip(gpu.coo(mg), oracle.coo(mo)):
            assert bit_equal(a, b)
        rng = np.random.default_rng(7)
        x = rng.standard_normal(lp.ncols)
        y = rng.standard_normal(lp.nrows)
        assert bit_equal(gpu.multiply(mg, x, 0), oracle.multiply(mo, x, 0))
        assert bit_equal(gpu.multiply(mg, y, 1), oracle.multiply(mo, y, 1))
        x2 = rng.standard_normal(lp.ncols)
        g0, g1 = gpu.multiply_2(mg, x, x2, 0)
        o0, o1 = oracle.multiply_2(mo, x, x2, 0)
        assert bit_equal(g0, o0) and bit_equal(g1, o1)
        y2 = rng.standard_normal(lp.nrows)
        g0, g1 = gpu.multiply_2(mg, y, y2, 1)
        o0, o1 = oracle.multiply_2(mo, y, y2, 1)
        assert bit_equal(g0, o0) and bit_equal(g1, o1)
        for row in (0, 1):
            n = lp.nrows if row else lp.ncols
            v = rng.standard_normal(n)
            for d in (1, -1):
                rg, pg = gpu.permute(mg, v, row, d)
                ro, po = oracle.permute(mo, v, row, d)
                assert rg == ro == 0 and bit_equal(pg, po)
            # a length mismatch is reported, not acted on
            assert gpu.permute(mg, np.zeros(n + 1), row, 1)[0] == 1
    finally:
        gpu.free_matrix(mg)
        oracle.free_matrix(mo)


def test_copy_renumbers_again(gpu, oracle):
    lp = I.random_lp(500, 900, 5, 21)
    mg, mo = gpu.make_matrix(lp, 1), oracle.make_matrix(lp, 1)
    cg, co = gpu.copy_matrix(mg, 1), oracle.copy_matrix(mo, 1)
    try:
        for a, b in zip(gpu.coo(cg), oracle.coo(co)):
            assert bit_equal(a, b)
    finally:
        for m in (mg, cg):
            gpu.free_matrix(m)
        for m in (mo, co):
            oracle.free_matrix(m)


def test_dense_agreement_like_reference_test(gpu):
    """spmv/spmv.c:354-366 checks sparse against dense products to 1e-4;
    the bit-exact tests above are stronger, this one keeps the original
    criterion on the GPU path with the norm-scaled 1e-13 bound."""
    for seed in range(5):
        lp = dense_pm1(17, 13, seed)
        m = gpu.make_matrix(lp, 0)
        rng = np.random.default_rng(seed)
        x = rng.uniform(-1, 1, 13)
        y = gpu.multiply(m, x, 0)
        d = lp.dense()
        assert np.linalg.norm(y - d @ x) < 1e-4
        assert spmv_scaled_error(y, d @ x, lp.rows, lp.cols, lp.vals, x,
                                 17) < 1e-13
        gpu.free_matrix(m)


def test_autotuned_variants_give_identical_products(gpu, oracle):
    """The autotuner only changes scheduling: every variant it may keep
    sums each row in the same order, so products stay bit-identical."""
    import ctypes
    lp = I.set_cover(3000, 12000, 4)
    mg, mo = gpu.make_matrix(lp, 1), oracle.make_matrix(lp, 1)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(lp.ncols)
    y = rng.standard_normal(lp.nrows)
    for v in range(13):
        gpu.lib.acd_spmv_set_variant(v)
        assert bit_equal(gpu.multiply(mg, x, 0), oracle.multiply(mo, x, 0)), v
        assert bit_equal(gpu.multiply(mg, y, 1), oracle.multiply(mo, y, 1)), v
    gpu.lib.acd_spmv_set_variant(1)
    tuned = (ctypes.c_int * 2)()
    assert gpu.lib.acd_spmv_autotune(mg, 3, tuned) == 0
    assert 0 <= tuned[0] < 13 and 0 <= tuned[1] < 13
    assert bit_equal(gpu.multiply(mg, x, 0), oracle.multiply(mo, x, 0))
    assert bit_equal(gpu.multiply(mg, y, 1), oracle.multiply(mo, y, 1))
    gpu.free_matrix(mg)
    oracle.free_matrix(mo)
