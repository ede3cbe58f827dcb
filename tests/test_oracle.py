"""Pins the oracle (oracle/*.c) before it is trusted as the GPU checker.

1. Golden fixtures written from the reference built from its own sources
   (tests/golden/make_golden.py): the oracle must reproduce them bit for
   bit -- triplet order, permutations, products, whole APPROX / SPG / ALM
   trajectories including the log text, and the randomised CD solver.
2. Closed forms and invariants that a plausible mistake would break:
   Z-order keys, the windowed count sort, CSR order, dense products,
   step sizes sum_j w_j beta_j A_ji^2, the optimum of a hand-solved LP and
   of LPs solved independently by HiGHS.
"""
import ctypes
import glob
import os

import numpy as np
import pytest

from acdalm import instances as I
from helpers import bit_equal

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


SPMV = {"small": I.random_lp(301, 555, 7, 12), "dense": I.random_lp(19, 13, 9, 5),
        "setcover": I.set_cover(300, 1200, 5), "mcf": I.multicommodity(6, 3, 3),
        "c0": I.config0(1)}
SOLVE = {"small": I.random_lp(60, 150, 4, 21), "setcover": I.set_cover(80, 300, 6),
         "mcf": I.multicommodity(4, 2, 7)}


def test_fixture_files_present():
    assert len(glob.glob(os.path.join(GOLD, "*.npz"))) >= 14


@pytest.mark.parametrize("name", sorted(SPMV))
def test_spmv_fixture(oracle, name):
    lp = SPMV[name]
    f = load(f"spmv_{name}.npz")
    assert str(f["digest"]) == lp.digest(), "instance generator drifted"
    for perm in (0, 1):
        m = oracle.make_matrix(lp, perm)
        r, c, v = oracle.coo(m)
        assert bit_equal(r, f[f"rows{perm}"]) and bit_equal(c, f[f"cols{perm}"])
        assert bit_equal(v, f[f"vals{perm}"])
        assert bit_equal(oracle.permute(m, np.arange(lp.nrows, dtype=float), 1, 1)[1],
                         f[f"rowperm{perm}"])
        assert bit_equal(oracle.permute(m, np.arange(lp.ncols, dtype=float), 0, 1)[1],
                         f[f"colperm{perm}"])
        assert bit_equal(oracle.multiply(m, f["x"], 0), f[f"ax{perm}"])
        assert bit_equal(oracle.multiply(m, f["y"], 1), f[f"aty{perm}"])
        assert bit_equal(np.stack(oracle.multiply_2(m, f["x"], f["x2"], 0)), f[f"ax2_{perm}"])
        assert bit_equal(np.stack(oracle.multiply_2(m, f["y"], f["y2"], 1)), f[f"aty2_{perm}"])
        oracle.free_matrix(m)


@pytest.mark.parametrize("name", sorted(SOLVE))
def test_approx_and_mpg_fixture(oracle, name):
    lp = SOLVE[name]
    f = load(f"approx_{name}.npz")
    assert str(f["digest"]) == lp.digest()
    m = oracle.make_matrix(lp, 0)
    a = oracle.approx_make(m, lp)
    for solver in ("approx", "mpg"):
        fn = oracle.approx_solve if solver == "approx" else oracle.mpg_solve
        for cap in f["caps"]:
            rc, x, diag, log = fn(a, np.zeros(lp.ncols), int(cap), log=True,
                                  period=10, offset=0.5)
            assert rc == int(f[f"{solver}_rc_{cap}"])
            assert bit_equal(x, f[f"{solver}_x_{cap}"])
            assert bit_equal(diag, f[f"{solver}_diag_{cap}"])
            assert log == str(f[f"{solver}_log_{cap}"])
    rc, x, diag, log = oracle.approx_solve(a, np.zeros(lp.ncols), 100000,
                                           max_pg=1e-3, min_delta=1e-12,
                                           log=True, period=1000)
    assert rc == int(f["approx_pg_rc"]) and bit_equal(x, f["approx_pg_x"])
    assert bit_equal(diag, f["approx_pg_diag"]) and log == str(f["approx_pg_log"])
    oracle.approx_free(a)
    oracle.free_matrix(m)


@pytest.mark.parametrize("name", sorted(SOLVE))
def test_alm_fixture(oracle, name):
    lp = SOLVE[name]
    f = load(f"alm_{name}.npz")
    m = oracle.make_matrix(lp, 0)
    alm = oracle.alm_make(m, lp)
    rc, x, lam, diag, log = oracle.alm_solve(alm, 1000, np.zeros(lp.ncols),
                                             np.zeros(lp.nrows), log=True)
    assert rc == int(f["rc"])
    assert bit_equal(x, f["x"]) and bit_equal(lam, f["lam"])
    assert bit_equal(diag, f["diag"])
    assert log == str(f["log"])
    oracle.alm_free(alm)
    oracle.free_matrix(m)


def test_cd_fixture(oracle):
    lp = SOLVE["small"]
    f = load("cd_small.npz")
    m = oracle.make_matrix(lp, 0)
    a = oracle.approx_make(m, lp)
    ctypes.CDLL(None).srandom(1)
    rc, x, diag, log = oracle.cd_solve(a, np.zeros(lp.ncols), 5, log=True,
                                       period=10)
    assert rc == int(f["rc"]) and bit_equal(x, f["x"])
    assert bit_equal(diag, f["diag"]) and log == str(f["log"])
    oracle.approx_free(a)
    oracle.free_matrix(m)


# ---- closed forms ----

def test_morton_key_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        c, r = (int(v) for v in rng.integers(0, 2**32, size=2))
        want = 0
        for b in range(32):
            want |= ((c >> b) & 1) << (2 * b)
            want |= ((r >> b) & 1) << (2 * b + 1)
        assert oracle.morton_key(c, r) == want


def test_triplets_are_z_ordered_and_permutation_is_windowed_count_sort(oracle):
    lp = I.random_lp(700, 1100, 5, 8)
    m = oracle.make_matrix(lp, 1)
    r, c, _ = oracle.coo(m)
    keys = [oracle.morton_key(int(cc), int(rr)) for rr, cc in zip(r, c)]
    assert all(a < b for a, b in zip(keys, keys[1:]))
    for row, n, orig in ((1, lp.nrows, lp.rows), (0, lp.ncols, lp.cols)):
        idx, ridx = oracle.perm(m, row)
        count = np.bincount(orig, minlength=n)
        assert np.array_equal(ridx[idx], np.arange(n))
        for w in range(0, n, 256):
            win = np.arange(w, min(n, w + 256))
            want = win[np.lexsort((win, count[win]))]   # by count, then index
            assert np.array_equal(idx[w:w + 256], want)
    oracle.free_matrix(m)


def test_csr_and_products_against_dense(oracle):
    lp = I.random_lp(40, 90, 6, 4)
    m = oracle.make_matrix(lp, 0)
    d = lp.dense()
    for t in (0, 1):
        ptr, col, val = oracle.csr(m, t)
        dd = d.T if t else d
        for i in range(dd.shape[0]):
            cols = col[ptr[i]:ptr[i + 1]]
            assert np.all(np.diff(cols.astype(np.int64)) > 0)
            assert np.array_equal(val[ptr[i]:ptr[i + 1]], dd[i, cols])
    x = np.random.default_rng(1).standard_normal(90)
    np.testing.assert_allclose(oracle.multiply(m, x, 0), d @ x, rtol=1e-13, atol=1e-13)
    oracle.free_matrix(m)


def test_step_sizes_closed_form(oracle):
    lp = I.random_lp(50, 120, 3, 17)
    m = oracle.make_matrix(lp, 0)
    w = np.random.default_rng(2).uniform(0.5, 3, 50)
    a = oracle.approx_make(m, lp, weight=w)
    v, iv = oracle.step_sizes(a, 120, permuted=0)
    d = lp.dense()
    beta = (d != 0).sum(axis=1)
    want = ((w * beta)[:, None] * d * d).sum(axis=0)
    np.testing.assert_allclose(v, want, rtol=1e-14)
    np.testing.assert_allclose(iv, 1 / want, rtol=1e-14)
    oracle.approx_free(a)
    oracle.free_matrix(m)


def test_tiny_lp_optimum(oracle):
    lp = I.tiny_lp()
    m = oracle.make_matrix(lp, 0)
    alm = oracle.alm_make(m, lp)
    rc, x, lam, diag, _ = oracle.alm_solve(alm, 200, np.zeros(2), np.zeros(1))
    assert rc == 0
    np.testing.assert_allclose(x, [1.0, 0.0], atol=2e-5)
    assert abs(diag[0] - 1.0) < 2e-5
    oracle.alm_free(alm)
    oracle.free_matrix(m)


@pytest.mark.parametrize("name", sorted(SOLVE))
def test_alm_reaches_highs_optimum(oracle, name):
    """With tightened stopping constants the ALM reaches the LP optimum
    that HiGHS computed independently (fixture lp_optimum.npz)."""
    lp = SOLVE[name]
    opt = float(load("lp_optimum.npz")[name])
    m = oracle.make_matrix(lp, 0)
    alm = oracle.alm_make(m, lp)
    rc, x, lam, diag, _ = oracle.alm_solve(
        alm, 5000, np.zeros(lp.ncols), np.zeros(lp.nrows),
        stop_violation=1e-9, stop_pg=1e-9, precision_floor=1e-12)
    assert abs(diag[0] - opt) <= 1e-6 * max(1.0, abs(opt)), (diag[0], opt)
    oracle.alm_free(alm)
    oracle.free_matrix(m)


def test_kkt_closed_form(oracle):
    """ora_alm_kkt on the hand-solved LP min x1 + 2 x2 s.t. x1 + x2 = 1,
    0 <= x <= 10: zero at the optimum x = (1, 0), lambda = 1 (reduced costs
    d = (0, 1)); at x = (0.9, 0) the primal residual is 0.1 / (1 + 1) =
    0.05, the gap |0.9 - 1| / (1 + 0.9 + 1) = 0.0345, the dual residual 0,
    so KKT = 0.05."""
    lp = I.tiny_lp()
    m = oracle.make_matrix(lp, 0)
    a = oracle.alm_make(m, lp)
    assert oracle.alm_kkt(a, np.array([1.0, 0.0]), np.array([1.0])) == 0.0
    k = oracle.alm_kkt(a, np.array([0.9, 0.0]), np.array([1.0]))
    assert abs(k - 0.05) <= 1e-15, k
    # lambda = 1.5: d = (-0.5, 0.5); x = (1, 0) is feasible (r_p = 0);
    # clamp(x - d) = (1.5, 0), so r_d = 0.5 / (1 + sqrt(5)); the dual bound
    # b lambda + sum(l d+ - u d-) = 1.5 - 10 * 0.5 = -3.5 against c'x = 1
    # gives r_gap = 4.5 / (1 + 1 + 3.5) = 9/11, the maximum.
    k = oracle.alm_kkt(a, np.array([1.0, 0.0]), np.array([1.5]))
    assert abs(k - 9.0 / 11.0) <= 1e-15, k
    oracle.alm_free(a)
    oracle.free_matrix(m)


def test_warm_start_equals_the_reference_build(oracle):
    """Edits through the accessors, approx_update and a warm approx_solve,
    then an ALM warm re-solve after editing costs and bounds: the oracle
    and the reference built from its own sources agree bit for bit."""
    import dataclasses
    import pyoracle
    from acdalm.capi import CLib
    if not os.path.exists(pyoracle.REF_APPROX_SO):
        pytest.skip("reference build absent (oracle/_ref)")
    ref = CLib(pyoracle.REF_APPROX_SO, gpu_extras=False)
    lp = SOLVE["small"]
    rng = np.random.default_rng(7)
    w = np.linspace(0.5, 2.0, lp.nrows)
    rhs2 = lp.rhs * 1.05
    lin2 = lp.linear + rng.uniform(-0.3, 0.3, lp.ncols)
    up2 = lp.upper.copy()
    up2[rng.choice(lp.ncols, 10, replace=False)] = 7.5
    view = lambda p, n: np.ctypeslib.as_array(p, (n,))
    mr, mo = ref.make_matrix(lp, 0), oracle.make_matrix(lp, 0)
    ar, ao = ref.approx_make(mr, lp, weight=w), oracle.approx_make(mo, lp, weight=w)
    r1 = ref.approx_solve(ar, np.zeros(lp.ncols), 60)
    o1 = oracle.approx_solve(ao, np.zeros(lp.ncols), 60)
    assert r1[0] == o1[0] and bit_equal(r1[1], o1[1])
    R, O = ref.lib, oracle.lib
    for gr, go, vals, n in [(R.approx_rhs, O.ora_approx_rhs, rhs2, lp.nrows),
                            (R.approx_linear, O.ora_approx_linear, lin2, lp.ncols),
                            (R.approx_upper, O.ora_approx_upper, up2, lp.ncols)]:
        view(gr(ar), n)[:] = vals
        view(go(ao), n)[:] = vals
    R.approx_update(ar)
    O.ora_approx_update(ao)
    r2 = ref.approx_solve(ar, r1[1], 60)
    o2 = oracle.approx_solve(ao, o1[1], 60)
    assert r2[0] == o2[0] and bit_equal(r2[1], o2[1])
    assert bit_equal(r2[2], o2[2])
    ref.approx_free(ar)
    oracle.approx_free(ao)
    # ALM warm re-solve
    lr, lo = ref.alm_make(mr, lp), oracle.alm_make(mo, lp)
    c1 = ref.alm_solve(lr, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    d1 = oracle.alm_solve(lo, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    assert bit_equal(c1[1], d1[1]) and bit_equal(c1[2], d1[2])
    view(R.alm_linear(lr), lp.ncols)[:] = lin2
    view(R.alm_upper(lr), lp.ncols)[:] = up2
    view(O.ora_alm_linear(lo), lp.ncols)[:] = lin2
    view(O.ora_approx_upper(O.ora_alm_approx(lo)), lp.ncols)[:] = up2
    c2 = ref.alm_solve(lr, 1000, c1[1], c1[2])
    d2 = oracle.alm_solve(lo, 1000, d1[1], d1[2])
    assert c2[0] == d2[0]
    assert bit_equal(c2[1], d2[1]) and bit_equal(c2[2], d2[2])
    assert bit_equal(c2[3], d2[3])
    ref.alm_free(lr)
    oracle.alm_free(lo)
    ref.free_matrix(mr)
    oracle.free_matrix(mo)


def test_threaded_kkt_solves_match_serial(oracle, tmp_path):
    """The CPU baseline (ora_bench alm) solves several LPs at once, one per
    thread, each stopping on the KKT target with the inner-checkpoint test:
    every LP meets the target and ends where the same solve on one thread
    (the Python binding) ends."""
    import json
    import subprocess
    import pyoracle
    bench = os.path.join(os.path.dirname(pyoracle.ORACLE_SO), "ora_bench")
    lps = I.batch(12, seed=11, mmin=30, mmax=150)
    files = []
    for k, lp in enumerate(lps):
        f = str(tmp_path / f"lp{k}.lpb")
        lp.write_binary(f)
        files.append(f)
    r = subprocess.run([bench, "alm", "4", "1e-6"] + files, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-500:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert max(out["kkt"]) <= 1e-6
    for lp, obj in zip(lps, out["objective"]):
        m = oracle.make_matrix(lp, 0)
        a = oracle.alm_make(m, lp)
        rc, x, lam, diag, _ = oracle.alm_solve(a, 1000, np.zeros(lp.ncols),
                                               np.zeros(lp.nrows), stop_kkt=1e-6)
        oracle.alm_free(a)
        oracle.free_matrix(m)
        assert rc == 0
        assert diag[0] == obj
