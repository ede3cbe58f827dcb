"""Batched solves (alm_solve_batch): many independent LPs, one workgroup
each, in repeated launches (32 inner iterations while the tuner compares
workgroup sizes, then to completion).

Each LP's arithmetic depends only on that LP (its own workgroup, its own
fixed-order reductions), so a result is bit-identical alone and in any
mix; against the single-LP solver (grid-wide reductions in another
order) and the reference's golden runs it agrees to the stopping rule.
"""
import os

import numpy as np
import pytest

from acdalm import instances as I
from acdalm.kkt import kkt

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def make(gpu, lps):
    mats = [gpu.make_matrix(lp, 0) for lp in lps]
    alms = [gpu.alm_make(m, lp) for m, lp in zip(mats, lps)]
    return mats, alms


def free(gpu, mats, alms):
    for a in alms:
        gpu.alm_free(a)
    for m in mats:
        gpu.free_matrix(m)


def batch(gpu, lps, precision, niter=1000):
    mats, alms = make(gpu, lps)
    out, done, info = gpu.alm_solve_batch(
        alms, niter, [np.zeros(lp.ncols) for lp in lps],
        [np.zeros(lp.nrows) for lp in lps], precision=precision)
    free(gpu, mats, alms)
    return out, done, info


def same(r1, r2):
    return (r1[0] == r2[0] and np.array_equal(r1[1], r2[1])
            and np.array_equal(r1[2], r2[2]) and np.array_equal(r1[3], r2[3]))


@pytest.mark.parametrize("precision", [0, 1])
def test_result_is_the_same_alone_and_in_the_mix(gpu, precision):
    lps = I.batch(12, seed=5, mmin=20, mmax=150)
    print(f"[batch] mix precision={precision}", flush=True)
    mixed, done, info = batch(gpu, lps, precision)
    assert info.launches > 0 and np.all(done > 0)
    for i in (0, 5, 11):
        alone, _, _ = batch(gpu, [lps[i]], precision)
        assert same(mixed[i], alone[0]), i
    # another mix: reversed order with two extra LPs in front
    extra = I.batch(2, seed=6, mmin=20, mmax=150)
    other, _, _ = batch(gpu, extra + lps[::-1], precision)
    for i in range(len(lps)):
        assert same(mixed[i], other[2 + len(lps) - 1 - i]), i


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_agrees_with_single_solver(gpu, precision):
    """Against the multi-workgroup solver (engine=1), whose reductions run
    in another order."""
    lps = I.batch(8, seed=7, mmin=30, mmax=200)
    print(f"[batch] vs single precision={precision}", flush=True)
    res, _, _ = batch(gpu, lps, precision)
    for lp, (ret, x, lam, diag) in zip(lps, res):
        m = gpu.make_matrix(lp, 0)
        a = gpu.alm_make(m, lp)
        rs, xs, ls, ds, _ = gpu.alm_solve_ex(a, 1000, np.zeros(lp.ncols),
                                             np.zeros(lp.nrows),
                                             precision=precision, engine=1)
        gpu.alm_free(a)
        gpu.free_matrix(m)
        assert ret == rs == 0
        assert abs(diag[0] - ds[0]) <= 1e-6 * max(1.0, abs(ds[0]))
        k, ks = kkt(lp, x, lam)["kkt"], kkt(lp, xs, ls)["kkt"]
        assert k <= max(1e-6, 10 * ks), (k, ks)


def test_batch_fp64_matches_reference_objectives(gpu):
    cases = {"small": I.random_lp(60, 150, 4, 21),
             "setcover": I.set_cover(80, 300, 6),
             "mcf": I.multicommodity(4, 2, 7)}
    names = sorted(cases)
    print("[batch] reference objectives", flush=True)
    res, _, _ = batch(gpu, [cases[n] for n in names], 0)
    for n, (ret, x, lam, diag) in zip(names, res):
        ref = np.load(os.path.join(GOLD, f"alm_{n}.npz"), allow_pickle=False)
        assert ret == int(ref["rc"]) == 0
        assert abs(diag[0] - ref["diag"][0]) <= 1e-6 * max(1.0, abs(ref["diag"][0]))
        assert diag[1] < 1e-5


def test_batch_kkt_stopping_rule(gpu):
    lps = I.batch(6, seed=8, mmin=30, mmax=200)
    print("[batch] stop_kkt", flush=True)
    mats, alms = make(gpu, lps)
    out, done, info = gpu.alm_solve_batch(
        alms, 1000, [np.zeros(lp.ncols) for lp in lps],
        [np.zeros(lp.nrows) for lp in lps], precision=0, stop_kkt=1e-6)
    free(gpu, mats, alms)
    for lp, (ret, x, lam, diag) in zip(lps, out):
        assert ret == 0
        assert kkt(lp, x, lam)["kkt"] <= 1e-6


def test_a_failing_lp_is_reported_and_the_others_are_unaffected(gpu):
    """A NaN cost makes that LP's iterate non-finite: its return code is
    -ACD_SOLVE_UNBOUNDED while the other LPs solve as if alone."""
    lps = I.batch(4, seed=9, mmin=30, mmax=100)
    bad = I.random_lp(40, 100, 4, 10)
    bad.linear = bad.linear.copy()
    bad.linear[7] = np.nan
    print("[batch] failing LP", flush=True)
    res, _, _ = batch(gpu, lps[:2] + [bad] + lps[2:], 0)
    assert res[2][0] == -1
    ok = [r for i, r in enumerate(res) if i != 2]
    for lp, r in zip(lps, ok):
        alone, _, _ = batch(gpu, [lp], 0)
        assert same(r, alone[0])


def test_single_small_lp_routes_to_the_workgroup_solver(gpu):
    """alm_solve_ex picks the one-workgroup solver for small LPs solved
    without a log; its result is the batch-of-one result."""
    lp = I.batch(1, seed=12, mmin=60, mmax=80)[0]
    m = gpu.make_matrix(lp, 0)
    a = gpu.alm_make(m, lp)
    r_auto = gpu.alm_solve_ex(a, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    gpu.lib.acd_alm_reset(a)
    r_wg = gpu.alm_solve_ex(a, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows),
                            engine=2)
    gpu.alm_free(a)
    gpu.free_matrix(m)
    alone, _, _ = batch(gpu, [lp], 0)
    assert same(r_auto[:4], r_wg[:4]) and same(r_auto[:4], alone[0])


def test_writing_a_log_does_not_change_the_result(gpu):
    """The engine follows the LP's size only: a small LP runs on the
    one-workgroup solver with or without a log (the log is replayed from
    its event ring), so logging leaves x, lambda and the diagnosis
    unchanged; the log has one line per outer iteration in the
    reference's format (alm.c:266-285)."""
    import re
    lp = I.batch(1, seed=15, mmin=60, mmax=80)[0]
    m = gpu.make_matrix(lp, 0)
    a = gpu.alm_make(m, lp)
    print("[batch] log vs no log", flush=True)
    quiet = gpu.alm_solve_ex(a, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    gpu.lib.acd_alm_reset(a)
    loud = gpu.alm_solve_ex(a, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows),
                            log=True)
    gpu.alm_free(a)
    gpu.free_matrix(m)
    assert same(quiet[:4], loud[:4])
    outer = [ln for ln in loud[4].splitlines() if re.match(r"^\s*\d+: ", ln)]
    assert len(outer) >= 1
    for ln in outer:
        assert ln.rstrip().endswith("-stalled-") or re.search(r"\(\s*\S+\)$", ln.rstrip()), ln


def test_result_does_not_depend_on_the_workgroup_size(gpu, monkeypatch):
    """Every op is reduced as a fixed number of virtual waves per LP (4, or
    8 for the largest size class: operators of more than 64 slices), so
    workgroups of 1, 2, 4 (or 8) wavefronts give bit-identical results."""
    lps = I.batch(10, seed=13, mmin=20, mmax=600)
    lps.append(I.random_lp(1800, 4800, 5, 17))  # 75 slices: 8 virtual waves
    outs = []
    for w in (1, 2, 4, 8):
        monkeypatch.setenv("ACD_BATCH_WAVES", str(w))
        print(f"[batch] waves={w}", flush=True)
        out, _, info = batch(gpu, lps, 1)
        # 8 physical waves only where there are 8 virtual ones
        assert set(info.waves_per_lp) <= ({0, 4, 8} if w == 8 else {0, w})
        assert info.waves_per_lp[7] == w
        outs.append(out)
    for o in outs[1:]:
        for i, (r0, r) in enumerate(zip(outs[0], o)):
            assert same(r0, r), i


def test_tuned_batch_gives_the_forced_results(gpu, monkeypatch):
    """A batch large enough to tune switches workgroup sizes between rounds
    (one trial round per candidate); its results are those of a fixed
    size."""
    lps = I.batch(200, seed=14, mmin=20, mmax=200)
    monkeypatch.delenv("ACD_BATCH_WAVES", raising=False)
    print("[batch] tuned 200", flush=True)
    tuned, _, info = batch(gpu, lps, 1)
    assert any(v in (1, 2, 4) for v in info.waves_per_lp)
    assert all(v in (0, 1, 2, 4, 8) for v in info.waves_per_lp)
    monkeypatch.setenv("ACD_BATCH_WAVES", "4")
    sample = [0, 7, 50, 123, 199]
    fixed, _, _ = batch(gpu, [lps[i] for i in sample], 1)
    for k, i in enumerate(sample):
        assert same(tuned[i], fixed[k]), i


def test_16_bit_columns_give_the_same_results(gpu, monkeypatch):
    """In mixed precision, a batch whose operators all have at most 65,536
    columns reads 16-bit column indices (6 bytes per entry instead of 8);
    the products, and so the whole trajectories, are bit-identical to those
    with 32-bit indices, at 2 bytes fewer per entry and matrix pass."""
    lps = I.batch(12, seed=21, mmin=20, mmax=900)
    lps.append(I.random_lp(1800, 4800, 5, 17))  # 8 virtual waves
    monkeypatch.delenv("ACD_BATCH_WAVES", raising=False)
    monkeypatch.setenv("ACD_BATCH_COL16", "0")
    print("[batch] 32-bit columns", flush=True)
    wide, _, info32 = batch(gpu, lps, 1)
    monkeypatch.delenv("ACD_BATCH_COL16")
    print("[batch] 16-bit columns", flush=True)
    narrow, _, info16 = batch(gpu, lps, 1)
    for i, (a, b) in enumerate(zip(wide, narrow)):
        assert a[0] == 0, i
        assert same(a, b), i
    assert info16.inner_iterations == info32.inner_iterations
    assert info16.alg_bytes < info32.alg_bytes
