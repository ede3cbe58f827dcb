"""Re-solving after edits through the mutable accessors (approx.h and
alm.h accessors; approx_update, approx/approx.c:188-215): the column-
generation workflow of the reference's note changes costs, bounds and
right-hand sides between solves and restarts from the previous point.

The device copy must then be the edited problem exactly: a warm re-solve
after approx_update is bit-identical to a fresh object made from the
edited data, started from the same point; an ALM re-solve after editing
alm_linear / alm_upper reaches the edited LP's optimum."""
import dataclasses

import numpy as np
import pytest

from acdalm import instances as I
from acdalm.kkt import kkt

pytestmark = pytest.mark.gpu


def view(ptr, n):
    return np.ctypeslib.as_array(ptr, (n,))


def edited(lp, seed):
    rng = np.random.default_rng(seed)
    rhs = lp.rhs * 1.05
    lin = lp.linear + rng.uniform(-0.3, 0.3, lp.ncols) * (rng.random(lp.ncols) < 0.2)
    up = lp.upper.copy()
    up[rng.choice(lp.ncols, 10, replace=False)] = 7.5
    return dataclasses.replace(lp, rhs=rhs, linear=lin, upper=up)


def test_approx_update_equals_a_fresh_object(gpu):
    lp = I.random_lp(200, 500, 5, 31)
    w = np.linspace(0.5, 2.0, lp.nrows)
    m = gpu.make_matrix(lp, 0)
    a = gpu.approx_make(m, lp, weight=w)
    print("[warm] approx cold", flush=True)
    _, x1, _, _ = gpu.approx_solve(a, np.zeros(lp.ncols), 300)
    lp2 = edited(lp, 5)
    w2 = w.copy()
    w2[::7] *= 3.0
    L = gpu.lib
    view(L.approx_rhs(a), lp.nrows)[:] = lp2.rhs
    view(L.approx_weight(a), lp.nrows)[:] = w2
    view(L.approx_linear(a), lp.ncols)[:] = lp2.linear
    view(L.approx_upper(a), lp.ncols)[:] = lp2.upper
    gpu.approx_update(a)
    print("[warm] approx warm", flush=True)
    r2, x2, d2, _ = gpu.approx_solve(a, x1, 300)
    fresh = gpu.approx_make(m, lp2, weight=w2)
    rf, xf, df, _ = gpu.approx_solve(fresh, x1, 300)
    gpu.approx_free(fresh)
    gpu.approx_free(a)
    gpu.free_matrix(m)
    assert r2 == rf
    assert np.array_equal(x2, xf)
    assert np.array_equal(d2, df, equal_nan=True)


def test_alm_resolve_after_editing_costs_and_bounds(gpu):
    lp = I.random_lp(150, 400, 5, 32)
    m = gpu.make_matrix(lp, 0)
    alm = gpu.alm_make(m, lp)
    print("[warm] alm cold", flush=True)
    rc, x1, l1, d1, _ = gpu.alm_solve_ex(alm, 1000, np.zeros(lp.ncols),
                                         np.zeros(lp.nrows), stop_kkt=1e-7)
    assert rc == 0 and kkt(lp, x1, l1)["kkt"] <= 2e-7
    cold_inner = gpu.stats().inner_iterations
    lp2 = edited(lp, 6)
    L = gpu.lib
    view(L.alm_rhs(alm), lp.nrows)[:] = lp2.rhs
    view(L.alm_linear(alm), lp.ncols)[:] = lp2.linear
    view(L.alm_upper(alm), lp.ncols)[:] = lp2.upper
    print("[warm] alm warm", flush=True)
    rc2, x2, l2, d2, _ = gpu.alm_solve_ex(alm, 1000, x1, l1, stop_kkt=1e-7)
    warm_inner = gpu.stats().inner_iterations
    assert rc2 == 0
    k2 = kkt(lp2, x2, l2)
    assert k2["kkt"] <= 2e-7, k2
    # warm start with the penalty weights restored (acd_alm_reset): the
    # re-solve above continues the first solve's weights, as the
    # reference's alm_solve does
    L.acd_alm_reset(alm)
    rc4, x4, l4, d4, _ = gpu.alm_solve_ex(alm, 1000, x1, l1, stop_kkt=1e-7)
    reset_inner = gpu.stats().inner_iterations
    assert rc4 == 0 and kkt(lp2, x4, l4)["kkt"] <= 2e-7
    # the same LP solved cold on a fresh object: same optimum
    alm3 = gpu.alm_make(m, lp2)
    rc3, x3, l3, d3, _ = gpu.alm_solve_ex(alm3, 1000, np.zeros(lp.ncols),
                                          np.zeros(lp.nrows), stop_kkt=1e-7)
    fresh_inner = gpu.stats().inner_iterations
    gpu.alm_free(alm3)
    gpu.alm_free(alm)
    gpu.free_matrix(m)
    assert rc3 == 0
    assert abs(d2[0] - d3[0]) <= 1e-6 * max(1.0, abs(d3[0]))
    assert abs(d4[0] - d3[0]) <= 1e-6 * max(1.0, abs(d3[0]))
    print(f"[warm] inner iterations: first {cold_inner}, warm re-solve "
          f"{warm_inner}, warm after reset {reset_inner}, fresh cold "
          f"{fresh_inner}", flush=True)


def test_warm_start_matches_the_oracle(gpu, oracle):
    """The same edits through the accessors on the device library and on
    the oracle (the reference's algorithm, pinned bit for bit to it),
    followed by approx_update and a warm approx_solve from the previous
    point: both land on the same point (short runs, whose trajectories do
    not depend on the summation order, agree to 1e-9 as in the cold
    solver tests)."""
    lp = I.random_lp(60, 150, 4, 21)
    w = np.linspace(0.5, 2.0, lp.nrows)
    mg, mo = gpu.make_matrix(lp, 0), oracle.make_matrix(lp, 0)
    ag, ao = gpu.approx_make(mg, lp, weight=w), oracle.approx_make(mo, lp, weight=w)
    print("[warm] oracle: cold", flush=True)
    rg, xg, dg, _ = gpu.approx_solve(ag, np.zeros(lp.ncols), 60)
    ro, xo, do, _ = oracle.approx_solve(ao, np.zeros(lp.ncols), 60)
    scale = max(1.0, float(np.max(np.abs(xo))))
    assert rg == ro and np.max(np.abs(xg - xo)) <= 1e-9 * scale
    lp2 = edited(lp, 7)
    w2 = w.copy()
    w2[::5] *= 2.0
    G, O = gpu.lib, oracle.lib
    for get_g, get_o, vals, n in [
            (G.approx_rhs, O.ora_approx_rhs, lp2.rhs, lp.nrows),
            (G.approx_weight, O.ora_approx_weight, w2, lp.nrows),
            (G.approx_linear, O.ora_approx_linear, lp2.linear, lp.ncols),
            (G.approx_upper, O.ora_approx_upper, lp2.upper, lp.ncols)]:
        view(get_g(ag), n)[:] = vals
        view(get_o(ao), n)[:] = vals
    gpu.approx_update(ag)
    oracle.approx_update(ao)
    print("[warm] oracle: warm", flush=True)
    rg2, xg2, dg2, _ = gpu.approx_solve(ag, xg, 60)
    ro2, xo2, do2, _ = oracle.approx_solve(ao, xo, 60)
    scale = max(1.0, float(np.max(np.abs(xo2))))
    assert rg2 == ro2
    assert np.max(np.abs(xg2 - xo2)) <= 1e-9 * scale
    assert np.allclose(dg2[:4], do2[:4], rtol=1e-9, atol=1e-12)
    gpu.approx_free(ag)
    oracle.approx_free(ao)
    gpu.free_matrix(mg)
    oracle.free_matrix(mo)


def test_alm_warm_resolve_matches_the_oracle(gpu, oracle):
    """ALM: solve, edit the LP costs and upper bounds through alm_linear /
    alm_upper, and solve again from the previous (x, lambda) with the
    penalty weights carried over (as the reference does), on the device and
    on the oracle: the same objective (to the reference's stopping rule,
    1e-6 relative, as in the cold ALM tests) and the edited LP's optimum."""
    from scipy.optimize import linprog
    lp = I.random_lp(60, 150, 4, 21)
    mg, mo = gpu.make_matrix(lp, 0), oracle.make_matrix(lp, 0)
    ag, ao = gpu.alm_make(mg, lp), oracle.alm_make(mo, lp)
    print("[warm] oracle: alm cold", flush=True)
    rg, xg, lg, dg, _ = gpu.alm_solve(ag, 1000, np.zeros(lp.ncols),
                                      np.zeros(lp.nrows))
    ro, xo, lo, do, _ = oracle.alm_solve(ao, 1000, np.zeros(lp.ncols),
                                         np.zeros(lp.nrows))
    assert rg == ro == 0
    lp2 = edited(lp, 8)
    view(gpu.lib.alm_linear(ag), lp.ncols)[:] = lp2.linear
    view(gpu.lib.alm_upper(ag), lp.ncols)[:] = lp2.upper
    view(oracle.lib.ora_alm_linear(ao), lp.ncols)[:] = lp2.linear
    view(oracle.lib.ora_approx_upper(oracle.lib.ora_alm_approx(ao)),
         lp.ncols)[:] = lp2.upper
    print("[warm] oracle: alm warm", flush=True)
    rg2, xg2, lg2, dg2, _ = gpu.alm_solve(ag, 1000, xg, lg)
    ro2, xo2, lo2, do2, _ = oracle.alm_solve(ao, 1000, xo, lo)
    assert rg2 == ro2 == 0
    assert abs(dg2[0] - do2[0]) <= 1e-6 * max(1.0, abs(do2[0])), (dg2, do2)
    bounds = list(zip(lp2.lower, lp2.upper))
    opt = linprog(lp2.linear, A_eq=lp.dense(), b_eq=lp.rhs, bounds=bounds,
                  method="highs")
    assert opt.status == 0
    assert abs(dg2[0] - opt.fun) <= 1e-5 * max(1.0, abs(opt.fun))
    gpu.alm_free(ag)
    oracle.alm_free(ao)
    gpu.free_matrix(mg)
    oracle.free_matrix(mo)
