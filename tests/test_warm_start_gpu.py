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
