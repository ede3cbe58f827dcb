"""Relative KKT residual of an LP point (x, lambda).

The reference stops on absolute thresholds (alm.c:306-308) and defines no
KKT measure; this one is used to report convergence of GPU and CPU results
alike.  With the reference's sign convention (Lagrangian c'x - lambda'(Ax-b),
alm.c:106-119) the reduced costs are d = c - A'lambda, and

  r_p = |Ax - b|_2 / (1 + |b|_2)
  r_d = |x - clamp(x - d, l, u)|_2 / (1 + |c|_2)        (+ dual infeasibility)
  r_g = |c'x - dual| / (1 + |c'x| + |dual|),
        dual = b'lambda + sum_i (l_i d_i+ - u_i d_i-)

where a reduced cost pushing against an infinite bound cannot be priced and
is counted as dual infeasibility instead.  KKT = max(r_p, r_d, r_g).
"""
import numpy as np


def matvec(lp, x):
    return np.bincount(lp.rows, weights=lp.vals * x[lp.cols], minlength=lp.nrows)


def rmatvec(lp, y):
    return np.bincount(lp.cols, weights=lp.vals * y[lp.rows], minlength=lp.ncols)


def kkt(lp, x, lam):
    x = np.asarray(x, float)
    lam = np.asarray(lam, float)
    r = matvec(lp, x) - lp.rhs
    rp = np.linalg.norm(r) / (1 + np.linalg.norm(lp.rhs))
    d = lp.linear - rmatvec(lp, lam)
    proj = np.clip(x - d, lp.lower, lp.upper)
    dpos, dneg = np.maximum(d, 0), np.maximum(-d, 0)
    lo_inf, up_inf = ~np.isfinite(lp.lower), ~np.isfinite(lp.upper)
    dual_infeas = np.linalg.norm(np.where(lo_inf, dpos, 0) +
                                 np.where(up_inf, dneg, 0))
    rd = (np.linalg.norm(x - proj) + dual_infeas) / (1 + np.linalg.norm(lp.linear))
    lo = np.where(lo_inf, 0.0, lp.lower)
    up = np.where(up_inf, 0.0, lp.upper)
    dual = lp.rhs @ lam + np.sum(lo * dpos) - np.sum(up * dneg)
    primal = lp.linear @ x
    rg = abs(primal - dual) / (1 + abs(primal) + abs(dual))
    return {"kkt": float(max(rp, rd, rg)), "r_primal": float(rp),
            "r_dual": float(rd), "r_gap": float(rg), "primal": float(primal),
            "dual": float(dual)}
