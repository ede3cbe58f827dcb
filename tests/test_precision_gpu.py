"""Reduced-precision working copies, each checked for its error against the
FP64 path on the same seeded instance (the reference's own FP64 objective
under the same stopping rule, from the golden fixtures, and the HiGHS
optimum).

ACD_MIXED stores the matrix values in BF16 when every value is exact there
(0/+-1 set-cover and flow matrices; the result is then the FP64 problem's)
or in FP32 otherwise (random coefficients), and the working copy's costs,
bounds and step sizes in FP32; iterates, rhs, weights, multipliers and all
accumulation stay FP64, and the ALM residual reads the same stored values
as the inner solver.  Stated error: objective within 1e-6 relative of the
reference's FP64 objective and of the optimum, relative KKT residual
<= max(1e-6, 10x the FP64 path's).

ACD_FP32 also keeps the iterates in FP32.  Its rounding floor stops the
inner solver early (a stall at ~8 FP32 ulps of movement) and the multiplier
noise it leaves limits the KKT residual; stated error: objective within
1e-4 relative of the reference's FP64 objective, KKT residual <= 1e-2.
"""
import os

import numpy as np
import pytest

from acdalm import instances as I
from acdalm.kkt import kkt

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"small": I.random_lp(60, 150, 4, 21), "setcover": I.set_cover(80, 300, 6),
         "mcf": I.multicommodity(4, 2, 7), "c0": I.config0(1)}


def reference(name):
    """The reference's FP64 objective under the same stopping rule, and the
    LP optimum (HiGHS)."""
    f = np.load(os.path.join(GOLD, f"alm_{name}.npz"), allow_pickle=False)
    opt = np.load(os.path.join(GOLD, "lp_optimum.npz"), allow_pickle=False)
    return float(f["diag"][0]), float(opt[name])


def solve(gpu, lp, precision):
    m = gpu.make_matrix(lp, 0)
    alm = gpu.alm_make(m, lp)
    out = gpu.alm_solve_ex(alm, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows),
                           precision=precision)
    gpu.alm_free(alm)
    gpu.free_matrix(m)
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_precision_error_against_fp64(gpu, name):
    lp = CASES[name]
    ref, opt = reference(name)
    print(f"[precision] mixed {name}", flush=True)
    rc64, x64, lam64, d64, _ = solve(gpu, lp, 0)
    rc, x, lam, diag, _ = solve(gpu, lp, 1)
    k64 = kkt(lp, x64, lam64)["kkt"]
    k = kkt(lp, x, lam)["kkt"]
    rel = abs(diag[0] - ref) / max(1.0, abs(ref))
    rel_opt = abs(diag[0] - opt) / max(1.0, abs(opt))
    print(f"[precision] mixed {name}: objective error {rel:.3g} vs FP64, "
          f"{rel_opt:.3g} vs optimum; KKT {k:.3g} (FP64 {k64:.3g})",
          flush=True)
    assert rc == rc64 == 0
    assert rel <= 1e-6 and rel_opt <= 1e-6
    assert k <= max(1e-6, 10 * k64)


def test_mixed_precision_is_exact_for_bf16_matrices(gpu):
    """With a BF16-exact matrix only the FP32 costs and bounds differ from
    FP64: the solve follows the same path to within the cost rounding."""
    lp = I.set_cover(2000, 8000, 9)
    r64 = solve(gpu, lp, 0)
    rmx = solve(gpu, lp, 1)
    assert r64[0] == rmx[0] == 0
    assert abs(r64[3][0] - rmx[3][0]) <= 1e-7 * abs(r64[3][0])


@pytest.mark.parametrize("name", ["small", "c0"])
def test_fp32_iterates_error_is_bounded(gpu, name):
    lp = CASES[name]
    ref, _ = reference(name)
    print(f"[precision] fp32 {name}", flush=True)
    rc, x, lam, diag, _ = solve(gpu, lp, 2)
    k = kkt(lp, x, lam)["kkt"]
    rel = abs(diag[0] - ref) / max(1.0, abs(ref))
    print(f"[precision] fp32 {name}: objective error {rel:.3g} vs FP64, "
          f"KKT {k:.3g}", flush=True)
    assert np.isfinite(diag[0])
    assert rel <= 1e-4
    assert k <= 1e-2


@pytest.mark.parametrize("name", ["small", "c0", "mcf"])
def test_kkt_stopping_rule(gpu, name):
    """stop_kkt: the device evaluates the relative KKT residual after every
    outer iteration (the same formula as acdalm.kkt) and stops at the
    target; the host evaluation of the returned (x, lambda) agrees."""
    lp = CASES[name]
    print(f"[precision] stop_kkt {name}", flush=True)
    m = gpu.make_matrix(lp, 0)
    alm = gpu.alm_make(m, lp)
    rc, x, lam, diag, _ = gpu.alm_solve_ex(
        alm, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows), stop_kkt=1e-6)
    dev = gpu.stats().kkt
    gpu.alm_free(alm)
    gpu.free_matrix(m)
    host = kkt(lp, x, lam)["kkt"]
    assert rc == 0
    assert np.isfinite(dev) and dev <= 1e-6
    assert abs(dev - host) <= 1e-6 * host + 1e-13, (dev, host)


@pytest.mark.parametrize("name", ["setcover", "mcf"])
def test_value_free_operator_equals_bf16(gpu, name, monkeypatch):
    """0/+-1 operators in mixed precision drop their value stream (the sign
    travels in the column word): the multi-workgroup solve is the BF16
    form's bit for bit, and each A pass moves 2 bytes per nonzero less."""
    lp = {"setcover": I.set_cover(3000, 12000, 5),
          "mcf": I.multicommodity(8, 4, 3)}[name]
    print(f"[precision] value-free {name}", flush=True)
    m = gpu.make_matrix(lp, 0)
    alm = gpu.alm_make(m, lp)
    out = {}
    for form in ("bf16", "value-free"):
        if form == "bf16":
            monkeypatch.setenv("ACD_VALUE_FREE", "0")
        else:
            monkeypatch.delenv("ACD_VALUE_FREE")
        gpu.lib.acd_alm_reset(alm)
        r = gpu.alm_solve_ex(alm, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows),
                             precision=1, engine=1, stop_kkt=1e-6)
        cnt, _, by = gpu.profile_bytes()["pass2"]
        out[form] = (r, by / cnt)
    gpu.alm_free(alm)
    gpu.free_matrix(m)
    (rb, bb), (rv, bv) = out["bf16"], out["value-free"]
    assert rb[0] == rv[0] == 0
    assert np.array_equal(rb[1], rv[1]) and np.array_equal(rb[2], rv[2])
    assert np.array_equal(rb[3], rv[3])
    assert bb - bv == 2 * lp.nnz, (bb, bv, lp.nnz)
    assert kkt(lp, rv[1], rv[2])["kkt"] <= 1e-6
