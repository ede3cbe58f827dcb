"""GPU parity of cd_solve, the randomised coordinate descent that
pkhuong/acd-alm ships as its ALM inner solver (approx/cd.c:27-385).

Every sum on the device path runs in the reference's order and the sampled
column order is drawn from the process's random() exactly as the original
does, so x, the diagnosis vector and the log are compared bit for bit with
the reference's golden run (srandom(1)) and with the oracle under the same
seed.
"""
import ctypes
import os

import numpy as np
import pytest

from acdalm import instances as I
from helpers import bit_equal

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIBC = ctypes.CDLL(None)


def run(lib, lp, seed, niter, x0=None, weight=None, **kw):
    m = lib.make_matrix(lp, 0)
    a = lib.approx_make(m, lp, weight=weight)
    LIBC.srandom(seed)
    x0 = np.zeros(lp.ncols) if x0 is None else x0
    out = lib.cd_solve(a, x0, niter, **kw)
    lib.approx_free(a)
    lib.free_matrix(m)
    return out


def test_cd_matches_reference_fixture(gpu):
    lp = I.random_lp(60, 150, 4, 21)
    f = np.load(os.path.join(GOLD, "cd_small.npz"), allow_pickle=False)
    assert str(f["digest"]) == lp.digest()
    rc, x, diag, log = run(gpu, lp, 1, 5, log=True, period=10)
    assert rc == int(f["rc"])
    assert bit_equal(x, f["x"]) and bit_equal(diag, f["diag"])
    assert log == str(f["log"])


@pytest.mark.parametrize("seed", [3, 11])
def test_cd_matches_oracle_with_weights_and_start(gpu, oracle, seed):
    lp = I.random_lp(90, 240, 5, seed)
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 4.0, lp.nrows)
    x0 = rng.uniform(-1.0, 12.0, lp.ncols)   # outside the box on purpose
    kw = dict(log=True, period=20, offset=1.5)
    got = run(gpu, lp, seed, 4, x0=x0, weight=w, **kw)
    want = run(oracle, lp, seed, 4, x0=x0, weight=w, **kw)
    assert got[0] == want[0]
    assert bit_equal(got[1], want[1]) and bit_equal(got[2], want[2])
    assert got[3] == want[3]


def test_cd_stops_and_zero_iterations(gpu, oracle):
    lp = I.set_cover(80, 300, 6)
    for kw in (dict(niter=0), dict(niter=50, max_pg=1e30),
               dict(niter=50, max_value=1e30)):
        got = run(gpu, lp, 5, **kw)
        want = run(oracle, lp, 5, **kw)
        assert got[0] == want[0]
        assert bit_equal(got[1], want[1]) and bit_equal(got[2], want[2])
