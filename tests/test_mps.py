"""MPS input (acd_mps_parse, acd_alm_read_mps in include/acd_gpu.h): the
interchange format of LP test sets, read on the host into the ALM form
(slack columns for inequality and ranged rows).  The fixture
tests/golden/small.mps exercises every supported feature; its parsed form
and optimum are derived by hand in its header and below."""
import os

import numpy as np
import pytest

from acdalm import capi
from acdalm import instances as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SMALL = os.path.join(ROOT, "tests", "golden", "small.mps")
INF = np.inf


@pytest.fixture(scope="module")
def host():
    """The library bound without device initialisation (host code only)."""
    return capi.CLib(capi.LIB_PATH, init=False)


def write_mps(lp, path):
    """Free MPS of an equality-form LP: rows E, bounds LO/UP/MI/PL."""
    def num(v):
        return repr(float(v))
    order = np.lexsort((lp.rows, lp.cols))
    with open(path, "w") as f:
        f.write("NAME RANDOM\nROWS\n N OBJ\n")
        for i in range(lp.nrows):
            f.write(f" E R{i}\n")
        f.write("COLUMNS\n")
        for j in range(lp.ncols):
            if lp.linear[j] != 0:
                f.write(f"    C{j} OBJ {num(lp.linear[j])}\n")
        for k in order:
            f.write(f"    C{lp.cols[k]} R{lp.rows[k]} {num(lp.vals[k])}\n")
        f.write("RHS\n")
        for i in range(lp.nrows):
            if lp.rhs[i] != 0:
                f.write(f"    RHS R{i} {num(lp.rhs[i])}\n")
        f.write("BOUNDS\n")
        for j in range(lp.ncols):
            lo, up = lp.lower[j], lp.upper[j]
            if lo == -INF:
                f.write(f" MI BND C{j}\n")
            elif lo != 0:
                f.write(f" LO BND C{j} {num(lo)}\n")
            if up != INF:
                f.write(f" UP BND C{j} {num(up)}\n")
        f.write("ENDATA\n")


def test_parse_small_fixture(host):
    p = host.mps_parse(SMALL)
    assert (p["nrows"], p["ncols"], p["nstructural"]) == (4, 7, 4)
    assert p["maximize"] and p["objective_offset"] == -10.0
    coo = list(zip(p["rows"].tolist(), p["cols"].tolist(),
                   p["values"].tolist()))
    assert coo == [(0, 0, 1.0), (1, 0, 1.0), (0, 1, 1.0), (2, 1, -1.0),
                   (2, 2, 1.0), (3, 2, 1.0), (3, 3, 2.0),
                   (0, 4, -1.0), (1, 5, -1.0), (3, 6, -1.0)]
    assert p["rhs"].tolist() == [1.5, 1.0, 7.0, 1.5]
    assert p["linear"].tolist() == [-1.0, -2.0, 1.0, -0.5, 0.0, 0.0, 0.0]
    assert p["lower"].tolist() == [0.0, -1.0, -INF, -INF, 0.0, 0.0, 0.0]
    assert p["upper"].tolist() == [4.0, 1.0, 8.0, INF, 2.5, INF, 1.5]


@pytest.mark.parametrize("text, message", [
    ("ROWS\n N C\n E R\nCOLUMNS\n    X C 1 Q 2\nENDATA\n", "unknown row"),
    ("ROWS\n N C\n E R\nCOLUMNS\n    X R 1\n", "missing ENDATA"),
    ("ROWS\n N C\n E R\n E R\nENDATA\n", "duplicate row"),
    ("ROWS\n N C\n E R\nCOLUMNS\n    X R 1x\nENDATA\n", "bad number"),
    ("ROWS\n N C\n E R\nCOLUMNS\n    X R 1\nBOUNDS\n SC B X 4\nENDATA\n",
     "unsupported bound"),
    ("ROWS\n E R\nCOLUMNS\n    X R 1\nENDATA\n", "no objective"),
    ("ROWS\n N C\nFOO\nENDATA\n", "unknown section"),
])
def test_parse_errors(host, tmp_path, text, message):
    path = tmp_path / "bad.mps"
    path.write_text(text)
    with pytest.raises(ValueError, match=message):
        host.mps_parse(str(path))


def test_zero_ranges_on_inequality_rows_make_equalities(host, tmp_path):
    """A G or L row with RANGES 0 is the interval [b, b]: an equality with
    no slack column (the counts of columns and entries follow)."""
    path = tmp_path / "r0.mps"
    path.write_text(
        "NAME R0\nROWS\n N OBJ\n G RG\n L RL\n G RS\nCOLUMNS\n"
        "    X OBJ 1 RG 1\n    X RL 2\n    Y RL 1 RS 3\n"
        "RHS\n    RHS RG 4 RL 5\n    RHS RS 6\n"
        "RANGES\n    RNG RG 0 RL 0\nENDATA\n")
    p = host.mps_parse(str(path))
    # RG: x = 4; RL: 2x + y = 5; RS (no range): 3y - s = 6, s >= 0
    assert (p["nrows"], p["ncols"], p["nstructural"]) == (3, 3, 2)
    coo = sorted(zip(p["rows"].tolist(), p["cols"].tolist(),
                     p["values"].tolist()))
    assert coo == [(0, 0, 1.0), (1, 0, 2.0), (1, 1, 1.0), (2, 1, 3.0),
                   (2, 2, -1.0)]
    assert p["rhs"].tolist() == [4.0, 5.0, 6.0]
    assert p["lower"].tolist()[2] == 0.0 and p["upper"].tolist()[2] == INF


def test_parse_round_trip_of_a_random_lp(host, tmp_path):
    lp = I.random_lp(300, 700, 5, 44)
    path = str(tmp_path / "r.mps")
    write_mps(lp, path)
    p = host.mps_parse(path)
    assert (p["nrows"], p["ncols"], p["nstructural"]) == (lp.nrows,
                                                          lp.ncols, lp.ncols)
    got = sorted(zip(p["rows"].tolist(), p["cols"].tolist(),
                     p["values"].tolist()))
    want = sorted(zip(lp.rows.tolist(), lp.cols.tolist(), lp.vals.tolist()))
    assert got == want
    for k in ("rhs", "linear", "lower", "upper"):
        assert np.array_equal(p[k], getattr(lp, k)), k


@pytest.mark.gpu
def test_read_mps_and_solve_the_small_fixture(gpu):
    alm, info = gpu.alm_read_mps(SMALL)
    n, m = info.ncols, info.nrows
    print("[mps] small", flush=True)
    # the reference's stopping rule (the reference build itself stops at
    # x = (4, -8.8e-7, 6.9999985, -1.9999979), objective 3.9999992)
    rc, x, lam, diag, _ = gpu.alm_solve_ex(alm, 1000, np.zeros(n),
                                           np.zeros(m))
    mat = gpu.lib.alm_matrix(alm)
    gpu.alm_free(alm)
    gpu.free_matrix(mat)
    assert rc == 0
    assert np.allclose(x[:4], [4.0, 0.0, 7.0, -2.0], atol=1e-4)
    # minimised form: -c'x = 4; the maximised objective is c'x - (-10) = 6
    assert abs(diag[0] - 4.0) < 1e-5
    assert abs(-(diag[0] + info.objective_offset) - 6.0) < 1e-5


@pytest.mark.gpu
def test_read_mps_solves_like_alm_make(gpu, tmp_path):
    lp = I.random_lp(300, 700, 5, 45)
    path = str(tmp_path / "r.mps")
    write_mps(lp, path)
    alm, info = gpu.alm_read_mps(path)
    print("[mps] random via MPS", flush=True)
    r1 = gpu.alm_solve_ex(alm, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    mat = gpu.lib.alm_matrix(alm)
    gpu.alm_free(alm)
    gpu.free_matrix(mat)
    m2 = gpu.make_matrix(lp, 0)
    a2 = gpu.alm_make(m2, lp)
    r2 = gpu.alm_solve_ex(a2, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows))
    gpu.alm_free(a2)
    gpu.free_matrix(m2)
    assert r1[0] == r2[0] == 0
    assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
    assert np.array_equal(r1[3], r2[3])
