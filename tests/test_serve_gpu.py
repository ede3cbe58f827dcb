"""The request loop (acd_server_*): continuous batching with admission by
device footprint.  Requests arrive while others are running; small LPs
share the batched kernel, large ones run the multi-workgroup solver on
their own streams at the same time.  Each LP's arithmetic depends only on
that LP, so every result must be bit-identical alone and in the mix."""
import time

import numpy as np
import pytest

from acdalm import instances as I
from acdalm.kkt import kkt

pytestmark = pytest.mark.gpu
LARGE_ROWS = 150   # LPs with at least this many rows take the large lane


def problems():
    small = I.batch(10, seed=11, mmin=20, mmax=120)
    large = [I.random_lp(160, 400, 5, 12), I.random_lp(200, 420, 5, 13)]
    return small + large


def serve(gpu, lps, budget=1 << 40, waves=((0, None),), **opts):
    """Submit lps in waves [(round, indices)], step until drained; return
    {index: (ret, x, lam, diag)}."""
    mats = [gpu.make_matrix(lp, 0) for lp in lps]
    alms = [gpu.alm_make(m, lp) for m, lp in zip(mats, lps)]
    srv = gpu.server(budget, slots=64, large_rows=LARGE_ROWS, **opts)
    out, rnd = {}, 0
    todo = list(waves)
    t0 = time.time()
    while todo or srv.pending():
        while todo and todo[0][0] <= rnd:
            _, idx = todo.pop(0)
            for i in (range(len(lps)) if idx is None else idx):
                srv.submit(i, alms[i], np.zeros(lps[i].ncols),
                           np.zeros(lps[i].nrows))
        for tag, ret, x, lam, diag in srv.step():
            out[tag] = (ret, x, lam, diag)
        rnd += 1
        assert time.time() - t0 < 120, "request loop did not drain"
    srv.close()
    for a in alms:
        gpu.alm_free(a)
    for m in mats:
        gpu.free_matrix(m)
    return out


def same(a, b):
    return (a[0] == b[0] and np.array_equal(a[1], b[1])
            and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]))


@pytest.mark.parametrize("precision", [0, 1])
def test_each_result_is_the_same_alone_and_in_the_mix(gpu, precision):
    lps = problems()
    print(f"[serve] mix precision={precision}", flush=True)
    n = len(lps)
    # arrivals spread over rounds, large LPs in the middle of the small ones
    waves = [(0, [0, 1, 2, 10]), (3, [3, 4, 5]), (7, [11, 6, 7]),
             (20, [8, 9])]
    mixed = serve(gpu, lps, waves=waves, precision=precision, stop_kkt=1e-6)
    assert sorted(mixed) == list(range(n))
    for i in range(n):
        alone = serve(gpu, [lps[i]], precision=precision, stop_kkt=1e-6)[0]
        assert same(mixed[i], alone), i
        ret, x, lam, _ = mixed[i]
        assert ret == 0
        assert kkt(lps[i], x, lam)["kkt"] <= 1e-6 * (1.12 if precision else 1.0)


def test_small_budget_admits_one_at_a_time(gpu):
    lps = problems()[:6]
    print("[serve] budget", flush=True)
    wide = serve(gpu, lps, precision=0)
    narrow = serve(gpu, lps, budget=1, precision=0)
    for i in range(len(lps)):
        assert same(wide[i], narrow[i]), i


def test_a_failing_request_completes_with_an_error(gpu):
    lps = problems()[:3]
    bad = I.random_lp(40, 100, 4, 10)
    bad.linear = bad.linear.copy()
    bad.linear[7] = np.nan
    print("[serve] failing request", flush=True)
    out = serve(gpu, lps + [bad], precision=0)
    assert out[3][0] == -1
    assert all(out[i][0] == 0 for i in range(3))
