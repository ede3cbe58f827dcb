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


def serve_host_data(gpu, lps, budget, waves=((0, None),), **opts):
    """Like serve(), with the LPs submitted as host data: each is built on
    the device at admission and freed at completion.  Returns the results
    and the peak device bytes in use above the level after the server was
    made (sampled after every step)."""
    # the device allocator's first block and the library's one-time
    # buffers are not the requests': made and freed before the baseline
    m = gpu.make_matrix(lps[0], 0)
    a = gpu.alm_make(m, lps[0])
    gpu.alm_free(a)
    gpu.free_matrix(m)
    srv = gpu.server(budget, slots=64, large_rows=LARGE_ROWS, **opts)
    free0, _ = gpu.memory_info()
    out, rnd, peak = {}, 0, 0
    todo = list(waves)
    t0 = time.time()
    while todo or srv.pending():
        while todo and todo[0][0] <= rnd:
            _, idx = todo.pop(0)
            for i in (range(len(lps)) if idx is None else idx):
                srv.submit_lp(i, lps[i], np.zeros(lps[i].ncols),
                              np.zeros(lps[i].nrows))
        for tag, ret, x, lam, diag in srv.step():
            out[tag] = (ret, x, lam, diag)
        free, _ = gpu.memory_info()
        peak = max(peak, free0 - free)
        rnd += 1
        assert time.time() - t0 < 120, "request loop did not drain"
    srv.close()
    return out, peak


def test_host_data_requests_are_admitted_by_footprint(gpu):
    """Requests given as host arrays hold device memory only while admitted:
    with a budget of a quarter of their summed footprint estimates, the
    device memory in use stays within the budget, and every result equals
    the result of the same LP served alone."""
    lps = I.batch(24, seed=16, mmin=150, mmax=900)
    est = [int(gpu.lib.acd_server_footprint(lp.nrows, lp.ncols, lp.nnz))
           + (8 << 20 if lp.nrows >= LARGE_ROWS else 0) for lp in lps]
    budget = max(sum(est) // 4, max(est))
    print(f"[serve] host data: {len(lps)} LPs, estimates {sum(est)} B, "
          f"budget {budget} B", flush=True)
    waves = [(0, list(range(12))), (3, list(range(12, 24)))]
    mixed, peak = serve_host_data(gpu, lps, budget, waves=waves, precision=1,
                                  stop_kkt=0.9e-6)
    print(f"[serve] peak device bytes in use {peak}", flush=True)
    assert sorted(mixed) == list(range(len(lps)))
    assert peak <= budget, (peak, budget)
    for i in (0, 7, 13, 23):
        alone, _ = serve_host_data(gpu, [lps[i]], 1 << 40, precision=1,
                                   stop_kkt=0.9e-6)
        assert same(mixed[i], alone[0]), i
    for i, lp in enumerate(lps):
        assert mixed[i][0] == 0
        assert kkt(lp, mixed[i][1], mixed[i][2])["kkt"] <= 1e-6
