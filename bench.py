#!/usr/bin/env python3
"""bench.py -- LP solve throughput of the device-resident ALM + APPROX path
on one MI355X.

A step is one complete ``alm_solve`` (ALM outer loop, APPROX inner loop,
SpMV) of the set-cover LP relaxation (m = 200,000 rows, n = 1,000,000
columns plus m slack columns, 0/1 coefficients, 3-8 nonzeros per column,
c ~ U[1,100], 0 <= x <= 1; seeded synthetic instance), from x = 0 and
lambda = 0 until the relative KKT residual of (x, lambda), evaluated on
the device after each outer iteration, is at most 1e-6 (--stop-kkt; 0
selects the reference's absolute rule, alm.c:303-319, instead).  The matrix,
the problem vectors and all solver state are resident in HBM before the
timed region; each step restores the multiplier weights and penalised cost
(acd_alm_reset) and solves again.

Printed (one JSON line): solves/s over the K timed steps, p50/p99 latency
per solve, the relative KKT residual of the result (checked <= 1e-6), the
roofline of the solver's step kernel (algorithmic bytes / kernel time,
both accumulated by the solver's own control kernel over the timed solves),
the stand-alone SpMV kernel's HBM throughput on the same matrix (HIP
events), and the CPU baseline: the reference's thread_pool CPU path
(oracle/_ref/ref_bench, built from the reference sources) or, where that
build is absent, the oracle's C restatement, timed on a bounded sample of
APPROX iterations of the same instance and scaled by the number of inner
iterations the solve takes.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--precision mixed|fp64] [--cpu-iters N] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "acd-alm_amd"), os.path.join(HERE, "oracle")]

METRIC = ("solves/s + p50/p99 latency at 1e-6 rel KKT residual; "
          "SpMV HBM GB/s vs peak")
HBM_PEAK_GBS = 8000.0
KKT_TOL = 1e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class _StdoutToStderr:
    """Sends fd 1 to stderr while the work runs (the solver prints a
    newline on stdout after each restart, as the reference does), so the
    JSON line is the only thing on stdout."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        C.CDLL(None).fflush(None)
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def emit(line):
    print(json.dumps(line), flush=True)


def spmv_leg(lib, mat, lp, reps=50):
    """Stand-alone SpMV (y = A x and y = A' x) on the solver's matrix,
    timed with HIP events on the library stream, after the autotuner has
    picked the variant (unroll, rows per wavefront, load kind) per
    operator."""
    L = lib.lib
    tuned = (C.c_int * 2)()
    if L.acd_spmv_autotune(mat, 10, tuned):
        raise RuntimeError("SpMV autotune failed: %s" % (lib.last_error(),))
    out = {"variant_A": int(tuned[0]), "variant_AT": int(tuned[1])}
    for tr in (0, 1):
        ny, nx = (lp.ncols, lp.nrows) if tr else (lp.nrows, lp.ncols)
        dx = L.acd_device_alloc(8 * nx)
        dy = L.acd_device_alloc(8 * ny)
        x = np.random.default_rng(7).random(nx)
        L.acd_copy_to_device(dx, x.ctypes.data, 8 * nx)
        ms = C.c_double()
        rc = L.acd_spmv_device_time(mat, tr, dy, dx, reps, C.byref(ms))
        L.acd_device_free(dx)
        L.acd_device_free(dy)
        if rc:
            raise RuntimeError("SpMV timing failed: %s" % (lib.last_error(),))
        slices = (ny + 63) // 64
        alg = lp.nnz * (8 + 4) + 8 * (slices + 1) + 8 * nx + 8 * ny
        gbs = alg / (ms.value * 1e-3) / 1e9
        out["AT" if tr else "A"] = {
            "us": round(ms.value * 1e3, 2), "alg_bytes": int(alg),
            "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_spmv(lp, threads, repeat=10):
    """The reference's sparse_matrix_multiply (A and A', FP64, its thread
    pool with `threads` workers) on the same matrix, from the build in
    oracle/_ref; mean seconds per product over `repeat` products."""
    ref = os.path.join(HERE, "oracle", "_ref", "ref_bench")
    if not os.path.exists(ref):
        return {"error": "reference build absent"}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "lp.lpb")
        lp.write_binary(path)
        r = subprocess.run([ref, path, str(threads), "spmv", str(repeat)],
                           capture_output=True, text=True, timeout=900)
    if r.returncode:
        return {"error": r.stderr[-300:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    out = {"threads": threads, "repeat": repeat, "kind": "reference"}
    for key, tr in (("A", 0), ("AT", 1)):
        ny, nx = (lp.ncols, lp.nrows) if tr else (lp.nrows, lp.ncols)
        alg = lp.nnz * (8 + 4) + 8 * ((ny + 63) // 64 + 1) + 8 * nx + 8 * ny
        sec = res["s_per_at" if tr else "s_per_a"]
        out[key] = {"us": round(sec * 1e6, 1),
                    "GB/s": round(alg / sec / 1e9, 2)}
    return out


def cpu_baseline(lp, inner_iterations, iters, threads):
    """Bounded CPU sample: `iters` APPROX iterations of the first ALM
    subproblem on the reference build (or the oracle restatement), scaled to
    the inner iterations one solve takes."""
    ref = os.path.join(HERE, "oracle", "_ref", "ref_bench")
    with tempfile.TemporaryDirectory() as td:
        if os.path.exists(ref):
            path = os.path.join(td, "lp.lpb")
            lp.write_binary(path)
            r = subprocess.run([ref, path, str(threads), "approx", str(iters)],
                               capture_output=True, text=True, timeout=900)
            if r.returncode:
                raise RuntimeError("ref_bench failed: " + r.stderr[-400:])
            res = json.loads(r.stdout.strip().splitlines()[-1])
            spi = res["s_per_iteration"]
            kind, cores = "reference", threads
            # the same on one thread (a tenth of the iterations)
            r1 = subprocess.run([ref, path, "1", "approx",
                                 str(max(10, iters // 10))],
                                capture_output=True, text=True, timeout=900)
            spi1 = (json.loads(r1.stdout.strip().splitlines()[-1])
                    ["s_per_iteration"] if r1.returncode == 0 else None)
        else:
            from pyoracle import Oracle
            ora = Oracle()
            m = ora.make_matrix(lp, 0)
            a = ora.approx_make(m, lp)
            ora.approx_update(a)
            t = time.perf_counter()
            ora.approx_solve(a, np.zeros(lp.ncols), iters)
            spi = (time.perf_counter() - t) / iters
            ora.approx_free(a)
            ora.free_matrix(m)
            kind, cores = "port", 1
            spi1 = spi
    per_solve = spi * inner_iterations
    return {"value": 1.0 / per_solve, "unit": "solves/s", "cores": cores,
            "kind": kind,
            "value_1_thread": (1.0 / (spi1 * inner_iterations)
                               if spi1 else None),
            "sample": ("%d APPROX iterations of the first ALM subproblem of "
                       "the same instance (%.4g s/iteration on %d thread%s), "
                       "scaled by the %d inner iterations of one solve"
                       % (iters, spi, cores, "" if cores == 1 else "s",
                          inner_iterations))}


def batch_cpu_baseline(lps, sample=32, procs=16):
    """The reference's CPU path on the batched workload, LP-parallel (one
    LP per process, pool = NULL, as its thread pool cannot be shared by
    concurrent solves): `sample` LPs spread over the size-sorted batch,
    `procs` at a time; solves/s = sample / wall time.  It stops on the
    reference's absolute rule (its only rule)."""
    ref = os.path.join(HERE, "oracle", "_ref", "ref_bench")
    if not os.path.exists(ref):
        return {"value": None, "error": "reference build absent"}
    order = sorted(range(len(lps)), key=lambda i: lps[i].nnz)
    pick = [order[(k * len(order)) // sample] for k in range(sample)]
    with tempfile.TemporaryDirectory() as td:
        paths = []
        for k, i in enumerate(pick):
            path = os.path.join(td, "lp%d.lpb" % k)
            lps[i].write_binary(path)
            paths.append(path)
        t = time.perf_counter()
        running, fails = [], 0
        for path in paths:
            while len(running) >= procs:
                fails += running.pop(0).wait() != 0
            running.append(subprocess.Popen(
                [ref, path, "1", "alm"], stdout=subprocess.DEVNULL,
                stderr=subprocess.DEVNULL))
        for p in running:
            fails += p.wait() != 0
        wall = time.perf_counter() - t
    return {"value": sample / wall, "unit": "solves/s", "cores": procs,
            "kind": "reference", "failed": fails,
            "sample": ("%d LPs spread over the size-sorted batch, one per "
                       "process, %d processes at a time, reference stopping "
                       "rule (%.1f s wall)" % (sample, procs, wall))}


def serve_cpu_baseline(lps, large, arrivals, large_inner, procs=16,
                       sample=32, large_iters=20):
    """The reference's CPU path replayed on the same arrival trace: a
    server with `procs` host threads where a small LP takes one thread
    (pool = NULL) and a large LP all of them (its thread pool).  Service
    times: each small LP's from a log-log interpolation over `sample`
    small LPs solved by the reference (ALM, its stopping rule, one
    process each, `procs` at a time); a large LP's from `large_iters`
    APPROX iterations on `procs` threads scaled by `large_inner`, the
    inner iterations the device solve of one large LP took.  FIFO
    admission; p50/p99 of completion minus arrival."""
    ref = os.path.join(HERE, "oracle", "_ref", "ref_bench")
    if not os.path.exists(ref):
        return {"value": None, "error": "reference build absent"}
    small = [i for i in range(len(lps)) if not large[i]]
    order = sorted(small, key=lambda i: lps[i].nnz)
    pick = [order[(k * len(order)) // sample] for k in range(min(sample, len(order)))]
    times = {}
    with tempfile.TemporaryDirectory() as td:
        jobs = []
        for k, i in enumerate(pick):
            path = os.path.join(td, "lp%d.lpb" % k)
            lps[i].write_binary(path)
            jobs.append((i, path))
        running = []
        for i, path in jobs:
            while len(running) >= procs:
                j, t0, pr = running.pop(0)
                pr.wait()
                times[j] = time.perf_counter() - t0
            running.append((i, time.perf_counter(), subprocess.Popen(
                [ref, path, "1", "alm"], stdout=subprocess.DEVNULL,
                stderr=subprocess.DEVNULL)))
        for j, t0, pr in running:
            pr.wait()
            times[j] = time.perf_counter() - t0
        spi = None
        big = [i for i in range(len(lps)) if large[i]]
        if big:
            path = os.path.join(td, "large.lpb")
            lps[big[0]].write_binary(path)
            r = subprocess.run([ref, path, str(procs), "approx",
                                str(large_iters)], capture_output=True,
                               text=True, timeout=900)
            if r.returncode == 0:
                spi = json.loads(r.stdout.strip().splitlines()[-1])[
                    "s_per_iteration"]
    xs = np.log([lps[i].nnz for i in sorted(times, key=lambda i: lps[i].nnz)])
    ys = np.log([times[i] for i in sorted(times, key=lambda i: lps[i].nnz)])
    service = np.array([
        (spi * large_inner if spi else np.nan) if large[i]
        else float(np.exp(np.interp(np.log(lps[i].nnz), xs, ys)))
        for i in range(len(lps))])
    # replay: `procs` workers, FIFO; a large LP waits for all of them
    free_at = np.zeros(procs)
    done = np.zeros(len(lps))
    for i in np.argsort(arrivals, kind="stable"):
        if large[i]:
            start = max(arrivals[i], free_at.max())
            free_at[:] = start + service[i]
        else:
            w = int(np.argmin(free_at))
            start = max(arrivals[i], free_at[w])
            free_at[w] = start + service[i]
        done[i] = start + service[i]
    lat = (done - np.asarray(arrivals)) * 1e3
    return {"value": float(len(lps) / done.max()), "unit": "solves/s",
            "cores": procs, "kind": "reference",
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "large_service_s": (spi * large_inner) if spi else None,
            "sample": ("trace replay on %d host threads; small-LP service "
                       "times from %d reference ALM solves (one thread "
                       "each, median %.3g s), large-LP service time %.4g "
                       "s/iteration x %d inner iterations on %d threads"
                       % (procs, len(times), float(np.median(list(
                           times.values()))), spi or float("nan"),
                          large_inner, procs))}


def batch_leg(lib, count, prec, stop_kkt, steps=1, fp64_sample=8,
              cpu=True, mmin=50, mmax=2000):
    """Batched solve of `count` independent LPs (m log-uniform in [mmin,
    mmax], default [50, 2000], n = 2-3 m, 5 nonzeros per column, random
    +-U[0.5, 2] values; seeded),
    all device-resident, to the relative KKT target.  Also times the plain
    FP64 single-LP form (alm_solve_ex, one LP at a time) on the first
    `fp64_sample` LPs for the ratio."""
    from acdalm import instances, kkt
    L = lib.lib
    t0 = time.perf_counter()
    lps = instances.batch(count, seed=4, mmin=mmin, mmax=mmax)
    mats = [lib.make_matrix(lp, 0) for lp in lps]
    alms = [lib.alm_make(m, lp) for m, lp in zip(mats, lps)]
    setup = time.perf_counter() - t0
    x0s = [np.zeros(lp.ncols) for lp in lps]
    l0s = [np.zeros(lp.nrows) for lp in lps]
    dev_ms, done_all, kkts, inner, alg = 0.0, [], [], 0, 0.0
    for _ in range(steps):
        for a in alms:
            L.acd_alm_reset(a)
        # With FP32-rounded matrix values the device measures the KKT
        # residual of the rounded LP; a 10% margin keeps the true LP's
        # residual (checked below on the host) under the target.
        out, done, info = lib.alm_solve_batch(
            alms, 1000, x0s, l0s, precision=prec,
            stop_kkt=stop_kkt * (0.9 if prec else 1.0))
        dev_ms += info.device_ms
        done_all.append(done)
        inner += int(info.inner_iterations)
        alg += info.alg_bytes
        wpb = list(info.waves_per_lp)
        bad = [i for i, r in enumerate(out) if r[0] != 0]
        if bad:
            raise RuntimeError("batch: %d LPs did not converge" % len(bad))
    kkts = [kkt.kkt(lp, r[1], r[2])["kkt"] for lp, r in zip(lps, out)]
    done = np.concatenate(done_all)
    # the plain FP64 form on a sample, same target
    t = time.perf_counter()
    for a, lp in zip(alms[:fp64_sample], lps[:fp64_sample]):
        L.acd_alm_reset(a)
        lib.alm_solve_ex(a, 1000, np.zeros(lp.ncols), np.zeros(lp.nrows),
                         precision=0, stop_kkt=stop_kkt)
    single = fp64_sample / (time.perf_counter() - t)
    for a in alms:
        lib.alm_free(a)
    for m in mats:
        lib.free_matrix(m)
    cpu_b = None
    if cpu:
        try:
            cpu_b = batch_cpu_baseline(lps, procs=min(16, os.cpu_count() or 1))
        except Exception as e:  # reported, never silently replaced
            cpu_b = {"value": None, "error": str(e)[-300:]}
    rate = count * steps / (dev_ms * 1e-3)
    gbs = alg / (dev_ms * 1e-3) / 1e9
    return {
        "workload": "%d independent LPs, m log-uniform [%d,%d], n=2-3m, "
                    "5 nnz/column" % (count, mmin, mmax),
        "precision": "mixed" if prec == 1 else "fp64",
        "solves_per_s": rate, "device_ms_per_batch": dev_ms / steps,
        "p50_ms": float(np.percentile(done, 50)),
        "p99_ms": float(np.percentile(done, 99)),
        "kkt_max": float(max(kkts)), "inner_iterations": inner // steps,
        "waves_per_lp_by_class": wpb,
        "roofline": {"kernel": "k_batch", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4)},
        "fp64_single_lp_solves_per_s": single,
        "fp64_single_lp_sample": fp64_sample,
        "vs_fp64_single_lp": rate / single,
        "cpu_baseline": cpu_b,
        "setup_s": setup,
    }


def serve_leg(lib, rate, seconds, stop_kkt, prec=1, seed=9,
              large_frac=0.01, budget_gb=64.0, cpu=True):
    """Request loop (continuous batching): Poisson arrivals at `rate` per
    second for `seconds`; 99% LPs built like the batched workload, 1% large
    (m=100,000, n=300,000, 8 nonzeros per column).  Each request's ALM
    object is made before the clock starts (its data device-resident); the
    latency of a request runs from its arrival time to the step that
    retires it."""
    from acdalm import instances, kkt
    rng = np.random.default_rng(seed)
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        arrivals.append(t)
    large = rng.random(len(arrivals)) < large_frac
    ms = np.rint(np.exp(rng.uniform(np.log(50), np.log(2000), len(arrivals))))
    ns = np.rint(rng.uniform(2.0, 3.0, len(arrivals)) * ms)
    seeds = rng.integers(0, 2**62, size=len(arrivals))
    t0 = time.perf_counter()
    lps = [instances.random_lp(100000, 300000, 8, int(sd)) if lg else
           instances.random_lp(int(m), int(n), 5, int(sd))
           for lg, m, n, sd in zip(large, ms, ns, seeds)]
    mats = [lib.make_matrix(lp, 0) for lp in lps]
    alms = [lib.alm_make(m, lp) for m, lp in zip(mats, lps)]
    setup = time.perf_counter() - t0
    srv = lib.server(int(budget_gb * 2**30), slots=2048, large_rows=50000,
                     precision=prec, stop_kkt=stop_kkt * (0.9 if prec else 1.0))
    lat = np.zeros(len(lps))
    results = {}
    nxt = 0
    start = time.perf_counter()
    while nxt < len(lps) or srv.pending():
        now = time.perf_counter() - start
        while nxt < len(lps) and arrivals[nxt] <= now:
            srv.submit(nxt, alms[nxt], np.zeros(lps[nxt].ncols),
                       np.zeros(lps[nxt].nrows))
            nxt += 1
        done = srv.step()
        now = time.perf_counter() - start
        for tag, ret, x, lam, diag in done:
            lat[tag] = (now - arrivals[tag]) * 1e3
            results[tag] = (ret, x, lam)
        if not done and nxt < len(lps) and not srv.pending():
            time.sleep(max(0.0, min(1e-3, arrivals[nxt] - now)))
    end = time.perf_counter() - start
    srv.close()
    kk = [kkt.kkt(lp, results[i][1], results[i][2])["kkt"]
          for i, lp in enumerate(lps)]
    cpu_b = None
    if cpu:
        try:
            # inner iterations of one large LP solved alone on the device
            large_inner = 0
            if large.any():
                i = int(np.flatnonzero(large)[0])
                lib.lib.acd_alm_reset(alms[i])
                lib.alm_solve_ex(alms[i], 1000, np.zeros(lps[i].ncols),
                                 np.zeros(lps[i].nrows), precision=prec,
                                 stop_kkt=stop_kkt * (0.9 if prec else 1.0))
                large_inner = int(lib.stats().inner_iterations)
            cpu_b = serve_cpu_baseline(lps, large, arrivals, large_inner,
                                       procs=min(16, os.cpu_count() or 1))
        except Exception as e:  # reported, never silently replaced
            cpu_b = {"value": None, "error": str(e)[-300:]}
    for a in alms:
        lib.alm_free(a)
    for m in mats:
        lib.free_matrix(m)
    small = ~large
    return {
        "workload": "Poisson arrivals at %g/s for %g s: %d requests (%d "
                    "large m=100000 n=300000)" % (rate, seconds, len(lps),
                                                   int(large.sum())),
        "solves_per_s": len(lps) / end, "offered_per_s": rate,
        "makespan_s": end,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "small_p50_ms": float(np.percentile(lat[small], 50)) if small.any() else None,
        "small_p99_ms": float(np.percentile(lat[small], 99)) if small.any() else None,
        "large_p50_ms": float(np.percentile(lat[large], 50)) if large.any() else None,
        "kkt_max": float(max(kk)),
        "all_converged": all(r[0] == 0 for r in results.values()),
        "cpu_baseline": cpu_b,
        "setup_s": setup,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cpu-iters", type=int, default=300)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stop-kkt", type=float, default=1e-6,
                    help="stop once the relative KKT residual is at most "
                         "this (0: the reference's absolute rule)")
    ap.add_argument("--log", action="store_true",
                    help="stream the ALM log (one line per outer iteration) "
                         "to stderr")
    ap.add_argument("--outer", type=int, default=1000,
                    help="ALM outer-iteration cap (the reference uses 1000)")
    ap.add_argument("--workload", choices=["setcover", "mcf", "batch", "serve"],
                    default="setcover")
    ap.add_argument("--rate", type=float, default=2000.0,
                    help="serve: offered requests per second")
    ap.add_argument("--seconds", type=float, default=1.0,
                    help="serve: length of the arrival window")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-mmin", type=int, default=50,
                    help="batched LPs: smallest row count (log-uniform)")
    ap.add_argument("--batch-mmax", type=int, default=2000,
                    help="batched LPs: largest row count")
    ap.add_argument("--no-batch", action="store_true",
                    help="setcover workload: skip the batched-solve leg")
    ap.add_argument("--copy-check", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES "
                         "(calibrates hardware byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (each LP solve runs on one device).", file=sys.stderr)
        sys.exit(2)

    from acdalm import capi, instances, kkt
    lib = capi.load()
    L = lib.lib
    prec = {"mixed": 1, "fp64": 0}[args.precision]

    if args.copy_check:
        src = L.acd_device_alloc(args.copy_check)
        dst = L.acd_device_alloc(args.copy_check)
        L.acd_copy_device(dst, src, args.copy_check)
        L.acd_device_free(src)
        L.acd_device_free(dst)

    if args.workload == "serve":
        r = serve_leg(lib, args.rate, args.seconds, args.stop_kkt or 1e-6,
                      prec=prec, cpu=not args.no_cpu)
        return ({
            "metric": METRIC, "value": r["solves_per_s"], "unit": "solves/s",
            "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": r["makespan_s"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded generator)",
            "config": {"workload": r["workload"], "precision": args.precision,
                       "stop": "relative KKT <= %g" % (args.stop_kkt or 1e-6),
                       "parallelism": "single GPU"},
            "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
            "roofline": None, "cpu_baseline": r["cpu_baseline"], "serve": r})

    if args.workload == "batch":
        L.acd_synchronize()
        t = time.perf_counter()
        b = batch_leg(lib, args.batch_count, prec, args.stop_kkt,
                      steps=args.steps, cpu=not args.no_cpu,
                      mmin=args.batch_mmin, mmax=args.batch_mmax)
        total = time.perf_counter() - t
        return ({
            "metric": METRIC, "value": b["solves_per_s"], "unit": "solves/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": b["device_ms_per_batch"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded generator)",
            "config": {"workload": b["workload"], "precision": args.precision,
                       "stop": "relative KKT <= %g" % args.stop_kkt,
                       "parallelism": "single GPU"},
            "p50_ms": b["p50_ms"], "p99_ms": b["p99_ms"],
            "roofline": dict(b["roofline"], bound="hbm", traffic=None),
            "cpu_baseline": b["cpu_baseline"], "batch": b, "wall_s": total})

    t0 = time.perf_counter()
    if args.workload == "mcf":
        lp = instances.multicommodity()
        workload = ("multicommodity flow LP: 100x100 grid, 39,600 arcs, 64 "
                    "commodities, m=%d n=%d, +-1 A" % (lp.nrows, lp.ncols))
    else:
        lp = instances.set_cover()
        workload = ("set-cover LP relaxation m=200000 n=1000000+200000 "
                    "slack, 0/1 A, 3-8 nnz/column")
    mat = lib.make_matrix(lp, 0)
    alm = lib.alm_make(mat, lp)
    x0 = np.zeros(lp.ncols)
    lam0 = np.zeros(lp.nrows)
    # Untimed priming: one outer iteration builds the solver graphs and the
    # reduced-precision copies of the matrix.
    lib.alm_solve_ex(alm, 1, x0, lam0, precision=prec)
    setup_s = time.perf_counter() - t0
    log("setup %.1f s (m=%d n=%d nnz=%d)" % (setup_s, lp.nrows, lp.ncols, lp.nnz))

    def one_solve():
        L.acd_alm_reset(alm)
        t = time.perf_counter()
        rc, x, lam, diag, _ = lib.alm_solve_ex(
            alm, args.outer, x0, lam0, log="stderr" if args.log else False,
            precision=prec, stop_kkt=args.stop_kkt)
        return time.perf_counter() - t, rc, x, lam, diag

    for w in range(args.warmup):
        dt, rc, *_ = one_solve()
        log("warmup %d: %.2f s rc=%d" % (w, dt, rc))

    L.acd_synchronize()
    lat = []
    step_us = np.zeros(32)
    bytes_ = np.zeros(32)
    launches = 0
    inner = outer = 0
    device_ms = 0.0
    objs = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        dt, rc, x, lam, diag = one_solve()
        lat.append(dt)
        st = lib.stats()
        inner, outer = st.inner_iterations, st.outer_iterations
        device_ms += st.device_ms
        for name, (cnt, sus, by) in lib.profile_bytes().items():
            i = [L.acd_pass_name(j).decode() for j in range(32)].index(name)
            step_us[i] += sus
            bytes_[i] += by
            launches += cnt
        objs.append(diag[0])
        log("step %d: %.2f s rc=%d inner=%d outer=%d obj=%.12g"
            % (k, dt, rc, inner, outer, diag[0]))
    L.acd_synchronize()
    total = time.perf_counter() - t_start

    res = kkt.kkt(lp, x, lam)
    kkt_ok = bool(rc == 0 and res["kkt"] <= KKT_TOL)
    lat_ms = np.array(lat) * 1e3
    spmv = spmv_leg(lib, mat, lp)
    cpu_threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    if not args.no_cpu:
        spmv["cpu_reference"] = cpu_spmv(lp, cpu_threads)

    achieved = bytes_.sum() / (step_us.sum() * 1e-6) / 1e9
    names = [L.acd_pass_name(j).decode() for j in range(32)]
    top = np.argsort(-step_us)[:3]
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "kernel": "k_step (solver step kernel, all passes)",
        "launches": int(launches),
        "avg_us": round(step_us.sum() / max(launches, 1), 2),
        "alg_bytes_per_launch": round(bytes_.sum() / max(launches, 1)),
        "by_pass": {names[i]: {
            "GB/s": round(bytes_[i] / (step_us[i] * 1e-6) / 1e9, 1),
            "share": round(step_us[i] / step_us.sum(), 3)}
            for i in top if step_us[i] > 0},
    }
    # The stand-alone SpMV on the other headline matrix: the set-cover A is
    # bound by random gathers, the multicommodity operators stream.
    spmv_other = None
    if args.workload == "setcover" and not args.no_batch:
        try:
            lp2 = instances.multicommodity()
            mat2 = lib.make_matrix(lp2, 0)
            spmv_other = dict(spmv_leg(lib, mat2, lp2),
                              workload="multicommodity flow LP, m=%d n=%d "
                              "nnz=%d" % (lp2.nrows, lp2.ncols, lp2.nnz))
            lib.free_matrix(mat2)
            if not args.no_cpu:
                spmv_other["cpu_reference"] = cpu_spmv(lp2, cpu_threads)
        except Exception as e:  # reported, never silently replaced
            spmv_other = {"error": str(e)[-300:]}
    batch = None
    if not args.no_batch:
        try:
            batch = batch_leg(lib, args.batch_count, 1, args.stop_kkt or 1e-6,
                              cpu=not args.no_cpu)
        except Exception as e:  # reported, never silently replaced
            batch = {"error": str(e)[-300:]}
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(lp, inner, args.cpu_iters, threads)
        except Exception as e:  # reported, never silently replaced
            cpu = {"value": None, "error": str(e)[-300:]}

    line = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": workload,
                   "precision": args.precision,
                   "stop": "relative KKT <= %g" % args.stop_kkt
                           if args.stop_kkt > 0 else "reference rule",
                   "nnz": int(lp.nnz), "m": int(lp.nrows), "n": int(lp.ncols),
                   "parallelism": "single GPU"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "kkt": res["kkt"], "kkt_ok": kkt_ok,
        "objective": objs[-1], "deterministic": len(set(objs)) == 1,
        "inner_iterations": int(inner), "outer_iterations": int(outer),
        "solver_device_ms": device_ms / args.steps,
        "roofline": roofline,
        "spmv": spmv,
        "spmv_mcf": spmv_other,
        "batch": batch,
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    return line


if __name__ == "__main__":
    with _StdoutToStderr():
        result = main()
    emit(result)
