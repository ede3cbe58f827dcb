#!/usr/bin/env python3
"""bench.py -- LP solve throughput of the device-resident ALM + APPROX path
on one MI355X.

Default workload (one step = one batched solve): 4,096 independent LPs built
like the plain random LP (m log-uniform in [50, 2000], n = 2-3 m, 5
nonzeros per column at distinct uniform rows, values +-U[0.5, 2], b = A x0
with x0 ~ U[0, 10], c ~ U[-1, 1], 0 <= x <= 10; seeded generator
acdalm.instances.batch(4096, seed=4)), all solved together by
alm_solve_batch in mixed precision (FP32 matrix values and per-column
constants, FP64 iterates, accumulation, multipliers and scalars), from
x = 0 and lambda = 0 until each LP's relative KKT residual is at most 1e-6
(checked on the host afterwards).  Matrices, problem vectors and solver
state are resident in HBM before the timed region; each step restores the
row weights and penalised cost (acd_alm_reset) and solves all LPs again.

Printed (one JSON line on stdout):
  value        solves/s = LPs x steps / device time of the steps (HIP events
               around the batched launches; the host-side upload of x0 and
               lambda0 and the download of the results are outside; the
               rate including them is `wall_solves_per_s`);
  p50/p99      per-LP completion time from the start of its batch;
  roofline     the batched kernel k_batch: algorithmic bytes / the summed
               HIP-event time of its launch rounds, against 8 TB/s;
  fp64 ratio   the plain FP64 single-LP form (alm_solve_ex, one LP at a
               time) on a size-stratified sample of the same LPs;
  cpu_baseline the oracle (this repository's C restatement of the
               reference's CPU path, oracle/ora_bench) solving a
               size-stratified sample LP-parallel on the host cores, to
               the same KKT rule, and on one thread;
  spmv         stand-alone SpMV on the set-cover (C1) and multicommodity
               (C2) operators, rotating over several copies so the working
               set exceeds the 256 MB last-level cache, with the oracle's
               threaded SpMV on the host beside it.
Legs after the timed loop run only while the wall-clock budget
(--budget-s) allows; a leg that does not fit is reported as skipped.

--workload setcover|mcf makes one ALM solve of the set-cover or
multicommodity LP the step; --workload serve runs the request loop;
--workload cd times parallel coordinate descent against the sequential
device port of cd_solve (value: solves/s to a fixed objective).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W]
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
ORA_BENCH = os.path.join(HERE, "oracle", "_build", "ora_bench")
T_START = time.perf_counter()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def elapsed():
    return time.perf_counter() - T_START


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


def host_cpu():
    """Host cores this process may use (affinity, capped by the box's
    declared share in OMP_NUM_THREADS) and the CPU model."""
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": max(1, threads), "affinity_cpus": avail,
            "omp_num_threads": share or None, "model": model}


def stratified(lps, k):
    """k LPs spread evenly over the batch sorted by size (nonzeros)."""
    order = sorted(range(len(lps)), key=lambda i: (lps[i].nnz, i))
    k = min(k, len(order))
    return [order[(j * len(order)) // k + len(order) // (2 * k)]
            for j in range(k)]


# ---------------------------------------------------------------- CPU legs
def ora_alm(lps, idx, threads, stop_kkt):
    """The oracle's ALM (APPROX inner solver, niter 1000, from 0) on the
    LPs idx, LP-parallel on `threads` host threads; its JSON result."""
    with tempfile.TemporaryDirectory() as td:
        paths = []
        for k, i in enumerate(idx):
            p = os.path.join(td, "lp%d.lpb" % k)
            lps[i].write_binary(p)
            paths.append(p)
        r = subprocess.run([ORA_BENCH, "alm", str(threads), repr(stop_kkt)]
                           + paths, capture_output=True, text=True,
                           timeout=600)
    if r.returncode:
        raise RuntimeError("ora_bench alm: " + r.stderr[-300:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def batch_cpu_baseline(lps, stop_kkt, cpu, sample=256, sample1=16):
    """Oracle ALM on a size-stratified sample, LP-parallel on all host
    threads (solves/s = sample / wall), and on one thread."""
    idx = stratified(lps, sample)
    res = ora_alm(lps, idx, cpu["threads"], stop_kkt)
    idx1 = stratified(lps, sample1)
    res1 = ora_alm(lps, idx1, 1, stop_kkt)
    bad = sum(1 for r in res["ret"] if r != 0)
    return {"value": len(idx) / res["wall_s"], "unit": "solves/s",
            "cores": cpu["threads"], "kind": "port",
            "value_1_thread": len(idx1) / res1["wall_s"],
            "kkt_max": max(res["kkt"]), "not_converged": bad,
            "cpu_model": cpu["model"], "affinity_cpus": cpu["affinity_cpus"],
            "sample": ("%d LPs spread over the size-sorted batch, solved by "
                       "the oracle's ALM (APPROX inner solver, niter 1000) "
                       "to relative KKT <= %g, one LP per thread at a time "
                       "on %d threads (%.1f s wall); one thread: %d LPs "
                       "(%.1f s)" % (len(idx), stop_kkt, cpu["threads"],
                                     res["wall_s"], len(idx1),
                                     res1["wall_s"]))}


def cpu_spmv(lp, threads, repeat=10):
    """The oracle's CSR products y = A x and y = A' x, rows split over
    `threads` host threads; mean seconds per product."""
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "lp.lpb")
        lp.write_binary(path)
        r = subprocess.run([ORA_BENCH, "spmv", str(threads), str(repeat),
                            path], capture_output=True, text=True,
                           timeout=600)
    if r.returncode:
        return {"error": r.stderr[-300:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    out = {"threads": threads, "repeat": repeat, "kind": "port"}
    for key, tr in (("A", 0), ("AT", 1)):
        ny, nx = (lp.ncols, lp.nrows) if tr else (lp.nrows, lp.ncols)
        alg = spmv_bytes(lp, ny, nx)
        sec = res["s_per_at" if tr else "s_per_a"]
        out[key] = {"us": round(sec * 1e6, 1),
                    "GB/s": round(alg / sec / 1e9, 2)}
    return out


# ---------------------------------------------------------------- GPU legs
def spmv_bytes(lp, ny, nx):
    """Algorithmic bytes of one FP64 product: each nonzero (8-byte value,
    4-byte column) once, the slice offsets, x once, y once."""
    return lp.nnz * (8 + 4) + 8 * ((ny + 63) // 64 + 1) + 8 * nx + 8 * ny


def spmv_leg(lib, lp, copies=4, reps=20):
    """Stand-alone FP64 SpMV, y = A x and y = A' x, on `copies` device
    copies of the matrix used in rotation (their operators together exceed
    the 256 MB last-level cache), timed with HIP events on the library
    stream; and on one copy repeated (cache-assisted when it fits)."""
    L = lib.lib
    mats = [lib.make_matrix(lp, 0) for _ in range(copies)]
    tuned = (C.c_int * 2)()
    if L.acd_spmv_autotune(mats[0], 5, tuned):
        raise RuntimeError("SpMV autotune failed: %s" % (lib.last_error(),))
    for m in mats[1:]:
        L.acd_spmv_autotune(m, 1, None)
    out = {"variant_A": int(tuned[0]), "variant_AT": int(tuned[1]),
           "copies": copies}
    for tr in (0, 1):
        ny, nx = (lp.ncols, lp.nrows) if tr else (lp.nrows, lp.ncols)
        dxs = [L.acd_device_alloc(8 * nx) for _ in mats]
        dys = [L.acd_device_alloc(8 * ny) for _ in mats]
        x = np.random.default_rng(7).random(nx)
        for dx in dxs:
            L.acd_copy_to_device(dx, x.ctypes.data, 8 * nx)
        ms = C.c_double()
        arr_m = (C.c_void_p * copies)(*mats)
        rc = L.acd_spmv_rotate_time(arr_m, copies, tr,
                                    (C.c_void_p * copies)(*dys),
                                    (C.c_void_p * copies)(*dxs), reps,
                                    C.byref(ms))
        ms1 = C.c_double()
        rc = rc or L.acd_spmv_device_time(mats[0], tr, dys[0], dxs[0], reps,
                                          C.byref(ms1))
        for p in dxs + dys:
            L.acd_device_free(p)
        if rc:
            raise RuntimeError("SpMV timing failed: %s" % (lib.last_error(),))
        alg = spmv_bytes(lp, ny, nx)
        gbs = alg / (ms.value * 1e-3) / 1e9
        gbs1 = alg / (ms1.value * 1e-3) / 1e9
        stored = int(L.acd_matrix_stored_entries(mats[0], tr))
        out["AT" if tr else "A"] = {
            "us": round(ms.value * 1e3, 2), "alg_bytes": int(alg),
            "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "working_set_MB": round(copies * (stored * 12 + 8 * (nx + ny))
                                    / 1e6, 1),
            "one_copy_us": round(ms1.value * 1e3, 2),
            "one_copy_GB/s": round(gbs1, 1)}
    for m in mats:
        lib.free_matrix(m)
    return out


def fp64_single_leg(lib, lps, alms, idx, stop_kkt):
    """The plain FP64 GPU form: alm_solve_ex one LP at a time (the
    library's default engine) to the same KKT target, on the LPs idx."""
    from acdalm import kkt
    L = lib.lib
    t = time.perf_counter()
    worst = 0.0
    for i in idx:
        L.acd_alm_reset(alms[i])
        rc, x, lam, diag, _ = lib.alm_solve_ex(
            alms[i], 1000, np.zeros(lps[i].ncols), np.zeros(lps[i].nrows),
            precision=0, stop_kkt=stop_kkt)
        if rc != 0:
            raise RuntimeError("fp64 single-LP solve did not converge")
        worst = max(worst, kkt.kkt(lps[i], x, lam)["kkt"])
    sec = time.perf_counter() - t
    return {"solves_per_s": len(idx) / sec, "lps": len(idx),
            "seconds": sec, "kkt_max": worst,
            "sample": "%d LPs spread over the size-sorted batch" % len(idx)}


def run_batch(args, lib):
    from acdalm import instances, kkt
    L = lib.lib
    prec = {"mixed": 1, "fp64": 0}[args.precision]
    count = args.batch_count
    # With FP32-rounded matrix values the device measures the KKT residual
    # of the rounded LP; a 10% margin keeps the true LP's residual (checked
    # on the host) under the target.
    dev_kkt = args.stop_kkt * (0.9 if prec else 1.0)
    t0 = time.perf_counter()
    lps = instances.batch(count, seed=4, mmin=args.batch_mmin,
                          mmax=args.batch_mmax)
    mats = [lib.make_matrix(lp, 0) for lp in lps]
    alms = [lib.alm_make(m, lp) for m, lp in zip(mats, lps)]
    x0s = [np.zeros(lp.ncols) for lp in lps]
    l0s = [np.zeros(lp.nrows) for lp in lps]
    # niter = 0 runs no iteration; it makes the reduced-precision copies.
    lib.alm_solve_batch(alms, 0, x0s, l0s, precision=prec)
    setup = time.perf_counter() - t0
    log("batch setup %.1f s (%d LPs, %d nonzeros)"
        % (setup, count, sum(lp.nnz for lp in lps)))

    def step():
        for a in alms:
            L.acd_alm_reset(a)
        t = time.perf_counter()
        out, done, info = lib.alm_solve_batch(alms, 1000, x0s, l0s,
                                              precision=prec,
                                              stop_kkt=dev_kkt)
        wall = time.perf_counter() - t
        bad = [i for i, r in enumerate(out) if r[0] != 0]
        if bad:
            raise RuntimeError(
                "batch: %d LPs did not converge: %s" % (len(bad), "; ".join(
                    "LP %d (m=%d n=%d): ret %d, objective %.12g, max "
                    "violation %.3g, %d inner iterations" % (
                        i, lps[i].nrows, lps[i].ncols, out[i][0],
                        out[i][3][0], out[i][3][1], int(info.inner[i]))
                    for i in bad[:4])))
        return out, done, info, wall

    for w in range(args.warmup):
        _, _, info, wall = step()
        log("warmup %d: %.2f s device, %.2f s wall" % (w, info.device_ms / 1e3,
                                                       wall))
    L.acd_synchronize()
    dev_ms = ker_ms = alg = walls = ops_ms = ctl_ms = 0.0
    launches = inner = 0
    inner_max = (0, 0, 0)
    dones = []
    for k in range(args.steps):
        out, done, info, wall = step()
        dev_ms += info.device_ms
        ker_ms += info.kernel_ms
        ops_ms += info.lp_ops_ms
        ctl_ms += info.lp_ctl_ms
        alg += info.alg_bytes
        launches += int(info.launches)
        inner += int(info.inner_iterations)
        big = int(np.argmax(info.inner))
        if int(info.inner[big]) > inner_max[0]:
            inner_max = (int(info.inner[big]), lps[big].nrows, lps[big].ncols)
        walls += wall
        dones.append(done)
        wpl = list(info.waves_per_lp)
        log("step %d: %.2f s device (%.2f s kernel), %.2f s wall, %d launches"
            % (k, info.device_ms / 1e3, info.kernel_ms / 1e3, wall,
               info.launches))
    L.acd_synchronize()
    kk = [kkt.kkt(lp, r[1], r[2])["kkt"] for lp, r in zip(lps, out)]
    done = np.concatenate(dones)
    value = count * args.steps / (dev_ms * 1e-3)
    achieved = alg / (ker_ms * 1e-3) / 1e9
    res = {
        "value": value, "ms_per_step": dev_ms / args.steps,
        "p50_ms": float(np.percentile(done, 50)),
        "p99_ms": float(np.percentile(done, 99)),
        "kkt_max": float(max(kk)), "kkt_ok": bool(max(kk) <= KKT_TOL),
        "wall_solves_per_s": count * args.steps / walls,
        "inner_iterations_per_lp": inner / args.steps / count,
        "inner_iterations_max": {"iterations": inner_max[0],
                                 "m": inner_max[1], "n": inner_max[2]},
        "waves_per_lp_by_class": wpl,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_batch (batched solver kernel, one workgroup per "
                      "LP; 32 inner iterations per launch while tuning, "
                      "then each LP to completion)",
            "timing": "HIP events around each round; the workgroup-size "
                      "groups of a round run concurrently on their own "
                      "streams, so avg_us is round time per group launch",
            "launches": launches,
            "avg_us": round(ker_ms * 1e3 / max(launches, 1), 1),
            "alg_bytes_per_launch": round(alg / max(launches, 1)),
            "kernel_ms_per_step": ker_ms / args.steps,
            "lp_time_split": {"passes": round(ops_ms / (ops_ms + ctl_ms), 4),
                              "control": round(ctl_ms / (ops_ms + ctl_ms),
                                               4)},
            "lp_busy_ms_per_step": (ops_ms + ctl_ms) / args.steps},
        "done_ms_deciles": [round(float(np.percentile(done, q)), 1)
                            for q in range(0, 101, 10)],
        "setup_s": setup,
    }
    legs = {}
    budget = args.budget_s

    def leg(name, est_s, fn):
        if args.no_legs or name in args.skip:
            legs[name] = {"skipped": "not requested"}
            return
        if elapsed() + est_s > budget:
            legs[name] = {"skipped": "wall budget: %.0f s used of %.0f s, leg "
                                     "needs ~%.0f s" % (elapsed(), budget,
                                                        est_s)}
            return
        t = time.perf_counter()
        try:
            legs[name] = fn()
        except Exception as e:  # reported, never silently replaced
            legs[name] = {"error": str(e)[-300:]}
        log("leg %s: %.1f s" % (name, time.perf_counter() - t))

    cpu = host_cpu()
    idx64 = stratified(lps, args.fp64_sample)
    leg("fp64_single_lp", 15, lambda: fp64_single_leg(lib, lps, alms, idx64,
                                                      args.stop_kkt))
    if not args.no_cpu:
        leg("cpu_baseline", 45,
            lambda: batch_cpu_baseline(lps, args.stop_kkt, cpu))
    for a in alms:
        lib.alm_free(a)
    for m in mats:
        lib.free_matrix(m)

    def spmv_case(which):
        lp = instances.set_cover() if which == "setcover" else \
            instances.multicommodity()
        r = spmv_leg(lib, lp)
        r["workload"] = ("set-cover m=%d n=%d nnz=%d" if which == "setcover"
                         else "multicommodity flow m=%d n=%d nnz=%d") % (
                             lp.nrows, lp.ncols, lp.nnz)
        if not args.no_cpu:
            r["cpu_port"] = cpu_spmv(lp, cpu["threads"])
        return r

    leg("spmv_mcf", 40, lambda: spmv_case("mcf"))
    leg("spmv_setcover", 40, lambda: spmv_case("setcover"))
    f64 = legs.get("fp64_single_lp", {})
    res["vs_fp64_single_lp"] = (value / f64["solves_per_s"]
                                if "solves_per_s" in f64 else None)
    res["legs"] = legs
    res["cpu_baseline"] = legs.get("cpu_baseline")
    res["workload"] = ("%d independent LPs, m log-uniform [%d,%d], n=2-3m, "
                       "5 nnz/column, values +-U[0.5,2]" % (
                           count, args.batch_mmin, args.batch_mmax))
    return res


def run_single(args, lib):
    """One ALM solve of the set-cover or multicommodity LP per step."""
    from acdalm import instances, kkt
    L = lib.lib
    prec = {"mixed": 1, "fp64": 0}[args.precision]
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
    x0, lam0 = np.zeros(lp.ncols), np.zeros(lp.nrows)
    # One untimed outer iteration builds the solver graphs and the
    # reduced-precision copies of the matrix.
    lib.alm_solve_ex(alm, 1, x0, lam0, precision=prec)
    setup = time.perf_counter() - t0

    def one():
        L.acd_alm_reset(alm)
        t = time.perf_counter()
        r = lib.alm_solve_ex(alm, args.outer, x0, lam0,
                             log="stderr" if args.log else False,
                             precision=prec, stop_kkt=args.stop_kkt)
        return (time.perf_counter() - t,) + tuple(r)

    for w in range(args.warmup):
        dt, rc, *_ = one()
        log("warmup %d: %.2f s rc=%d" % (w, dt, rc))
    names = [L.acd_pass_name(j).decode() for j in range(32)]
    step_us, bytes_, lat, objs = np.zeros(32), np.zeros(32), [], []
    launches = 0
    L.acd_synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        dt, rc, x, lam, diag, _ = one()
        lat.append(dt)
        st = lib.stats()
        for name, (cnt, sus, by) in lib.profile_bytes().items():
            i = names.index(name)
            step_us[i] += sus
            bytes_[i] += by
            launches += cnt
        objs.append(diag[0])
        log("step %d: %.2f s rc=%d inner=%d outer=%d obj=%.12g"
            % (k, dt, rc, st.inner_iterations, st.outer_iterations, diag[0]))
    L.acd_synchronize()
    total = time.perf_counter() - t_start
    res_k = kkt.kkt(lp, x, lam)
    achieved = bytes_.sum() / (step_us.sum() * 1e-6) / 1e9
    top = np.argsort(-step_us)[:3]
    lat_ms = np.array(lat) * 1e3
    lib.alm_free(alm)
    lib.free_matrix(mat)
    return {
        "value": args.steps / total, "ms_per_step": total / args.steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "kkt_max": res_k["kkt"], "kkt_ok": bool(rc == 0 and
                                                 res_k["kkt"] <= KKT_TOL),
        "objective": objs[-1], "deterministic": len(set(objs)) == 1,
        "inner_iterations": int(st.inner_iterations),
        "outer_iterations": int(st.outer_iterations),
        "roofline": {
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
                for i in top if step_us[i] > 0}},
        "cpu_baseline": None, "setup_s": setup, "workload": workload,
        "nnz": int(lp.nnz), "m": int(lp.nrows), "n": int(lp.ncols),
    }


def run_cd(args, lib):
    """Parallel coordinate descent (acd_cd_solve_parallel, sample 1/16)
    against the sequential device port of cd_solve on the plain FP64 LP
    (m=2,000, n=5,000, 10 nonzeros per column): the first ALM subproblem
    (lambda = 0, unit weights, the LP's cost) from x = 0.  The target is
    the objective the sequential port reaches in 16 major steps (about 160
    epochs of single-coordinate updates; srandom(1)).  A step is one
    parallel solve to that target (max_value), seed = step number.  The
    oracle's cd_solve (the reference's algorithm) runs to the same target
    on one host core beside it."""
    import pyoracle
    from acdalm import instances
    L = lib.lib
    libc = C.CDLL(None)
    lp = instances.config0(1)
    mat = lib.make_matrix(lp, 0)
    a = lib.approx_make(mat, lp)
    x0 = np.zeros(lp.ncols)
    seq = {}
    for majors in (4, 16):
        libc.srandom(1)
        t = time.perf_counter()
        _, _, d, _ = lib.cd_solve(a, x0, majors)
        seq[majors] = (time.perf_counter() - t, d[0])
        log("sequential device cd, %d major steps: %.3f s value %.10g"
            % (majors, seq[majors][0], d[0]))
    target = seq[16][1]
    equal_updates = {}
    for epochs in (40, 160):
        _, _, d = lib.cd_solve_parallel(a, x0, epochs, seed=1)
        equal_updates[epochs] = d[0]
    names = [L.acd_pass_name(j).decode() for j in range(32)]

    def one(seed):
        t = time.perf_counter()
        rc, _, d = lib.cd_solve_parallel(a, x0, 100000, seed=seed,
                                         max_value=target)
        return time.perf_counter() - t, rc, d

    for w in range(args.warmup):
        one(1000 + w)
    lib.profile_bytes()
    step_us, bytes_, launches, lat, epochs = np.zeros(32), np.zeros(32), 0, [], []
    for k in range(args.steps):
        dt, rc, d = one(k + 1)
        if rc != 1:
            raise RuntimeError("parallel cd missed the target: rc %d" % rc)
        lat.append(dt)
        epochs.append(int(d[4]))
        for name, (cnt, sus, by) in lib.profile_bytes().items():
            i = names.index(name)
            step_us[i] += sus
            bytes_[i] += by
            launches += cnt
        log("step %d: %.3f s, %d epochs, value %.10g" % (k, dt, d[4], d[0]))
    total = float(np.sum(lat))
    i = names.index("pcd_step")
    achieved = bytes_[i] / (step_us[i] * 1e-6) / 1e9 if step_us[i] else 0.0
    ora = pyoracle.Oracle()
    om = ora.make_matrix(lp, 0)
    oa = ora.approx_make(om, lp)
    libc.srandom(1)
    t = time.perf_counter()
    _, _, d, _ = ora.cd_solve(oa, x0, 1000, max_value=target)
    cpu_s = time.perf_counter() - t
    ora.approx_free(oa)
    ora.free_matrix(om)
    lib.approx_free(a)
    lib.free_matrix(mat)
    lat_ms = np.array(lat) * 1e3
    return {
        "metric_note": "coordinate-descent solves/s to a fixed objective "
                       "(the sequential port's after 16 major steps)",
        "value": args.steps / total, "ms_per_step": total / args.steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "stop": "objective <= %.10g (max_value)" % target,
        "target_objective": target, "epochs_to_target": epochs,
        "sequential_device_port": {
            "s_4_major": seq[4][0], "s_16_major": seq[16][0],
            "value_4_major": seq[4][1], "value_16_major": seq[16][1],
            "speedup_of_parallel": seq[16][0] / (total / args.steps)},
        "parallel_at_equal_updates": {"40_epochs": equal_updates[40],
                                      "160_epochs": equal_updates[160]},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_step, pcd_step passes",
            "launches": int(launches),
            "share_of_step_kernel_time": round(step_us[i] / step_us.sum(), 3)
            if step_us.sum() else None},
        "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "solves/s",
                         "cores": 1, "kind": "port",
                         "sample": "oracle cd_solve (the reference's "
                                   "sequential coordinate descent), "
                                   "srandom(1), to the same target, one "
                                   "solve",
                         "major_steps": int(d[4]), "seconds": cpu_s},
        "workload": "cd_solve subproblem of the plain FP64 LP m=2000 "
                    "n=5000 10 nnz/column (lambda=0, unit weights)",
    }


def serve_cpu_baseline(lps, large, arrivals, large_inner, cpu, sample=32,
                       large_iters=20):
    """The oracle replayed on the same arrival trace: a server with
    `threads` host threads where a small LP takes one thread and a large LP
    all of them.  Service times: each small LP's from a log-log
    interpolation over `sample` small LPs solved by the oracle (ALM to the
    same KKT rule, LP-parallel); a large LP's from `large_iters` serial
    APPROX iterations scaled by `large_inner` (the inner iterations the
    device took for one large LP) and divided by the thread count (a
    generous, perfectly parallel CPU).  FIFO admission; p50/p99 of
    completion minus arrival."""
    threads = cpu["threads"]
    small = [i for i in range(len(lps)) if not large[i]]
    pick = [small[j] for j in stratified([lps[i] for i in small], sample)]
    res = ora_alm(lps, pick, threads, KKT_TOL)
    times = dict(zip(pick, res["seconds"]))
    spi = None
    big = [i for i in range(len(lps)) if large[i]]
    if big:
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "large.lpb")
            lps[big[0]].write_binary(path)
            r = subprocess.run([ORA_BENCH, "approx", str(large_iters), path],
                               capture_output=True, text=True, timeout=600)
        if r.returncode == 0:
            spi = json.loads(r.stdout.strip().splitlines()[-1])[
                "s_per_iteration"] / threads
    srt = sorted(times, key=lambda i: lps[i].nnz)
    xs = np.log([lps[i].nnz for i in srt])
    ys = np.log([max(times[i], 1e-6) for i in srt])
    service = np.array([
        (spi * large_inner if spi else np.nan) if large[i]
        else float(np.exp(np.interp(np.log(lps[i].nnz), xs, ys)))
        for i in range(len(lps))])
    free_at = np.zeros(threads)
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
            "cores": threads, "kind": "port",
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "sample": ("trace replay on %d host threads; small-LP service "
                       "times from %d oracle ALM solves to KKT %g; large-LP "
                       "service %s" % (threads, len(pick), KKT_TOL,
                                       "%.4g s/iteration (serial / %d) x %d "
                                       "inner iterations" % (
                                           spi * threads, threads,
                                           large_inner) if spi else "n/a"))}


def serve_leg(lib, rate, seconds, stop_kkt, prec=1, seed=9,
              large_frac=0.01, budget_gb=64.0, cpu=None):
    """Request loop (continuous batching): Poisson arrivals at `rate` per
    second for `seconds`; 99% LPs built like the batched workload, 1% large
    (m=100,000, n=300,000, 8 nonzeros per column).  Each request is
    submitted as host arrays at its arrival time (acd_server_submit_lp):
    the library builds its matrix and ALM object on the device when it is
    admitted (by estimated footprint against the byte budget) and frees
    them when it retires, so a request's latency, from its arrival to the
    step that retires it, includes its device preprocessing."""
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
    gen_s = time.perf_counter() - t0
    dev_kkt = stop_kkt * (0.9 if prec else 1.0)
    srv = lib.server(int(budget_gb * 2**30), slots=2048, large_rows=50000,
                     precision=prec, stop_kkt=dev_kkt)
    lat = np.zeros(len(lps))
    results = {}
    nxt = 0
    peak = 0
    start = time.perf_counter()
    while nxt < len(lps) or srv.pending():
        now = time.perf_counter() - start
        while nxt < len(lps) and arrivals[nxt] <= now:
            srv.submit_lp(nxt, lps[nxt], np.zeros(lps[nxt].ncols),
                          np.zeros(lps[nxt].nrows))
            nxt += 1
        done = srv.step()
        peak = max(peak, srv.bytes_admitted())
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
    if cpu is not None:
        try:
            large_inner = 0
            if large.any():
                i = int(np.flatnonzero(large)[0])
                m = lib.make_matrix(lps[i], 0)
                a = lib.alm_make(m, lps[i])
                lib.alm_solve_ex(a, 1000, np.zeros(lps[i].ncols),
                                 np.zeros(lps[i].nrows), precision=prec,
                                 stop_kkt=dev_kkt)
                large_inner = int(lib.stats().inner_iterations)
                lib.alm_free(a)
                lib.free_matrix(m)
            cpu_b = serve_cpu_baseline(lps, large, arrivals, large_inner, cpu)
        except Exception as e:  # reported, never silently replaced
            cpu_b = {"value": None, "error": str(e)[-300:]}
    small = ~large
    return {
        "workload": "Poisson arrivals at %g/s for %g s: %d requests (%d "
                    "large m=100000 n=300000), submitted as host data"
                    % (rate, seconds, len(lps), int(large.sum())),
        "value": len(lps) / end, "ms_per_step": end * 1e3,
        "offered_per_s": rate, "makespan_s": end,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "small_p50_ms": float(np.percentile(lat[small], 50)) if small.any() else None,
        "small_p99_ms": float(np.percentile(lat[small], 99)) if small.any() else None,
        "large_p50_ms": float(np.percentile(lat[large], 50)) if large.any() else None,
        "kkt_max": float(max(kk)), "kkt_ok": bool(max(kk) <= KKT_TOL),
        "all_converged": all(r[0] == 0 for r in results.values()),
        "peak_admitted_bytes": int(peak), "budget_bytes": int(budget_gb * 2**30),
        "roofline": None, "cpu_baseline": cpu_b, "generate_s": gen_s,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["batch", "setcover", "mcf",
                                           "serve", "cd"], default="batch")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--stop-kkt", type=float, default=KKT_TOL,
                    help="stop once the relative KKT residual is at most "
                         "this (0: the reference's absolute rule; single "
                         "LP workloads only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="batch: only the timed steps")
    ap.add_argument("--skip", default="",
                    help="batch: comma-separated legs to leave out "
                         "(fp64_single_lp, cpu_baseline, spmv_mcf, "
                         "spmv_setcover)")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget: legs after the timed steps run "
                         "only while they fit")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-mmin", type=int, default=50)
    ap.add_argument("--batch-mmax", type=int, default=2000)
    ap.add_argument("--fp64-sample", type=int, default=16)
    ap.add_argument("--outer", type=int, default=1000,
                    help="setcover/mcf: ALM outer-iteration cap")
    ap.add_argument("--log", action="store_true",
                    help="setcover/mcf: stream the ALM log to stderr")
    ap.add_argument("--copy-check", type=int, default=0, metavar="BYTES",
                    help="first run one plain device copy kernel of BYTES "
                         "(the reference for calibrating byte counters)")
    ap.add_argument("--rate", type=float, default=2000.0,
                    help="serve: offered requests per second")
    ap.add_argument("--seconds", type=float, default=1.0,
                    help="serve: length of the arrival window")
    ap.add_argument("--budget-gb", type=float, default=64.0,
                    help="serve: device byte budget of the request loop")
    args = ap.parse_args()
    args.skip = set(filter(None, args.skip.split(",")))
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (every LP solve runs on one device; how LPs and one "
              "large LP would shard is described in DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    from acdalm import capi
    lib = capi.load()
    if args.copy_check:
        L = lib.lib
        src = L.acd_device_alloc(args.copy_check)
        dst = L.acd_device_alloc(args.copy_check)
        if not src or not dst or L.acd_copy_device(dst, src, args.copy_check):
            raise RuntimeError("copy check failed: %s" % (lib.last_error(),))
        L.acd_device_free(src)
        L.acd_device_free(dst)
    if args.workload == "serve":
        r = serve_leg(lib, args.rate, args.seconds, args.stop_kkt,
                      prec={"mixed": 1, "fp64": 0}[args.precision],
                      budget_gb=args.budget_gb,
                      cpu=None if args.no_cpu else host_cpu())
        steps, warmup = 1, 0
    elif args.workload == "batch":
        r = run_batch(args, lib)
        steps, warmup = args.steps, args.warmup
    elif args.workload == "cd":
        r = run_cd(args, lib)
        steps, warmup = args.steps, args.warmup
    else:
        r = run_single(args, lib)
        steps, warmup = args.steps, args.warmup
    mixed = args.precision == "mixed" and args.workload != "cd"
    line = {
        "metric": r.get("metric_note", METRIC), "value": r["value"],
        "unit": "solves/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage / f64 compute" if mixed else "f64",
        "dtype_detail": ("matrix values and per-column constants (c, l, u, "
                         "step sizes) stored in FP32; iterates, products, "
                         "accumulation, multipliers, norms and the theta "
                         "recurrence in FP64" if mixed else
                         "FP64 everywhere"),
        "data": "synthetic (seeded generator, acdalm.instances)",
        "config": {"workload": r["workload"],
                   "precision": "fp64" if args.workload == "cd"
                                else args.precision,
                   "stop": r.get("stop") or (
                       "relative KKT <= %g" % args.stop_kkt
                       if args.stop_kkt > 0 else "reference rule"),
                   "parallelism": "single GPU"},
        "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
        "roofline": r["roofline"], "cpu_baseline": r["cpu_baseline"],
        "wall_s": elapsed(),
    }
    line.update({k: v for k, v in r.items() if k not in line and
                 k not in ("workload", "value", "metric_note", "stop")})
    return line


if __name__ == "__main__":
    with _StdoutToStderr():
        result = main()
    print(json.dumps(result), flush=True)
