"""Seeded synthetic LP instances:  min c'x  s.t.  A x = b,  l <= x <= u.

Each generator returns an :class:`LP` whose triplets (rows, columns,
values) follow the coordinate convention of spmv/spmv.h:15-24 of
pkhuong/acd-alm (no repeated (row, column) pair, any order).  The same seed
gives the same bytes on any machine with this numpy (PCG64 bit generator);
``LP.digest()`` fingerprints an instance so fixtures can detect drift.

Workloads:
  * ``random_lp``     -- equality LP with k distinct uniform rows per
                         column, values +-U[0.5, 2], feasible b = A x0.
  * ``set_cover``     -- covering relaxation A x - s = 1 with 0/1 columns.
  * ``multicommodity``-- grid-graph multicommodity flow with arc capacities.
  * ``batch``         -- many random LPs of log-uniform sizes.
"""
from __future__ import annotations

import hashlib
import struct
from dataclasses import dataclass, field

import numpy as np

INF = float("inf")


@dataclass
class LP:
    nrows: int
    ncols: int
    rows: np.ndarray      # uint32[nnz]
    cols: np.ndarray      # uint32[nnz]
    vals: np.ndarray      # float64[nnz]
    rhs: np.ndarray       # float64[nrows]
    linear: np.ndarray    # float64[ncols]
    lower: np.ndarray     # float64[ncols]
    upper: np.ndarray     # float64[ncols]
    lambda_lower: np.ndarray = None   # float64[nrows]
    lambda_upper: np.ndarray = None   # float64[nrows]
    name: str = ""
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.rows = np.ascontiguousarray(self.rows, dtype=np.uint32)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.uint32)
        self.vals = np.ascontiguousarray(self.vals, dtype=np.float64)
        for f in ("rhs", "linear", "lower", "upper"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f),
                                                  dtype=np.float64))
        if self.lambda_lower is None:
            self.lambda_lower = np.full(self.nrows, -INF)
        if self.lambda_upper is None:
            self.lambda_upper = np.full(self.nrows, INF)
        self.lambda_lower = np.ascontiguousarray(self.lambda_lower,
                                                 dtype=np.float64)
        self.lambda_upper = np.ascontiguousarray(self.lambda_upper,
                                                 dtype=np.float64)

    @property
    def nnz(self) -> int:
        return int(self.vals.shape[0])

    def digest(self) -> str:
        h = hashlib.sha256()
        h.update(struct.pack("<QQQ", self.nrows, self.ncols, self.nnz))
        for a in (self.rows, self.cols, self.vals, self.rhs, self.linear,
                  self.lower, self.upper, self.lambda_lower,
                  self.lambda_upper):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    def dense(self) -> np.ndarray:
        a = np.zeros((self.nrows, self.ncols))
        a[self.rows, self.cols] = self.vals
        return a

    def write_binary(self, path: str) -> None:
        """Instance file read by oracle/ref_harness.c ("ACDLP001")."""
        with open(path, "wb") as f:
            f.write(b"ACDLP001")
            f.write(struct.pack("<QQQ", self.nrows, self.ncols, self.nnz))
            for a in (self.rows, self.cols, self.vals, self.rhs, self.linear,
                      self.lower, self.upper, self.lambda_lower,
                      self.lambda_upper):
                f.write(np.ascontiguousarray(a).tobytes())

    def write_text(self, path: str) -> None:
        """Text format of alm.h:105-117 / spmv.h:76-83."""
        def fmt(v):
            if v == INF:
                return "inf"
            if v == -INF:
                return "-inf"
            return repr(float(v))
        with open(path, "w") as f:
            f.write(f"{self.nrows} {self.ncols} {self.nnz}\n")
            for r, c, v in zip(self.rows, self.cols, self.vals):
                f.write(f"{int(r)} {int(c)} {fmt(v)}\n")
            for arr in (self.rhs, self.linear, self.lower, self.upper,
                        self.lambda_lower, self.lambda_upper):
                f.write(" ".join(fmt(v) for v in arr) + "\n")


def _distinct_rows(rng: np.random.Generator, m: int, n: int, k: int):
    """(n, k) array: for every column, k distinct rows drawn uniformly."""
    if k > m:
        raise ValueError("more entries per column than rows")
    rows = rng.integers(0, m, size=(n, k), dtype=np.int64)
    while True:
        s = np.sort(rows, axis=1)
        bad = np.nonzero((s[:, 1:] == s[:, :-1]).any(axis=1))[0]
        if bad.size == 0:
            return rows
        rows[bad] = rng.integers(0, m, size=(bad.size, k), dtype=np.int64)


def _matvec(rows, cols, vals, x, m):
    """A x with a fixed summation order (bincount accumulates in entry
    order), so b is reproducible bit for bit."""
    return np.bincount(rows, weights=vals * x[cols], minlength=m)


def random_lp(m: int, n: int, k: int, seed: int, *, upper: float = 10.0,
              name: str = "") -> LP:
    """Feasible equality LP: each column has k distinct uniform rows with
    values sign*U[0.5, 2]; x0 ~ U[0, upper]; b = A x0; c ~ U[-1, 1];
    0 <= x <= upper."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = _distinct_rows(rng, m, n, k)
    cols = np.repeat(np.arange(n, dtype=np.int64), k)
    rows = r.reshape(-1)
    mag = rng.uniform(0.5, 2.0, size=rows.shape[0])
    sign = np.where(rng.random(rows.shape[0]) < 0.5, -1.0, 1.0)
    vals = sign * mag
    x0 = rng.uniform(0.0, upper, size=n)
    b = _matvec(rows, cols, vals, x0, m)
    c = rng.uniform(-1.0, 1.0, size=n)
    return LP(m, n, rows, cols, vals, b, c, np.zeros(n), np.full(n, upper),
              name=name or f"random_lp(m={m},n={n},k={k},seed={seed})",
              meta={"x0": x0})


def config0(seed: int = 1) -> LP:
    """Plain FP64 single LP: m=2000, n=5000, 10 entries per column."""
    return random_lp(2000, 5000, 10, seed, name=f"c0(seed={seed})")


def set_cover(m: int = 200_000, n: int = 1_000_000, seed: int = 2, *,
              kmin: int = 3, kmax: int = 8, sign_multipliers: bool = True
              ) -> LP:
    """Covering relaxation: each x column covers k ~ U{kmin..kmax} distinct
    rows with coefficient 1; row j has slack s_j (coefficient -1, [0, inf))
    so A x - s = 1 encodes A x >= 1.  c ~ U[1, 100] on x, 0 on slacks,
    0 <= x <= 1.  With sign_multipliers the multipliers are bounded below
    by 0 (the rows are >= constraints)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = rng.integers(kmin, kmax + 1, size=n)
    r = _distinct_rows(rng, m, n, kmax)
    keep = np.arange(kmax)[None, :] < k[:, None]
    xrows = r[keep]
    xcols = np.repeat(np.arange(n, dtype=np.int64), k)
    rows = np.concatenate([xrows, np.arange(m, dtype=np.int64)])
    cols = np.concatenate([xcols, n + np.arange(m, dtype=np.int64)])
    vals = np.concatenate([np.ones(xrows.shape[0]), -np.ones(m)])
    c = np.concatenate([rng.uniform(1.0, 100.0, size=n), np.zeros(m)])
    lower = np.zeros(n + m)
    upper = np.concatenate([np.ones(n), np.full(m, INF)])
    llo = np.zeros(m) if sign_multipliers else None
    return LP(m, n + m, rows, cols, vals, np.ones(m), c, lower, upper,
              lambda_lower=llo, name=f"set_cover(m={m},n={n},seed={seed})")


def multicommodity(grid: int = 100, ncommodities: int = 64, seed: int = 3
                   ) -> LP:
    """Multicommodity flow on a grid x grid graph with arcs both ways.

    Rows: per commodity and node, out - in = d (1[v=s] - 1[v=t]); per arc,
    sum_k x_ak + s_a = cap_a.  Columns: x_ak (arc-major within commodity
    blocks: column k*narcs + a), then one slack per arc.  Costs U[1, 5] per
    arc (shared by commodities), demands U[1, 10], capacities U[20, 80]
    raised where the routing of every commodity along an L-shaped path
    would exceed them, which certifies feasibility."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = grid
    nodes = g * g
    tails, heads = [], []
    for r in range(g):
        for c in range(g - 1):
            u, v = r * g + c, r * g + c + 1
            tails += [u, v]
            heads += [v, u]
    for r in range(g - 1):
        for c in range(g):
            u, v = r * g + c, (r + 1) * g + c
            tails += [u, v]
            heads += [v, u]
    tails = np.array(tails, dtype=np.int64)
    heads = np.array(heads, dtype=np.int64)
    narcs = tails.shape[0]
    arc_of = {(int(t), int(h)): i for i, (t, h) in enumerate(zip(tails, heads))}
    cost = rng.uniform(1.0, 5.0, size=narcs)
    cap = rng.uniform(20.0, 80.0, size=narcs)
    K = ncommodities
    src = rng.integers(0, nodes, size=K)
    dst = rng.integers(0, nodes, size=K)
    for k in range(K):
        while dst[k] == src[k]:
            dst[k] = rng.integers(0, nodes)
    demand = rng.uniform(1.0, 10.0, size=K)
    load = np.zeros(narcs)
    for k in range(K):   # row-first then column-first L path
        r0, c0 = divmod(int(src[k]), g)
        r1, c1 = divmod(int(dst[k]), g)
        cur = (r0, c0)
        step = 1 if c1 > c0 else -1
        while cur[1] != c1:
            nxt = (cur[0], cur[1] + step)
            load[arc_of[(cur[0] * g + cur[1], nxt[0] * g + nxt[1])]] += demand[k]
            cur = nxt
        step = 1 if r1 > r0 else -1
        while cur[0] != r1:
            nxt = (cur[0] + step, cur[1])
            load[arc_of[(cur[0] * g + cur[1], nxt[0] * g + nxt[1])]] += demand[k]
            cur = nxt
    cap = np.maximum(cap, load)
    ncons = K * nodes
    m = ncons + narcs
    n = K * narcs + narcs
    kk = np.repeat(np.arange(K, dtype=np.int64), narcs)
    aa = np.tile(np.arange(narcs, dtype=np.int64), K)
    xcol = kk * narcs + aa
    rows = np.concatenate([kk * nodes + tails[aa], kk * nodes + heads[aa],
                           ncons + aa, ncons + np.arange(narcs)])
    cols = np.concatenate([xcol, xcol, xcol, K * narcs + np.arange(narcs)])
    vals = np.concatenate([np.ones(K * narcs), -np.ones(K * narcs),
                           np.ones(K * narcs), np.ones(narcs)])
    b = np.zeros(m)
    b[np.arange(K) * nodes + src] += demand
    b[np.arange(K) * nodes + dst] -= demand
    b[ncons:] = cap
    c = np.concatenate([np.tile(cost, K), np.zeros(narcs)])
    lower = np.zeros(n)
    upper = np.full(n, INF)
    return LP(m, n, rows, cols, vals, b, c, lower, upper,
              name=f"mcf(grid={g},K={K},seed={seed})",
              meta={"load": load, "cap": cap})


def batch(count: int = 4096, seed: int = 4, *, mmin: int = 50,
          mmax: int = 2000, k: int = 5, nmin: float = 2.0, nmax: float = 3.0
          ) -> list:
    """Independent LPs built like random_lp: m log-uniform in [mmin, mmax],
    n = round(U[nmin, nmax] m), k entries per column."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ms = np.rint(np.exp(rng.uniform(np.log(mmin), np.log(mmax), size=count)))
    ns = np.rint(rng.uniform(nmin, nmax, size=count) * ms)
    seeds = rng.integers(0, 2**62, size=count)
    return [random_lp(int(m), int(n), k, int(s), name=f"batch[{i}]")
            for i, (m, n, s) in enumerate(zip(ms, ns, seeds))]


def tiny_lp() -> LP:
    """min x1 + 2 x2  s.t.  x1 + x2 = 1,  0 <= x <= 10:  x* = (1, 0),
    objective 1."""
    return LP(1, 2, [0, 0], [0, 1], [1.0, 1.0], [1.0], [1.0, 2.0],
              [0.0, 0.0], [10.0, 10.0], name="tiny")
