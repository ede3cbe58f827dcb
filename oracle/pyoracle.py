"""ctypes binding of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py as the checker; the product never loads
it.  Methods mirror acdalm.capi.CLib so a test can run the same calls on
the oracle, on the reference build (oracle/_ref/) and on the GPU library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_APPROX_SO = os.path.join(HERE, "_ref", "libacdref_approx.so")
REF_CD_SO = os.path.join(HERE, "_ref", "libacdref_cd.so")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")

vp = C.c_void_p
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
sz = C.c_size_t


class AlmOptions(C.Structure):
    _fields_ = [("stop_violation", C.c_double), ("stop_pg", C.c_double),
                ("precision_floor", C.c_double), ("min_delta", C.c_double),
                ("inner", C.c_int), ("stop_kkt", C.c_double)]


def _d(a):
    return a.ctypes.data_as(dp) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def build(quiet=True):
    """Compile the oracle (and, where the reference sources exist, the
    reference build) with oracle/Makefile."""
    targets = ["oracle"]
    if os.path.isdir(os.environ.get("REFDIR", "/root/reference")):
        targets.append("ref")
    subprocess.run(["make", "-C", HERE] + targets, check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class Oracle:
    """Same helper methods as acdalm.capi.CLib, on the ora_* symbols."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            build()
        mode = getattr(os, "RTLD_LOCAL", 0) | getattr(os, "RTLD_NOW", 2)
        self.lib = L = C.CDLL(path, mode=mode)
        solve = [dp, sz, vp, sz, C.c_double, C.c_double, C.c_double, vp, sz,
                 dp, C.c_double]
        sig = {
            "ora_sparse_matrix_make": (vp, [sz, sz, sz, u32p, u32p, dp, C.c_int]),
            "ora_sparse_matrix_copy": (vp, [vp, C.c_int]),
            "ora_sparse_matrix_free": (C.c_int, [vp]),
            "ora_sparse_matrix_ncolumns": (sz, [vp]),
            "ora_sparse_matrix_nrows": (sz, [vp]),
            "ora_sparse_matrix_nnz": (sz, [vp]),
            "ora_sparse_matrix_rows": (u32p, [vp]),
            "ora_sparse_matrix_columns": (u32p, [vp]),
            "ora_sparse_matrix_values": (dp, [vp]),
            "ora_sparse_matrix_row_permute": (C.c_int, [vp, dp, sz, dp, C.c_int]),
            "ora_sparse_matrix_col_permute": (C.c_int, [vp, dp, sz, dp, C.c_int]),
            "ora_sparse_matrix_multiply": (C.c_int, [dp, sz, vp, dp, sz, C.c_int]),
            "ora_sparse_matrix_multiply_2": (C.c_int, [C.POINTER(dp), sz, vp,
                                                        C.POINTER(dp), sz, C.c_int]),
            "ora_csr_row_ptr": (u32p, [vp, C.c_int]),
            "ora_csr_columns": (u32p, [vp, C.c_int]),
            "ora_csr_values": (dp, [vp, C.c_int]),
            "ora_perm_idx": (u32p, [vp, C.c_int]),
            "ora_perm_ridx": (u32p, [vp, C.c_int]),
            "ora_morton_key": (C.c_uint64, [C.c_uint32, C.c_uint32]),
            "ora_approx_make": (vp, [vp, sz, dp, dp, sz, dp, dp, dp]),
            "ora_approx_free": (C.c_int, [vp]),
            "ora_approx_update": (C.c_int, [vp]),
            "ora_approx_rhs": (dp, [vp]),
            "ora_approx_weight": (dp, [vp]),
            "ora_approx_linear": (dp, [vp]),
            "ora_approx_lower": (dp, [vp]),
            "ora_approx_upper": (dp, [vp]),
            "ora_approx_step_v": (dp, [vp, C.c_int]),
            "ora_approx_step_inv_v": (dp, [vp, C.c_int]),
            "ora_approx_solve": (C.c_int, solve),
            "ora_mpg_solve": (C.c_int, solve),
            "ora_cd_solve": (C.c_int, solve),
            "ora_alm_make": (vp, [vp, sz, dp, sz, dp, dp, dp, dp, dp]),
            "ora_alm_free": (C.c_int, [vp]),
            "ora_alm_solve": (C.c_int, [vp, sz, dp, sz, dp, sz, vp, dp,
                                        C.POINTER(AlmOptions)]),
            "ora_alm_default_options": (None, [C.POINTER(AlmOptions)]),
            "ora_alm_kkt": (C.c_double, [vp, dp, dp]),
            "ora_alm_linear": (dp, [vp]),
            "ora_alm_approx": (vp, [vp]),
            "ora_set_reverse_sums": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    # ---- matrices ----
    def make_matrix(self, lp, permute=0):
        rows = np.ascontiguousarray(lp.rows, np.uint32)
        cols = np.ascontiguousarray(lp.cols, np.uint32)
        vals = _f64(lp.vals)
        return self.lib.ora_sparse_matrix_make(
            lp.ncols, lp.nrows, vals.shape[0], rows.ctypes.data_as(u32p),
            cols.ctypes.data_as(u32p), _d(vals), permute)

    def copy_matrix(self, m, permute):
        return self.lib.ora_sparse_matrix_copy(m, permute)

    def free_matrix(self, m):
        return self.lib.ora_sparse_matrix_free(m)

    def shape(self, m):
        return (self.lib.ora_sparse_matrix_nrows(m),
                self.lib.ora_sparse_matrix_ncolumns(m))

    def coo(self, m):
        L = self.lib
        nnz = L.ora_sparse_matrix_nnz(m)
        if nnz == 0:
            return (np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                    np.zeros(0))
        return (np.ctypeslib.as_array(L.ora_sparse_matrix_rows(m), (nnz,)).copy(),
                np.ctypeslib.as_array(L.ora_sparse_matrix_columns(m), (nnz,)).copy(),
                np.ctypeslib.as_array(L.ora_sparse_matrix_values(m), (nnz,)).copy())

    def csr(self, m, transpose):
        L = self.lib
        nr, nc = self.shape(m)
        n = nc if transpose else nr
        nnz = L.ora_sparse_matrix_nnz(m)
        ptr = np.ctypeslib.as_array(L.ora_csr_row_ptr(m, transpose), (n + 1,)).copy()
        if nnz == 0:
            return ptr, np.zeros(0, np.uint32), np.zeros(0)
        return (ptr,
                np.ctypeslib.as_array(L.ora_csr_columns(m, transpose), (nnz,)).copy(),
                np.ctypeslib.as_array(L.ora_csr_values(m, transpose), (nnz,)).copy())

    def perm(self, m, row):
        n = self.shape(m)[0 if row else 1]
        if n == 0:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        return (np.ctypeslib.as_array(self.lib.ora_perm_idx(m, row), (n,)).copy(),
                np.ctypeslib.as_array(self.lib.ora_perm_ridx(m, row), (n,)).copy())

    def permute(self, m, src, row, direction):
        src = _f64(src)
        dest = np.zeros_like(src)
        fn = (self.lib.ora_sparse_matrix_row_permute if row
              else self.lib.ora_sparse_matrix_col_permute)
        return fn(m, _d(dest), src.shape[0], _d(src), direction), dest

    def multiply(self, m, x, transpose=0, pool=None):
        nr, nc = self.shape(m)
        ny, nx = (nc, nr) if transpose else (nr, nc)
        x = _f64(x)
        y = np.zeros(ny)
        self.lib.ora_sparse_matrix_multiply(_d(y), ny, m, _d(x), nx, transpose)
        return y

    def multiply_2(self, m, x0, x1, transpose=0, pool=None):
        nr, nc = self.shape(m)
        ny, nx = (nc, nr) if transpose else (nr, nc)
        x0, x1 = _f64(x0), _f64(x1)
        y0, y1 = np.zeros(ny), np.zeros(ny)
        self.lib.ora_sparse_matrix_multiply_2((dp * 2)(_d(y0), _d(y1)), ny, m,
                                              (dp * 2)(_d(x0), _d(x1)), nx,
                                              transpose)
        return y0, y1

    def morton_key(self, col, row):
        return self.lib.ora_morton_key(col, row)

    # ---- APPROX ----
    def approx_make(self, m, lp, weight=None, linear="lp"):
        lin = lp.linear if isinstance(linear, str) else linear
        return self.lib.ora_approx_make(
            m, lp.nrows, _d(_f64(lp.rhs)),
            _d(_f64(weight)) if weight is not None else None, lp.ncols,
            _d(_f64(lin)) if lin is not None else None, _d(_f64(lp.lower)),
            _d(_f64(lp.upper)))

    def approx_free(self, a):
        return self.lib.ora_approx_free(a)

    def approx_update(self, a):
        return self.lib.ora_approx_update(a)

    def step_sizes(self, a, n, permuted=1):
        return (np.ctypeslib.as_array(self.lib.ora_approx_step_v(a, permuted), (n,)).copy(),
                np.ctypeslib.as_array(self.lib.ora_approx_step_inv_v(a, permuted), (n,)).copy())

    def _inner(self, fn, a, x0, niter, max_pg, max_value, min_delta, log,
               period, offset):
        from acdalm.capi import LogCapture
        x = _f64(x0).copy()
        diag = np.full(5, np.nan)
        cap = LogCapture(log)
        rc = fn(_d(x), x.shape[0], a, niter, max_pg, max_value, min_delta,
                cap.fp, period, _d(diag), offset)
        return rc, x, diag, cap.close()

    def approx_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                     min_delta=0.0, log=False, period=0, offset=0.0, pool=None):
        return self._inner(self.lib.ora_approx_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset)

    def mpg_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                  min_delta=0.0, log=False, period=0, offset=0.0, pool=None):
        return self._inner(self.lib.ora_mpg_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset)

    def cd_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                 min_delta=0.0, log=False, period=0, offset=0.0, pool=None):
        return self._inner(self.lib.ora_cd_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset)

    # ---- ALM ----
    def alm_make(self, m, lp):
        return self.lib.ora_alm_make(
            m, lp.nrows, _d(_f64(lp.rhs)), lp.ncols, _d(_f64(lp.linear)),
            _d(_f64(lp.lower)), _d(_f64(lp.upper)), _d(_f64(lp.lambda_lower)),
            _d(_f64(lp.lambda_upper)))

    def alm_free(self, a):
        return self.lib.ora_alm_free(a)

    def alm_kkt(self, a, x, lam):
        x, lam = _f64(x), _f64(lam)
        return float(self.lib.ora_alm_kkt(a, _d(x), _d(lam)))

    def options(self, **kw):
        o = AlmOptions()
        self.lib.ora_alm_default_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def alm_solve(self, a, niter, x0, lam0, log=False, pool=None, **opts):
        from acdalm.capi import LogCapture
        x = _f64(x0).copy()
        lam = _f64(lam0).copy()
        diag = np.full(3, np.nan)
        cap = LogCapture(log)
        o = self.options(**opts)
        rc = self.lib.ora_alm_solve(a, niter, _d(x), x.shape[0], _d(lam),
                                    lam.shape[0], cap.fp, _d(diag), C.byref(o))
        return rc, x, lam, diag, cap.close()
