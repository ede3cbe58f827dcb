"""ctypes view of a shared library that exports the acd-alm C interface
(alm.h, approx/approx.h, spmv/spmv.h, thread_pool/thread_pool.h).

The same class drives this package's GPU library (lib/libacdalm.so) and,
in tests, the reference compiled from its own sources; it is also the
binding a Python embedder would write against the original library
(see INTEGRATION.md).  Arrays are numpy float64/uint32, copied in and out.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PACKAGE_ROOT, "lib", "libacdalm.so")

vp = C.c_void_p
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
sz = C.c_size_t

_libc = C.CDLL(None)
_libc.fopen.restype = vp
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [vp]
_libc.fflush.argtypes = [vp]
_libc.fdopen.restype = vp
_libc.fdopen.argtypes = [C.c_int, C.c_char_p]
_libc.setvbuf.argtypes = [vp, C.c_char_p, C.c_int, C.c_size_t]


class AlmOptions(C.Structure):
    """struct acd_alm_options (include/acd_gpu.h)."""
    _fields_ = [("stop_violation", C.c_double), ("stop_pg", C.c_double),
                ("precision_floor", C.c_double), ("min_delta", C.c_double),
                ("precision", C.c_int), ("engine", C.c_int),
                ("stop_kkt", C.c_double)]


class BatchInfo(C.Structure):
    """struct acd_batch_info (include/acd_gpu.h)."""
    _fields_ = [("device_ms", C.c_double), ("launches", C.c_uint64),
                ("inner_iterations", C.c_uint64), ("alg_bytes", C.c_double),
                ("waves_per_lp", C.c_int * 8), ("kernel_ms", C.c_double),
                ("lp_ops_ms", C.c_double), ("lp_ctl_ms", C.c_double)]


class LpData(C.Structure):
    """struct acd_lp_data (include/acd_gpu.h)."""
    _fields_ = [("nrows", C.c_size_t), ("ncols", C.c_size_t),
                ("nnz", C.c_size_t),
                ("rows", C.POINTER(C.c_uint32)),
                ("cols", C.POINTER(C.c_uint32)),
                ("values", C.POINTER(C.c_double)),
                ("rhs", C.POINTER(C.c_double)),
                ("linear", C.POINTER(C.c_double)),
                ("lower", C.POINTER(C.c_double)),
                ("upper", C.POINTER(C.c_double)),
                ("lambda_lower", C.POINTER(C.c_double)),
                ("lambda_upper", C.POINTER(C.c_double))]


class MpsLP(C.Structure):
    """struct acd_mps_lp (include/acd_gpu.h)."""
    _fields_ = [("nrows", C.c_size_t), ("ncols", C.c_size_t),
                ("nnz", C.c_size_t), ("nstructural", C.c_size_t),
                ("rows", C.POINTER(C.c_uint32)),
                ("cols", C.POINTER(C.c_uint32)),
                ("values", C.POINTER(C.c_double)),
                ("rhs", C.POINTER(C.c_double)),
                ("linear", C.POINTER(C.c_double)),
                ("lower", C.POINTER(C.c_double)),
                ("upper", C.POINTER(C.c_double)),
                ("objective_offset", C.c_double), ("maximize", C.c_int),
                ("error", C.c_char * 128)]


class SolveStats(C.Structure):
    """struct acd_solve_stats (include/acd_gpu.h)."""
    _fields_ = [("inner_iterations", C.c_uint64),
                ("outer_iterations", C.c_uint64), ("launches", C.c_uint64),
                ("device_ms", C.c_double), ("kkt", C.c_double)]


def _d(a):
    return a.ctypes.data_as(dp) if a is not None else None


def _u(a):
    return a.ctypes.data_as(u32p)


def f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape[0] != n:
        raise ValueError("length mismatch")
    return a


class LogCapture:
    """A C FILE* writing to a temporary file; .text after close().  With
    enabled="stderr" the log goes straight to the process's stderr (the
    solver writes each line as it is produced) and .text stays empty."""

    def __init__(self, enabled=True):
        self.enabled = enabled
        self.fp = None
        self.text = ""
        self.path = None
        if enabled == "stderr":
            self.fp = _libc.fdopen(os.dup(2), b"w")
            _libc.setvbuf(self.fp, None, 2, 0)  # unbuffered
        elif enabled:
            fd, self.path = tempfile.mkstemp(prefix="acdlog")
            os.close(fd)
            self.fp = _libc.fopen(self.path.encode(), b"w")

    def close(self):
        if self.fp:
            _libc.fflush(self.fp)
            _libc.fclose(self.fp)
            self.fp = None
            if self.path:
                with open(self.path) as f:
                    self.text = f.read()
                os.unlink(self.path)
        return self.text


class CLib:
    """Bindings for a library exporting the original interface."""

    def __init__(self, path: str = LIB_PATH, *, gpu_extras: bool = None,
                 init: bool = True):
        """init=False binds the functions without sparse_matrix_init (no
        device needed: host-only entry points such as the MPS parser)."""
        mode = getattr(os, "RTLD_LOCAL", 0) | getattr(os, "RTLD_NOW", 2)
        self.path = path
        self.lib = C.CDLL(path, mode=mode)
        L = self.lib
        sig = {
            "sparse_matrix_init": (None, []),
            "sparse_matrix_make": (vp, [sz, sz, sz, u32p, u32p, dp, C.c_int]),
            "sparse_matrix_copy": (vp, [vp, C.c_int]),
            "sparse_matrix_free": (C.c_int, [vp]),
            "sparse_matrix_ncolumns": (sz, [vp]),
            "sparse_matrix_nrows": (sz, [vp]),
            "sparse_matrix_nnz": (sz, [vp]),
            "sparse_matrix_rows": (u32p, [vp]),
            "sparse_matrix_columns": (u32p, [vp]),
            "sparse_matrix_values": (dp, [vp]),
            "sparse_matrix_row_permute": (C.c_int, [vp, dp, sz, dp, C.c_int]),
            "sparse_matrix_col_permute": (C.c_int, [vp, dp, sz, dp, C.c_int]),
            "sparse_matrix_multiply": (C.c_int, [dp, sz, vp, dp, sz, C.c_int, vp]),
            "sparse_matrix_multiply_2": (C.c_int, [C.POINTER(dp), sz, vp,
                                                    C.POINTER(dp), sz, C.c_int, vp]),
            "sparse_matrix_read": (vp, [vp, C.c_int]),
            "approx_make": (vp, [vp, sz, dp, dp, sz, dp, dp, dp]),
            "approx_free": (C.c_int, [vp]),
            "approx_matrix": (vp, [vp]),
            "approx_nrhs": (sz, [vp]),
            "approx_nvars": (sz, [vp]),
            "approx_rhs": (dp, [vp]),
            "approx_weight": (dp, [vp]),
            "approx_linear": (dp, [vp]),
            "approx_lower": (dp, [vp]),
            "approx_upper": (dp, [vp]),
            "approx_update": (C.c_int, [vp]),
            "approx_solve": (C.c_int, [dp, sz, vp, sz, C.c_double, C.c_double,
                                       C.c_double, vp, sz, dp, C.c_double, vp]),
            "mpg_solve": (C.c_int, [dp, sz, vp, sz, C.c_double, C.c_double,
                                    C.c_double, vp, sz, dp, C.c_double, vp]),
            "cd_solve": (C.c_int, [dp, sz, vp, sz, C.c_double, C.c_double,
                                   C.c_double, vp, sz, dp, C.c_double, vp]),
            "alm_make": (vp, [vp, sz, dp, sz, dp, dp, dp, dp, dp]),
            "alm_free": (C.c_int, [vp]),
            "alm_matrix": (vp, [vp]),
            "alm_nrhs": (sz, [vp]),
            "alm_nvars": (sz, [vp]),
            "alm_rhs": (dp, [vp]),
            "alm_linear": (dp, [vp]),
            "alm_lower": (dp, [vp]),
            "alm_upper": (dp, [vp]),
            "alm_lambda_lower": (dp, [vp]),
            "alm_lambda_upper": (dp, [vp]),
            "alm_solve": (C.c_int, [vp, sz, dp, sz, dp, sz, vp, dp, vp]),
            "alm_read": (vp, [vp]),
            "read_doubles": (None, [vp, dp, sz]),
            "thread_pool_init": (vp, [C.c_uint]),
            "thread_pool_free": (None, [vp]),
            "thread_pool_count": (sz, [vp]),
        }
        self.missing = []
        for name, (res, args) in sig.items():
            try:
                fn = getattr(L, name)
            except AttributeError:
                self.missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
        if gpu_extras is None:
            gpu_extras = hasattr(L, "acd_set_device")
        self.gpu = gpu_extras
        if gpu_extras:
            self._bind_gpu()
        if init:
            L.sparse_matrix_init()

    def _bind_gpu(self):
        L = self.lib
        for name, res, args in [
            ("acd_set_device", C.c_int, [C.c_int]),
            ("acd_last_error", C.c_int, [C.c_char_p, sz]),
            ("acd_synchronize", C.c_int, []),
            ("acd_device_alloc", vp, [sz]),
            ("acd_device_free", None, [vp]),
            ("acd_copy_to_device", C.c_int, [vp, vp, sz]),
            ("acd_copy_device", C.c_int, [vp, vp, sz]),
            ("acd_copy_to_host", C.c_int, [vp, vp, sz]),
            ("acd_spmv_device", C.c_int, [vp, C.c_int, vp, vp]),
            ("acd_spmv_device_time", C.c_int, [vp, C.c_int, vp, vp, C.c_int,
                                               C.POINTER(C.c_double)]),
            ("acd_matrix_stored_entries", C.c_uint64, [vp, C.c_int]),
            ("acd_alm_default_options", None, [C.POINTER(AlmOptions)]),
            ("alm_solve_ex", C.c_int, [vp, sz, dp, sz, dp, sz, vp, dp,
                                       C.POINTER(AlmOptions)]),
            ("acd_last_solve_stats", None, [C.POINTER(SolveStats)]),
            ("acd_server_make", vp, [sz, sz, sz, C.POINTER(AlmOptions)]),
            ("acd_server_free", None, [vp]),
            ("acd_server_submit", C.c_int, [vp, vp, dp, dp, dp,
                                            C.POINTER(C.c_int), C.c_uint64]),
            ("acd_server_step", C.c_int, [vp, C.POINTER(C.c_uint64), sz,
                                          C.POINTER(sz)]),
            ("acd_server_pending", sz, [vp]),
            ("acd_server_submit_lp", C.c_int, [vp, C.POINTER(LpData), dp, dp,
                                               dp, C.POINTER(C.c_int),
                                               C.c_uint64]),
            ("acd_server_footprint", sz, [sz, sz, sz]),
            ("acd_server_bytes_admitted", sz, [vp]),
            ("acd_device_memory_info", C.c_int, [C.POINTER(sz),
                                                 C.POINTER(sz)]),
            ("alm_solve_batch", C.c_int, [sz, C.POINTER(vp), sz,
                                          C.POINTER(dp), C.POINTER(dp),
                                          C.POINTER(dp), C.POINTER(C.c_int),
                                          C.POINTER(AlmOptions), dp,
                                          C.POINTER(C.c_uint64),
                                          C.POINTER(BatchInfo)]),
            ("acd_alm_reset", C.c_int, [vp]),
            ("acd_cd_solve_parallel", C.c_int, [dp, sz, vp, sz, C.c_double,
                                                C.c_double, C.c_double,
                                                C.c_double, C.c_uint64, dp]),
            ("acd_spmv_variant_time", C.c_int, [vp, C.c_int, vp, vp, C.c_int,
                                                C.c_int, C.POINTER(C.c_double)]),
            ("acd_spmv_set_variant", C.c_int, [C.c_int]),
            ("acd_spmv_autotune", C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
            ("acd_spmv_rotate_time", C.c_int, [C.POINTER(vp), sz, C.c_int,
                                               C.POINTER(vp), C.POINTER(vp),
                                               C.c_int, C.POINTER(C.c_double)]),
            ("acd_last_solve_profile", None, [C.POINTER(C.c_uint64),
                                              C.POINTER(C.c_double), C.c_int]),
            ("acd_last_solve_bytes", None, [C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.c_int]),
            ("acd_pass_name", C.c_char_p, [C.c_int]),
            ("acd_mps_parse", C.c_int, [vp, C.POINTER(MpsLP)]),
            ("acd_mps_free", None, [C.POINTER(MpsLP)]),
            ("acd_alm_read_mps", vp, [vp, C.POINTER(MpsLP)]),
        ]:
            fn = getattr(L, name, None)
            if fn is None:
                self.missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args

    # ---- matrices ----
    def make_matrix(self, lp, permute=0, rows=None, cols=None, vals=None,
                    nrows=None, ncols=None):
        rows = np.ascontiguousarray(lp.rows if rows is None else rows, np.uint32)
        cols = np.ascontiguousarray(lp.cols if cols is None else cols, np.uint32)
        vals = f64(lp.vals if vals is None else vals)
        nrows = lp.nrows if nrows is None else nrows
        ncols = lp.ncols if ncols is None else ncols
        m = self.lib.sparse_matrix_make(ncols, nrows, vals.shape[0], _u(rows),
                                        _u(cols), _d(vals), permute)
        if not m:
            raise MemoryError("sparse_matrix_make returned NULL")
        return m

    def copy_matrix(self, m, permute):
        return self.lib.sparse_matrix_copy(m, permute)

    def free_matrix(self, m):
        return self.lib.sparse_matrix_free(m)

    def coo(self, m):
        L = self.lib
        nnz = L.sparse_matrix_nnz(m)
        if nnz == 0:
            return (np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                    np.zeros(0))
        r = np.ctypeslib.as_array(L.sparse_matrix_rows(m), (nnz,)).copy()
        c = np.ctypeslib.as_array(L.sparse_matrix_columns(m), (nnz,)).copy()
        v = np.ctypeslib.as_array(L.sparse_matrix_values(m), (nnz,)).copy()
        return r, c, v

    def shape(self, m):
        return (self.lib.sparse_matrix_nrows(m),
                self.lib.sparse_matrix_ncolumns(m))

    def permute(self, m, src, row, direction):
        src = f64(src)
        dest = np.zeros_like(src)
        fn = (self.lib.sparse_matrix_row_permute if row
              else self.lib.sparse_matrix_col_permute)
        rc = fn(m, _d(dest), src.shape[0], _d(src), direction)
        return rc, dest

    def multiply(self, m, x, transpose=0, pool=None):
        nr, nc = self.shape(m)
        ny, nx = (nc, nr) if transpose else (nr, nc)
        x = f64(x, nx)
        y = np.zeros(ny)
        self.lib.sparse_matrix_multiply(_d(y), ny, m, _d(x), nx, transpose,
                                        pool)
        return y

    def multiply_2(self, m, x0, x1, transpose=0, pool=None):
        nr, nc = self.shape(m)
        ny, nx = (nc, nr) if transpose else (nr, nc)
        x0, x1 = f64(x0, nx), f64(x1, nx)
        y0, y1 = np.zeros(ny), np.zeros(ny)
        ys = (dp * 2)(_d(y0), _d(y1))
        xs = (dp * 2)(_d(x0), _d(x1))
        self.lib.sparse_matrix_multiply_2(ys, ny, m, xs, nx, transpose, pool)
        return y0, y1

    # ---- APPROX ----
    def approx_make(self, m, lp, weight=None, linear="lp"):
        lin = lp.linear if isinstance(linear, str) else linear
        a = self.lib.approx_make(
            m, lp.nrows, _d(f64(lp.rhs)),
            _d(f64(weight)) if weight is not None else None, lp.ncols,
            _d(f64(lin)) if lin is not None else None, _d(f64(lp.lower)),
            _d(f64(lp.upper)))
        if not a:
            raise MemoryError("approx_make returned NULL")
        return a

    def approx_free(self, a):
        return self.lib.approx_free(a)

    def approx_update(self, a):
        return self.lib.approx_update(a)

    def _inner(self, fn, a, x0, niter, max_pg, max_value, min_delta, log,
               period, offset, pool):
        x = f64(x0).copy()
        diag = np.full(5, np.nan)
        cap = LogCapture(log)
        rc = fn(_d(x), x.shape[0], a, niter, max_pg, max_value, min_delta,
                cap.fp, period, _d(diag), offset, pool)
        return rc, x, diag, cap.close()

    def approx_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                     min_delta=0.0, log=False, period=0, offset=0.0,
                     pool=None):
        return self._inner(self.lib.approx_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset, pool)

    def mpg_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                  min_delta=0.0, log=False, period=0, offset=0.0, pool=None):
        return self._inner(self.lib.mpg_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset, pool)

    def cd_solve(self, a, x0, niter, max_pg=0.0, max_value=-np.inf,
                 min_delta=0.0, log=False, period=0, offset=0.0, pool=None):
        return self._inner(self.lib.cd_solve, a, x0, niter, max_pg,
                           max_value, min_delta, log, period, offset, pool)

    def cd_solve_parallel(self, a, x0, niter, sample=1 / 16., seed=1,
                          max_pg=0.0, max_value=-np.inf, offset=0.0):
        """acd_cd_solve_parallel: (return code, x, diagnosis[5])."""
        x = f64(x0).copy()
        diag = np.full(5, np.nan)
        rc = self.lib.acd_cd_solve_parallel(_d(x), x.shape[0], a, niter,
                                            max_pg, max_value, offset,
                                            sample, seed, _d(diag))
        return rc, x, diag

    # ---- ALM ----
    def alm_make(self, m, lp):
        a = self.lib.alm_make(m, lp.nrows, _d(f64(lp.rhs)), lp.ncols,
                              _d(f64(lp.linear)), _d(f64(lp.lower)),
                              _d(f64(lp.upper)), _d(f64(lp.lambda_lower)),
                              _d(f64(lp.lambda_upper)))
        if not a:
            raise MemoryError("alm_make returned NULL")
        return a

    def alm_free(self, a):
        return self.lib.alm_free(a)

    def alm_solve(self, a, niter, x0, lam0, log=False, pool=None):
        x = f64(x0).copy()
        lam = f64(lam0).copy()
        diag = np.full(3, np.nan)
        cap = LogCapture(log)
        rc = self.lib.alm_solve(a, niter, _d(x), x.shape[0], _d(lam),
                                lam.shape[0], cap.fp, _d(diag), pool)
        return rc, x, lam, diag, cap.close()

    def alm_solve_ex(self, a, niter, x0, lam0, log=False, **opts):
        o = AlmOptions()
        self.lib.acd_alm_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        x = f64(x0).copy()
        lam = f64(lam0).copy()
        diag = np.full(3, np.nan)
        cap = LogCapture(log)
        rc = self.lib.alm_solve_ex(a, niter, _d(x), x.shape[0], _d(lam),
                                   lam.shape[0], cap.fp, _d(diag), C.byref(o))
        return rc, x, lam, diag, cap.close()

    def alm_solve_batch(self, alms, niter, x0s, lam0s, **opts):
        """alm_solve_batch over ALM objects; returns ([(ret, x, lam, diag)],
        done_ms, BatchInfo with .inner = per-LP inner iterations)."""
        n = len(alms)
        o = AlmOptions()
        self.lib.acd_alm_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        xs = [f64(v).copy() for v in x0s]
        ls = [f64(v).copy() for v in lam0s]
        ds = [np.full(3, np.nan) for _ in range(n)]
        arr_a = (vp * n)(*alms)
        arr_x = (dp * n)(*[_d(v) for v in xs])
        arr_l = (dp * n)(*[_d(v) for v in ls])
        arr_d = (dp * n)(*[_d(v) for v in ds])
        rets = (C.c_int * n)()
        done = np.zeros(n)
        inner = np.zeros(n, dtype=np.uint64)
        info = BatchInfo()
        self.lib.alm_solve_batch(n, arr_a, niter, arr_x, arr_l, arr_d, rets,
                                 C.byref(o), _d(done),
                                 inner.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 C.byref(info))
        info.inner = inner
        return ([(int(rets[i]), xs[i], ls[i], ds[i]) for i in range(n)],
                done, info)

    def server(self, hbm_budget, slots=2048, large_rows=50000, **opts):
        return Server(self, hbm_budget, slots, large_rows, **opts)

    def profile(self):
        """{pass name: (launches, microseconds)} of the last solve."""
        cnt = (C.c_uint64 * 32)()
        us = (C.c_double * 32)()
        self.lib.acd_last_solve_profile(cnt, us, 32)
        out = {}
        for i in range(32):
            if cnt[i]:
                out[self.lib.acd_pass_name(i).decode()] = (int(cnt[i]), float(us[i]))
        return out

    def profile_bytes(self):
        """{pass name: (launches, step microseconds, algorithmic bytes)}."""
        cnt = (C.c_uint64 * 32)()
        us = (C.c_double * 32)()
        sus = (C.c_double * 32)()
        by = (C.c_double * 32)()
        self.lib.acd_last_solve_profile(cnt, us, 32)
        self.lib.acd_last_solve_bytes(sus, by, 32)
        out = {}
        for i in range(32):
            if cnt[i]:
                out[self.lib.acd_pass_name(i).decode()] = (
                    int(cnt[i]), float(sus[i]), float(by[i]))
        return out

    # ---- MPS ----
    def _fopen(self, path):
        libc = C.CDLL(None)
        libc.fopen.restype = vp
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [vp]
        f = libc.fopen(path.encode(), b"r")
        if not f:
            raise OSError(f"cannot open {path}")
        return libc, f

    def mps_parse(self, path):
        """acd_mps_parse: dict of numpy arrays (copies), or raises
        ValueError with the parser's message."""
        libc, f = self._fopen(path)
        lp = MpsLP()
        rc = self.lib.acd_mps_parse(f, C.byref(lp))
        libc.fclose(f)
        if rc:
            raise ValueError(lp.error.decode())
        arr = lambda p, n: np.ctypeslib.as_array(p, (n,)).copy() if n else \
            np.zeros(0, np.dtype(p._type_))
        out = {"nrows": lp.nrows, "ncols": lp.ncols,
               "nstructural": lp.nstructural,
               "objective_offset": lp.objective_offset,
               "maximize": bool(lp.maximize),
               "rows": arr(lp.rows, lp.nnz), "cols": arr(lp.cols, lp.nnz),
               "values": arr(lp.values, lp.nnz),
               "rhs": arr(lp.rhs, lp.nrows),
               "linear": arr(lp.linear, lp.ncols),
               "lower": arr(lp.lower, lp.ncols),
               "upper": arr(lp.upper, lp.ncols)}
        self.lib.acd_mps_free(C.byref(lp))
        return out

    def alm_read_mps(self, path):
        """acd_alm_read_mps: (alm, info) or raises ValueError."""
        libc, f = self._fopen(path)
        info = MpsLP()
        a = self.lib.acd_alm_read_mps(f, C.byref(info))
        libc.fclose(f)
        if not a:
            raise ValueError(info.error.decode())
        return a, info

    def stats(self):
        st = SolveStats()
        self.lib.acd_last_solve_stats(C.byref(st))
        return st

    # ---- GPU extras ----
    def memory_info(self):
        """(free, total) device bytes."""
        f, t = sz(), sz()
        if self.lib.acd_device_memory_info(C.byref(f), C.byref(t)):
            raise RuntimeError("acd_device_memory_info failed")
        return f.value, t.value

    def last_error(self):
        buf = C.create_string_buffer(256)
        code = self.lib.acd_last_error(buf, 256)
        return code, buf.value.decode(errors="replace")


class Server:
    """acd_server_* (include/acd_gpu.h) with the request buffers kept alive
    until completion."""

    def __init__(self, lib, hbm_budget, slots, large_rows, **opts):
        self.lib = lib
        o = AlmOptions()
        lib.lib.acd_alm_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        self.s = lib.lib.acd_server_make(hbm_budget, slots, large_rows,
                                         C.byref(o))
        if not self.s:
            raise RuntimeError("acd_server_make failed: %s" % (lib.last_error(),))
        self.cap = slots + 8
        self.live = {}

    def submit(self, tag, alm, x0, lam0):
        x = f64(x0).copy()
        lam = f64(lam0).copy()
        diag = np.full(3, np.nan)
        ret = C.c_int(-1)
        self.live[tag] = (x, lam, diag, ret)
        rc = self.lib.lib.acd_server_submit(self.s, alm, _d(x), _d(lam),
                                            _d(diag), C.byref(ret), tag)
        if rc:
            raise RuntimeError("acd_server_submit failed")

    def step(self):
        """One round; returns [(tag, ret, x, lam, diag)] completed."""
        done = (C.c_uint64 * self.cap)()
        n = C.c_size_t(0)
        if self.lib.lib.acd_server_step(self.s, done, self.cap, C.byref(n)):
            raise RuntimeError("acd_server_step failed: %s"
                               % (self.lib.last_error(),))
        out = []
        for i in range(min(n.value, self.cap)):
            x, lam, diag, ret = self.live.pop(done[i])[:4]
            out.append((int(done[i]), ret.value, x, lam, diag))
        return out

    def submit_lp(self, tag, lp, x0, lam0):
        """Queue an LP given as host arrays; the library builds it on the
        device when it is admitted and frees it when it completes."""
        x = f64(x0).copy()
        lam = f64(lam0).copy()
        diag = np.full(3, np.nan)
        ret = C.c_int(-1)
        keep = [np.ascontiguousarray(lp.rows, np.uint32),
                np.ascontiguousarray(lp.cols, np.uint32)] + \
            [f64(a) for a in (lp.vals, lp.rhs, lp.linear, lp.lower, lp.upper,
                              lp.lambda_lower, lp.lambda_upper)]
        d = LpData(lp.nrows, lp.ncols, lp.nnz, _u(keep[0]), _u(keep[1]),
                   *[_d(a) for a in keep[2:]])
        self.live[tag] = (x, lam, diag, ret, keep, d)
        rc = self.lib.lib.acd_server_submit_lp(self.s, C.byref(d), _d(x),
                                               _d(lam), _d(diag),
                                               C.byref(ret), tag)
        if rc:
            raise RuntimeError("acd_server_submit_lp failed")

    def bytes_admitted(self):
        return self.lib.lib.acd_server_bytes_admitted(self.s)

    def pending(self):
        return self.lib.lib.acd_server_pending(self.s)

    def close(self):
        if self.s:
            self.lib.lib.acd_server_free(self.s)
            self.s = None


_LIB = None


def load(path: str = LIB_PATH) -> CLib:
    """The GPU library, loaded once per process.  Raises if it is not
    built: there is no fallback."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is not built; run __graft_entry__.build() or "
                "make -C acd-alm_amd")
        _LIB = CLib(path)
    return _LIB
