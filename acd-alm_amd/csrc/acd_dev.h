/* acd_dev.h -- the thin C ABI between the host C layer (host_*.c) and the
 * HIP device layer (dev_*.hip).
 *
 * Host C code never includes HIP headers: every device allocation, copy,
 * launch and synchronisation goes through the functions below, which the
 * .hip translation units implement with extern "C" linkage.  All device
 * work is ordered on one library stream (acd_dev_stream()).
 */
#ifndef ACD_DEV_H
#define ACD_DEV_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime (dev_runtime.hip) ---- */
int acd_dev_init(void);                 /* idempotent; 0 on success */
int acd_dev_set_device(int device);
int acd_dev_device(void);
void * acd_dev_malloc(size_t bytes);    /* zero-filled; NULL on failure */
void acd_dev_free(void * p);
int acd_dev_h2d(void * dst, const void * src, size_t bytes);
int acd_dev_d2h(void * dst, const void * src, size_t bytes);
int acd_dev_d2d(void * dst, const void * src, size_t bytes);
/* dst <- src by a plain HIP copy kernel (16 bytes per lane, grid-stride;
 * bytes and both addresses multiples of 16), on the library stream. */
int acd_dev_copy_kernel(void * dst, const void * src, size_t bytes);
int acd_dev_memset(void * dst, int byte, size_t bytes);
int acd_dev_sync(void);
int acd_dev_mem_info(size_t * free_bytes, size_t * total_bytes);
void * acd_dev_stream(void);            /* hipStream_t */
int acd_dev_set_stream(void * stream);
/* An extra stream (default priority, non-blocking) and its release. */
void * acd_dev_stream_create(void);
void acd_dev_stream_destroy(void * stream);
int acd_dev_cu_count(void);
/* First HIP error seen by the library since the last call (0 = none);
 * text copied into buf. */
int acd_dev_last_error(char * buf, size_t len);
void acd_dev_record_error(int code, const char * what);
/* Abort the process with a message if code != 0 (mirrors the original
 * library's assertion behaviour on unrecoverable states). */
void acd_dev_check(int code, const char * what);

/* ---- sliced ELL operator (one of A, A') ---- */
#define ACD_SLICE 64u      /* rows per slice = one wavefront */

struct acd_sell {
        uint32_t nrows;        /* rows of the operator */
        uint32_t ncols;        /* columns of the operator */
        uint32_t nslices;      /* ceil(nrows / 64) */
        uint32_t max_len;      /* longest row */
        uint64_t nent;         /* stored entries, padding included */
        uint64_t * off;        /* device [nslices+1]: first entry of slice */
        uint32_t * col;        /* device [nent] */
        double * val;          /* device [nent] */
        float * val32;         /* device [nent] or NULL (mixed precision) */
        uint16_t * val16;      /* device [nent] bf16 bits, or NULL */
        uint32_t * colsgn;     /* device [nent] column | sign bits when every
                                  value is -1, 0 or +1 (value-free), or NULL */
        uint16_t * col16;      /* device [nent] columns as 16 bits when the
                                  operator has at most 65,536 columns, or
                                  NULL */
        uint32_t * rowlen;     /* device [nrows]: entries per row */
        int32_t tuned;         /* SpMV variant + 1 picked by the autotuner,
                                  0 = library default */
};

/* ---- preprocessing (dev_preprocess.hip) ---- */
/* Device-side matrix build.  Inputs are host triplets; outputs:
 *   - row/col permutation idx/ridx (device and host copies),
 *   - A and A' in sliced ELL (device),
 *   - the renumbered triplets sorted in Z-order, copied back to the host
 *     arrays rows/cols/vals (which hold the input on entry).
 * Returns 0 on success. */
struct acd_dev_matrix {
        uint32_t nrows, ncols;
        uint64_t nnz;
        uint32_t * row_idx, * row_ridx;   /* device [nrows] */
        uint32_t * col_idx, * col_ridx;   /* device [ncols] */
        struct acd_sell a, at;
};
int acd_dev_matrix_build(struct acd_dev_matrix * out,
                         uint32_t nrows, uint32_t ncols, uint64_t nnz,
                         uint32_t * rows, uint32_t * cols, double * vals,
                         int permute,
                         uint32_t * h_row_idx, uint32_t * h_row_ridx,
                         uint32_t * h_col_idx, uint32_t * h_col_ridx);
void acd_dev_matrix_free(struct acd_dev_matrix *);
/* FP32 copy of the stored values (rounded to nearest). */
int acd_dev_sell_make_fp32(struct acd_sell *);
/* BF16 copy of the stored values, made only when every value is exactly
 * representable (0/+-1 and small dyadic coefficients).  Returns 1 when the
 * copy exists, 0 when some value is not exact (no copy), -1 on error. */
int acd_dev_sell_make_bf16(struct acd_sell *);
/* Value-free copy of a 0/+-1 operator: each stored entry as one word, the
 * column in bits 0-29, bit 31 for a value of -1, bit 30 for 0 (padding).
 * Returns 1 when made, 0 when some value is not -1, 0 or +1 or some
 * column does not fit in 30 bits (no copy), -1 on error. */
int acd_dev_sell_make_sign(struct acd_sell *);
/* 16-bit copy of the column indices (col16), made when the operator has
 * at most 65,536 columns (padding entries repeat a column of their row, or
 * 0, so every stored index is below ncols).  Returns 1 when made, 0 when
 * the operator is too wide (no copy), -1 on error. */
int acd_dev_sell_make_col16(struct acd_sell *);

/* ---- vector kernels (dev_spmv.hip) ---- */
/* y = op x on device pointers (op = A' when transpose). */
int acd_dev_spmv(const struct acd_dev_matrix *, int transpose,
                 double * y, const double * x);
enum { ACD_DEV_SPMV_VARIANTS = 13 };
/* Times every SpMV variant (unroll x slices per wavefront x load kind) on
 * both operators of m, `reps` products each, and keeps the fastest per
 * operator; out[0], out[1] receive the choices for A and A'.  Variants
 * differ only in scheduling: every row is summed in the same order. */
int acd_dev_spmv_autotune(struct acd_dev_matrix * m, int reps, int * out);
int acd_dev_spmv2(const struct acd_dev_matrix *, int transpose,
                  double * y0, double * y1, const double * x0,
                  const double * x1);
/* dest[i] = src[p[i]] over n doubles. */
int acd_dev_gather(double * dest, const double * src, const uint32_t * p,
                   uint32_t n);
/* Time `reps` back-to-back launches of the FP64 SpMV kernel with HIP
 * events on the library stream; returns mean milliseconds per launch. */
int acd_dev_spmv_time(const struct acd_dev_matrix *, int transpose,
                      double * y, const double * x, int reps,
                      double * OUT_ms);
int acd_dev_spmv_time_variant(const struct acd_dev_matrix *, int transpose,
                              double * y, const double * x, int reps,
                              int variant, double * OUT_ms);
int acd_dev_spmv_set_default_variant(int variant);
/* Allocation-free product (no hipMalloc/hipFree): host x is staged in and
 * host y out through a fixed device array of 2 x ACD_NOALLOC_MAX doubles
 * that the library holds per device.  Returns 0, or -1 (error recorded)
 * when a length exceeds ACD_NOALLOC_MAX or does not match the operator.
 * Calls that overlap on one device share the array and race. */
#define ACD_NOALLOC_MAX 4194304u
int acd_dev_spmv_noalloc(const struct acd_dev_matrix *, int transpose,
                         double * y, size_t ny, const double * x, size_t nx);
/* Looks up the staging array's device address for the current device
 * (once per device); 0 on success. */
int acd_dev_scratch_setup(void);
/* Mean milliseconds per product over `reps` rounds of k products, product
 * i on matrix m[i] with d_y[i] <- op d_x[i] (k distinct operators and
 * vectors, so a working set larger than the device's last-level cache is
 * streamed from HBM), timed with HIP events on the library stream. */
int acd_dev_spmv_time_rotate(const struct acd_dev_matrix * const * m,
                             size_t k, int transpose, double * const * y,
                             const double * const * x, int reps,
                             double * OUT_ms);

/* ---- device-resident solver (dev_solver.hip) ---- */
/* Original-order vectors of a device problem that the host uploads. */
enum acd_vec {
        ACD_V_RHS = 0,        /* b (approx rhs)               [m] */
        ACD_V_WEIGHT,         /* row weights                  [m] */
        ACD_V_LINEAR,         /* approx linear term c'        [n] */
        ACD_V_LOWER,          /* x lower bounds               [n] */
        ACD_V_UPPER,          /* x upper bounds               [n] */
        ACD_V_ALM_LINEAR,     /* LP objective c               [n] */
        ACD_V_LAM_LO,         /* multiplier bounds            [m] */
        ACD_V_LAM_HI,         /*                              [m] */
        ACD_V_X,              /* x in/out                     [n] */
        ACD_V_LAMBDA,         /* multipliers in/out           [m] */
        ACD_V_COUNT
};

enum acd_mode { ACD_MODE_APPROX = 1, ACD_MODE_MPG = 2, ACD_MODE_ALM = 3,
                ACD_MODE_PCD = 4 };
/* Working-copy precision (same values as acd_gpu.h's enum acd_precision). */
enum { ACD_DEV_FP64 = 0, ACD_DEV_MIXED = 1, ACD_DEV_FP32 = 2 };

struct acd_solve_request {
        int mode;                 /* enum acd_mode */
        int log;                  /* emit log events */
        uint64_t niter;           /* inner (APPROX/MPG) or outer (ALM) cap */
        double max_pg, max_value, min_delta, offset;
        uint64_t period;
        /* ALM only */
        double stop_violation, stop_pg, precision_floor, inner_min_delta;
        int precision;            /* ACD_DEV_FP64 / _MIXED / _FP32 */
        double stop_kkt;          /* > 0: KKT stopping target (ALM) */
        /* PCD only: sampling probability, seed, and steps per epoch
         * (niter counts epochs) */
        double pcd_p;
        uint64_t pcd_seed;
        uint32_t pcd_steps;
};

struct acd_solve_result {
        int ret;                  /* solver return code */
        int error;                /* 0, or an ACD_ERR_* code */
        double diag[5];
        uint64_t inner_iterations, outer_iterations, launches;
        double device_ms;
        double kkt;               /* last device KKT residual or NaN */
};

enum acd_err {
        ACD_ERR_NONE = 0,
        ACD_ERR_UNBOUNDED = 1,    /* max |zp| not finite (step_project.inc:90) */
        ACD_ERR_ESTIMATE = 2,     /* expected improvement > 0 (approx.c:376) */
        ACD_ERR_BARRIER = 3,      /* grid barrier timed out */
        ACD_ERR_LAUNCH = 4
};

/* Log events, replayed by the host in order. */
enum acd_event_type {
        ACD_EV_APPROX_BEGIN = 1,  /* new inner solve: reset restart flag */
        ACD_EV_RESTART,           /* a[0] = iteration */
        ACD_EV_NEWLINE,           /* the stdout newline after restarts */
        ACD_EV_PRINT,             /* k, value, ng, pg, step, delta */
        ACD_EV_MPG,               /* k, pg, value */
        ACD_EV_ALM_ITER,          /* k, max_viol, |r|, pg, value, c'x */
        ACD_EV_ALM_TAIL           /* stalled, mean weight */
};

struct acd_event {
        int32_t type, pad;
        uint64_t k;
        double a[6];
};

typedef void (*acd_event_sink)(void * ctx, const struct acd_event * ev);

struct acd_dev_lp;
struct acd_dev_lp * acd_dev_lp_create(const struct acd_dev_matrix * orig,
                                      const struct acd_dev_matrix * perm);
void acd_dev_lp_free(struct acd_dev_lp *);
double * acd_dev_lp_vec(struct acd_dev_lp *, int which);
/* Permute b, w, c', l, u into the working copy and recompute the step
 * sizes there (approx/approx.c:188-216). */
int acd_dev_approx_update(struct acd_dev_lp *);
int acd_dev_solve(struct acd_dev_lp *, const struct acd_solve_request *,
                  struct acd_solve_result *, acd_event_sink, void * ctx);
/* Per-op launch counts and microseconds (step + control) of the last
 * solve, indexed by op code. */
void acd_dev_last_profile(uint64_t * counts, double * us, int n);
/* Per op of the last solve: microseconds from the step kernel's start to
 * the control kernel's start (the step kernel's duration plus the control
 * launch latency), and algorithmic bytes moved. */
void acd_dev_last_profile_ex(double * step_us, double * bytes, int n);
/* Control kernel time split over the last solve: microseconds in the
 * loads, the combine and the decision, and the number of calls. */
void acd_dev_last_ctl_phases(double * us4);

/* ---- batched solves (dev_solver.hip) ---- */
/* Runs request q (ALM mode) on every LP, one workgroup per LP, in repeated
 * launches of 32 inner iterations until all are done.  res[i] receives
 * each LP's result; done_ms[i] (optional) the device time from the batch
 * start to that LP's completion. */
#define ACD_DEV_BATCH_CLASSES 8
struct acd_batch_stats {
        double device_ms;          /* HIP events around all launches */
        double kernel_ms;          /* HIP events around each round of
                                      launches, summed (host gaps between
                                      rounds excluded) */
        uint64_t launches;
        uint64_t inner_iterations; /* summed over the LPs */
        double alg_bytes;          /* algorithmic bytes, summed */
        /* LP-milliseconds (summed over the LPs' workgroups) spent in ops
         * (passes and their reductions) and in the control continuation */
        double lp_ops_ms, lp_ctl_ms;
        /* workgroup size / 64 used for each size class (0: no LP) */
        int waves_per_lp[ACD_DEV_BATCH_CLASSES];
};
/* sink (optional, with q->log set and count == 1): log events of the LP,
 * replayed between launches and at the end, in order. */
int acd_dev_solve_batch(struct acd_dev_lp * const * lps, size_t count,
                        const struct acd_solve_request * q,
                        struct acd_solve_result * res, double * done_ms,
                        struct acd_batch_stats * stats, acd_event_sink sink,
                        void * ctx);

/* ---- continuous batching (dev_solver.hip) ---- */
/* A pool of `cap` slots for the batched kernel: LPs are admitted into free
 * slots between rounds; a round launches the batched kernel `launches`
 * times over the listed slots (each advancing by up to `budget` inner
 * iterations) and returns the per-slot done flags (pinned host array). */
struct acd_dev_pool;
struct acd_dev_pool * acd_dev_pool_create(uint32_t cap, int precision);
void acd_dev_pool_free(struct acd_dev_pool *);
int acd_dev_pool_admit(struct acd_dev_pool *, uint32_t slot,
                       struct acd_dev_lp *, const struct acd_solve_request *);
int acd_dev_pool_round(struct acd_dev_pool *, const uint32_t * slots,
                       uint32_t n, uint32_t launches, uint64_t budget,
                       uint32_t ** flags);
int acd_dev_pool_result(struct acd_dev_pool *, uint32_t slot,
                        struct acd_solve_result *);
/* Non-blocking single-LP solve on the LP's own stream (no log): begin,
 * then poll until it returns 1 (done, result filled); 0 = still running,
 * -1 = error.  Uploads/downloads of the LP's vectors are the caller's. */
int acd_dev_lp_async_begin(struct acd_dev_lp *,
                           const struct acd_solve_request *);
int acd_dev_lp_async_poll(struct acd_dev_lp *, struct acd_solve_result *);
/* Block until the chunk in flight has finished. */
int acd_dev_lp_async_wait(struct acd_dev_lp *);

/* ---- randomised coordinate descent (dev_cd.hip, approx/cd.c) ---- */
/* Works on the caller-order vectors of an acd_dev_lp (x, rhs, weight,
 * linear, lower, upper) and the unpermuted matrix. */
struct acd_dev_cd;
struct acd_dev_cd * acd_dev_cd_create(const struct acd_dev_matrix *,
                                      struct acd_dev_lp *);
void acd_dev_cd_free(struct acd_dev_cd *);
/* Column set-up from the current weights; copies the n sampling
 * probabilities to h_prob. */
int acd_dev_cd_setup(struct acd_dev_cd *, double * h_prob);
int acd_dev_cd_clamp(struct acd_dev_cd *);
/* g, violation and the objective value of the current x. */
int acd_dev_cd_gradient(struct acd_dev_cd *, double * value);
/* ||projected gradient||, ||g||, ||violation|| of the last gradient. */
int acd_dev_cd_norms(struct acd_dev_cd *, double * pg, double * ng,
                     double * nviol);
/* The coordinate steps for the sampled columns, in order. */
int acd_dev_cd_steps(struct acd_dev_cd *, const uint32_t * h_cols,
                     size_t ncols, double scale);

#ifdef __cplusplus
}
#endif
#endif
