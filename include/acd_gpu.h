/* acd_gpu.h -- additions of the MI355X library that have no counterpart
 * in pkhuong/acd-alm.  Everything declared in alm.h, approx/approx.h and
 * spmv/spmv.h keeps its original meaning; nothing here is required to use
 * them.
 */
#ifndef ACD_GPU_H
#define ACD_GPU_H
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include "spmv/spmv.h"
#include "approx/approx.h"
#include "alm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device, errors ---- */
/* Choose the HIP device before sparse_matrix_init().  Returns 0, or -1
 * once the library is already bound to another device. */
int acd_set_device(int device);
/* First HIP error recorded since the previous call (0 if none); its text
 * is copied into buf (may be NULL). */
int acd_last_error(char * buf, size_t len);
/* Wait for all work queued by the library. */
int acd_synchronize(void);
/* Free and total bytes of the library's device (hipMemGetInfo). */
int acd_device_memory_info(size_t * OUT_free, size_t * OUT_total);

/* ---- device-resident data ---- */
void * acd_device_alloc(size_t bytes);          /* zero-filled */
void acd_device_free(void * p);
int acd_copy_to_device(void * dst, const void * src, size_t bytes);
int acd_copy_to_host(void * dst, const void * src, size_t bytes);
/* Device-to-device copy on the library stream (a plain copy kernel of
 * known size, used to calibrate hardware byte counters). */
int acd_copy_device(void * dst, const void * src, size_t bytes);
/* d_y <- op(A) d_x on device pointers, enqueued on the library stream. */
int acd_spmv_device(const sparse_matrix_t *, int transpose,
                    double * d_y, const double * d_x);
/* Mean milliseconds per launch of the device product over reps
 * back-to-back launches, timed with HIP events on the library stream. */
int acd_spmv_device_time(const sparse_matrix_t *, int transpose,
                         double * d_y, const double * d_x, int reps,
                         double * OUT_ms_per_launch);
/* SpMV kernel variants (unroll depth x slices per wavefront x load kind x
 * XCD-contiguous row tiles; all give
 * bit-identical results).  Times variant `variant` like
 * acd_spmv_device_time; acd_spmv_set_variant makes one the default for
 * subsequent products. */
int acd_spmv_variant_time(const sparse_matrix_t *, int transpose,
                          double * d_y, const double * d_x, int reps,
                          int variant, double * OUT_ms_per_launch);
int acd_spmv_set_variant(int variant);
enum { ACD_SPMV_VARIANTS = 13 };
/* Autotuner: times every variant on both operators of the matrix (`reps`
 * products each) and keeps the fastest per operator for
 * sparse_matrix_multiply and acd_spmv_device.  Variants only schedule
 * differently (unroll depth, rows per wavefront, load kind): every row is
 * summed in the same order, so results do not change.  OUT_variants
 * (optional, int[2]) receives the choices for A and A'. */
int acd_spmv_autotune(sparse_matrix_t *, int reps, int * OUT_variants);

/* sparse_matrix_multiply without any device allocation or free: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  OUT_y[ny] <- op(a) x[nx] (op = A' when
 * transpose), host arrays, staged through a fixed device array of
 * 2 x 4,194,304 doubles (64 MiB) that the library holds on each device it
 * runs on until the process ends (sparse_matrix_multiply instead
 * allocates staging buffers per matrix on first use).  Returns 0; or -1,
 * with acd_last_error() describing it, when nx or ny exceeds 4,194,304,
 * a length does not match the operator, or an argument is NULL.  The
 * result is that of sparse_matrix_multiply bit for bit.  Calls that
 * overlap on one device (several host threads) share the array and
 * race; serialise them. */
int acd_sparse_matrix_multiply_noalloc(double * OUT_y, size_t ny,
                                       const sparse_matrix_t * a,
                                       const double * x, size_t nx,
                                       int transpose);

/* Mean milliseconds per product over `reps` rounds of k products: in
 * each round, d_y[i] <- op(mats[i]) d_x[i] for i = 0..k-1, timed with HIP
 * events on the library stream.  With k distinct matrices whose operators
 * together exceed the device's last-level cache (256 MB on MI355X), every
 * product streams its matrix from HBM.  k <= 64. */
int acd_spmv_rotate_time(const sparse_matrix_t * const * mats, size_t k,
                         int transpose, double * const * d_y,
                         const double * const * d_x, int reps,
                         double * OUT_ms_per_launch);

/* Stored entries of the operator (padding to the slice length included). */
uint64_t acd_matrix_stored_entries(const sparse_matrix_t *, int transpose);

/* ---- solver options ---- */
enum acd_precision {
        ACD_FP64 = 0,    /* FP64 everywhere: parity form */
        ACD_MIXED = 1,   /* matrix values in BF16 when every value is exact
                            there, else FP32; rhs, weights, bounds, costs and
                            step sizes of the working copy in FP32; iterates,
                            row accumulation, duals, norms and scalars FP64 */
        ACD_FP32 = 2     /* ACD_MIXED plus FP32 iterates (x, z, gradients,
                            violation vectors): fewest bytes, but the FP32
                            rounding of the iterates limits the reachable
                            KKT residual (about 1e-5..1e-3 relative) */
};

/* Which device solver runs a single LP: all workgroups of the GPU
 * (step + control kernels), or one workgroup (the batched kernel with a
 * batch of one, no log).  AUTO picks the workgroup solver for small LPs
 * solved without a log (rows + columns <= 4,000), where it is 1.3-2x
 * faster per iteration; both follow the same algorithm and differ only in
 * the order of their reductions. */
enum acd_engine {
        ACD_ENGINE_AUTO = 0,
        ACD_ENGINE_GRID = 1,
        ACD_ENGINE_WORKGROUP = 2
};

struct acd_alm_options {
        double stop_violation;   /* 1e-5  (alm.c:306) */
        double stop_pg;          /* 1e-5  (alm.c:308) */
        double precision_floor;  /* 1e-7  (alm.c:244) */
        double min_delta;        /* 1e-11 (alm.c:256) */
        int precision;           /* enum acd_precision */
        int engine;              /* ACD_ENGINE_AUTO (default), _GRID, _WORKGROUP */
        double stop_kkt;         /* 0: the reference's rule above.  > 0:
                                    instead stop once the relative KKT
                                    residual max(r_p, r_d, r_gap) of
                                    (x, lambda), evaluated on the device after
                                    each outer iteration, is <= stop_kkt */
};
void acd_alm_default_options(struct acd_alm_options *);
/* alm_solve with explicit stopping constants and precision. */
int alm_solve_ex(alm_t *, size_t niter, double * x, size_t nvars,
                 double * lambda, size_t nconstraints, FILE * log,
                 double * OUT_diagnosis, const struct acd_alm_options *);

/* ---- parallel coordinate descent ---- */
/* cd_solve's problem (minimise c'x + 1/2 sum_i w_i (A x - b)_i^2 + offset
 * over l <= x <= u, from x) by parallel randomised coordinate descent on
 * the device; no counterpart in the reference, whose cd_solve takes one
 * coordinate at a time (and stays available, bit for bit).  Each step
 * moves a random sample of the coordinates (each with probability
 * `sample`, 0 < sample <= 1, drawn from `seed`) by the step of the
 * expected separable overapproximation, x_j <- clamp(x_j - g_j / v'_j),
 * v'_j = sum_i w_i a_ij^2 (1 + sample (omega_i - 1)), omega_i the nonzeros
 * of row i.  niter counts epochs of
 * ceil(1 / sample) steps.  Stops like cd_solve (value < max_value: 1,
 * projected-gradient norm < max_pg: 2, else 0 after niter epochs);
 * OUT_diagnosis (double[5], may be NULL) as cd_solve's: value, |g|,
 * projected-gradient norm, 0, epochs.  Returns -1 for a sample outside
 * (0, 1].  The result depends on the seed, not on the device. */
int acd_cd_solve_parallel(double * x, size_t n, approx_t * approx,
                          size_t niter, double max_pg, double max_value,
                          double offset, double sample, uint64_t seed,
                          double * OUT_diagnosis);

/* ---- batched solves and the request loop ---- */
/* A batched or served LP whose solve fails gets the negated code as its
 * return value (the single-LP entry points abort instead, as the
 * reference's asserts do). */
enum acd_solve_error {
        ACD_SOLVE_UNBOUNDED = 1,   /* an iterate is not finite */
        ACD_SOLVE_ESTIMATE = 2,    /* a long step's model estimate > 0 */
        ACD_SOLVE_NO_MEMORY = 3    /* a served LP given as host data could
                                      not be built on the device */
};
#define ACD_BATCH_CLASSES 8
struct acd_batch_info {
        double device_ms;            /* device time of the whole batch */
        uint64_t launches;           /* batch kernel launches */
        uint64_t inner_iterations;   /* summed over the LPs */
        double alg_bytes;            /* algorithmic bytes moved, summed */
        /* Wavefronts per LP workgroup (1, 2 or 4) used for each size
         * class: class k holds the LPs whose wider operator (A or A') has
         * at most 2^k slices of 64 rows, the last class the rest; 0 where
         * the batch had no LP of the class.  Batches of 192 LPs or more
         * tune it on their own rounds (each candidate assignment runs one
         * round, the fastest is kept; again when half of the LPs have
         * finished) and report the first choice; smaller batches reuse
         * the last one; ACD_BATCH_WAVES=1|2|4 forces one size.  Results do
         * not depend on it. */
        int waves_per_lp[ACD_BATCH_CLASSES];
        /* Device time of the batched-kernel launches: HIP events around
         * each round of launches on the library stream, summed, so the
         * host's work between rounds is excluded (device_ms includes it). */
        double kernel_ms;
        /* Time the LPs' workgroups spent, summed over the LPs (LP-ms), in
         * the data-parallel passes with their reductions and in the
         * one-thread control step that takes the reference's decisions
         * between passes. */
        double lp_ops_ms, lp_ctl_ms;
};
/* alm_solve_ex on `count` independent ALM objects at once: one workgroup
 * per LP, every LP on the device together, repeated launches of 32 inner
 * iterations until all are done (no extension of the reference; a loop
 * of alm_solve_ex calls gives the same results up to reduction order).
 * x[i], lambda[i] (in/out) and OUT_diagnosis[i] (double[3], may be NULL)
 * are per LP; OUT_ret[i] receives alm_solve's return code, or
 * -enum acd_solve_error.  OUT_done_ms
 * and OUT_inner (optional, count entries) receive each LP's completion
 * time from the batch start and its inner iteration count.  Each LP's result does not depend on the other LPs in the
 * batch. */
int alm_solve_batch(size_t count, alm_t * const * alms, size_t niter,
                    double * const * x, double * const * lambda,
                    double * const * OUT_diagnosis, int * OUT_ret,
                    const struct acd_alm_options * options,
                    double * OUT_done_ms, uint64_t * OUT_inner,
                    struct acd_batch_info * OUT_info);

/* ---- MPS input ---- */
/* An LP read from (free or blank-free fixed) MPS: sections NAME, ROWS,
 * COLUMNS (integrality markers ignored), RHS, RANGES, BOUNDS, OBJSENSE,
 * ENDATA.  It is put in the ALM form min c'x s.t. Ax = b, l <= x <= u:
 * each L, G or ranged row whose interval [lo, hi] is not a single point
 * gets a slack column, after the structural columns: a'x - s = lo with
 * 0 <= s <= hi - lo when lo is finite, else a'x + s = hi with s >= 0.
 * OBJSENSE MAX negates the costs.  Arrays are malloc'ed; the COO holds
 * the COLUMNS entries in input order, then one entry per slack. */
struct acd_mps_lp {
        size_t nrows, ncols, nnz;       /* ncols includes the slacks */
        size_t nstructural;             /* columns named in COLUMNS */
        uint32_t * rows, * cols;
        double * values;
        double * rhs, * linear, * lower, * upper;
        double objective_offset;        /* constant term of c'x (minus
                                           the objective row's RHS) */
        int maximize;                   /* OBJSENSE MAX: costs negated */
        char error[128];                /* why parsing failed */
};
/* 0, or -1 with OUT_lp->error set (nothing to free then). */
int acd_mps_parse(FILE * stream, struct acd_mps_lp * OUT_lp);
void acd_mps_free(struct acd_mps_lp * lp);
/* acd_mps_parse, sparse_matrix_make (unpermuted) and alm_make with free
 * multipliers; like alm_read, the ALM object references the matrix
 * (release it with sparse_matrix_free(alm_matrix(alm)) taken before
 * alm_free).  OUT_info (optional) receives the sizes, offset and error
 * text; its arrays are already freed.  NULL on error. */
alm_t * acd_alm_read_mps(FILE * stream, struct acd_mps_lp * OUT_info);

/* ---- request loop (continuous batching) ---- */
typedef struct acd_server acd_server_t;
/* A request loop on the current device.  hbm_budget bounds the summed
 * device working set of the requests running at once (a request larger
 * than the budget runs alone); `slots` is the number of small LPs that run
 * together in the batched kernel (0: 2048); LPs with at least `large_rows`
 * rows run the multi-workgroup solver on their own stream instead (up to
 * 4 at once).  Options as for alm_solve_ex (niter = 1000). */
acd_server_t * acd_server_make(size_t hbm_budget, size_t slots,
                               size_t large_rows,
                               const struct acd_alm_options * options);
void acd_server_free(acd_server_t *);
/* Queue one solve of `alm` from x / lambda (caller-owned, in/out, like
 * alm_solve; they and OUT_diagnosis (double[3] or NULL) and OUT_ret must
 * stay valid until the request completes).  The ALM object must not be
 * used elsewhere meanwhile. */
int acd_server_submit(acd_server_t *, alm_t * alm, double * x,
                      double * lambda, double * OUT_diagnosis, int * OUT_ret,
                      uint64_t tag);
/* One round: admit queued requests that fit, advance running ones, retire
 * finished ones.  Writes up to `cap` completed tags (cap >= slots + 4
 * never drops one) and their count to *OUT_ndone. */
int acd_server_step(acd_server_t *, uint64_t * OUT_done, size_t cap,
                    size_t * OUT_ndone);
/* An LP as host arrays, in alm_make's conventions: A is nrows x ncols
 * with nnz (row, column, value) triplets, no repeated pair; linear, lower,
 * upper, lambda_lower and lambda_upper may be NULL (0, -inf, +inf, -inf,
 * +inf).  The arrays must stay valid until the request completes. */
struct acd_lp_data {
        size_t nrows, ncols, nnz;
        const uint32_t * rows, * cols;
        const double * values;
        const double * rhs, * linear, * lower, * upper;
        const double * lambda_lower, * lambda_upper;
};
/* Queue one solve of an LP given as host data.  Nothing is allocated on
 * the device until the request is admitted: then its matrix and ALM object
 * are built there (sparse_matrix_make, unpermuted, and alm_make), and they
 * are freed when it completes, so the device holds only the admitted
 * requests (their estimated footprints summed stay within the budget,
 * except that one request larger than the budget runs alone).  A request
 * whose LP cannot be built completes with -ACD_SOLVE_NO_MEMORY.  The
 * latency of a request therefore includes its device preprocessing. */
int acd_server_submit_lp(acd_server_t *, const struct acd_lp_data * lp,
                         double * x, double * lambda, double * OUT_diagnosis,
                         int * OUT_ret, uint64_t tag);
/* The footprint estimate admission uses for an LP of m rows, n columns
 * and nnz entries: an upper estimate of its device bytes, in whole 2 MiB
 * blocks of the device allocator plus one. */
size_t acd_server_footprint(size_t m, size_t n, size_t nnz);
/* Sum of the footprint estimates of the requests admitted now. */
size_t acd_server_bytes_admitted(const acd_server_t *);
/* Requests queued or running. */
size_t acd_server_pending(const acd_server_t *);

/* Restore the row weights (1) and penalised linear term (0) that alm_make
 * starts from, so the next alm_solve repeats a fresh solve; the original
 * keeps them across alm_solve calls. */
int acd_alm_reset(alm_t *);

/* Counters of the most recent alm_solve / approx_solve / mpg_solve. */
struct acd_solve_stats {
        uint64_t inner_iterations;   /* APPROX / SPG iterations */
        uint64_t outer_iterations;   /* ALM iterations */
        uint64_t launches;           /* solver kernel launches */
        double device_ms;            /* time inside solver launches */
        double kkt;                  /* last device KKT residual (stop_kkt
                                        > 0), else NaN */
};
void acd_last_solve_stats(struct acd_solve_stats *);
/* Time split of the last solve by device pass: for pass p < n, the number
 * of launches and the microseconds from its issue to the next decision.
 * Pass names are listed by acd_pass_name(). */
void acd_last_solve_profile(uint64_t * counts, double * us, int n);
/* For pass p < n of the last solve: microseconds from the start of its
 * step kernels to the start of the following control kernels (kernel
 * duration plus one launch latency), and the algorithmic bytes those
 * launches moved (each stored nonzero, slice offset and vector element
 * read or written once). */
void acd_last_solve_bytes(double * step_us, double * bytes, int n);
const char * acd_pass_name(int p);

#ifdef __cplusplus
}
#endif
#endif
