/* approx.h -- box-constrained weighted least squares,
 *
 *     min_x  sum_i weight_i/2 (A_i x - rhs_i)^2 + linear' x
 *     s.t.   lower <= x <= upper,
 *
 * solved on the GPU.  Drop-in for approx/approx.h of pkhuong/acd-alm: same
 * prototypes, same defaults, same return codes and log lines.  The object
 * keeps host copies of its vectors (the accessors return them, mutable, as
 * the original does) and device copies used by the solvers; approx_update()
 * is the point where host edits reach the device.
 */
#ifndef ACD_APPROX_H
#define ACD_APPROX_H
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include "../spmv/spmv.h"
#include "../thread_pool/thread_pool.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct approx approx_t;

/* Copies every vector; borrows constraints, which must outlive the
 * object.  weight NULL -> all ones; linear NULL -> zeros; lower/upper NULL
 * -> -HUGE_VAL / HUGE_VAL.  nrhs and nvars must match the matrix.  Builds
 * the row/column-permuted working copy of the matrix on the device and
 * computes the per-coordinate step sizes.  (approx/approx.h:41-45) */
approx_t * approx_make(sparse_matrix_t * constraints,
                       size_t nrhs, const double * rhs, const double * weight,
                       size_t nvars,
                       const double * linear,
                       const double * lower, const double * upper);

/* Frees the object and its permuted matrix copy, not constraints.  NULL is
 * ignored.  Returns 0.  (approx/approx.h:54) */
int approx_free(approx_t *);

/* Accessors; the vectors are mutable host arrays.  (approx/approx.h:58-65) */
sparse_matrix_t * approx_matrix(approx_t *);
size_t approx_nrhs(approx_t *);
double * approx_rhs(approx_t *);
double * approx_weight(approx_t *);
size_t approx_nvars(approx_t *);
double * approx_linear(approx_t *);
double * approx_lower(approx_t *);
double * approx_upper(approx_t *);

/* Publish host-side edits of rhs/weight/linear/bounds to the device,
 * permute them into the working copy and recompute step sizes.  Returns 0.
 * (approx/approx.h:72) */
int approx_update(approx_t *);

/* Accelerated proximal coordinate-gradient (APPROX) with adaptive restart
 * and adaptive long steps, from x (n doubles, updated in place with the
 * best point).  Stops after niter iterations, when the projected-gradient
 * 2-norm drops below max_pg (returns 2), the objective below max_value
 * (returns 1), or when the point moved by less than min_delta (relative)
 * over the last 100 iterations (returns 3); 0 when niter ran out.
 * OUT_diagnosis (NULL or double[5]): objective, gradient norm, projected
 * gradient norm, relative move since the last multiple of 100, iteration
 * count.  offset is added to every reported objective value.  Logs to log
 * (NULL: silent) on the first and last iteration and every period
 * iterations.  Every iteration runs on the device; only convergence
 * scalars come back.  pool is accepted and ignored.
 * (approx/approx.h:113-117) */
int approx_solve(double * x, size_t n, approx_t * approx, size_t niter,
                 double max_pg, double max_value, double min_delta,
                 FILE * log, size_t period,
                 double * OUT_diagnosis /* NULL or double[5] */,
                 double offset, thread_pool_t * pool);
/* Spectral projected gradient (Barzilai-Borwein step).  Always returns 2,
 * as the original does.  (approx/approx.h:118-122) */
int mpg_solve(double * x, size_t n, approx_t * approx, size_t niter,
              double max_pg, double max_value, double min_delta,
              FILE * log, size_t period,
              double * OUT_diagnosis /* NULL or double[5] */,
              double offset, thread_pool_t * pool);
/* Randomised coordinate descent (the original's default ALM inner solver).
 * Returns 0, 1 or 2 like approx_solve.  (approx/approx.h:123-127) */
int cd_solve(double * x, size_t n, approx_t * approx, size_t niter,
             double max_pg, double max_value, double min_delta,
             FILE * log, size_t period,
             double * OUT_diagnosis /* NULL or double[5] */,
             double offset, thread_pool_t * pool);

#ifdef __cplusplus
}
#endif
#endif
