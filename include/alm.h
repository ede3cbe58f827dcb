/* alm.h -- linear programs by augmented Lagrangian, on the GPU:
 *
 *     min_x  linear' x   s.t.   A x = rhs,   lower <= x <= upper.
 *
 * Drop-in for alm.h of pkhuong/acd-alm (same prototypes, defaults, return
 * codes and log format).  Every outer iteration -- penalised objective,
 * inner APPROX solve, residual, multiplier/weight/precision updates and
 * the stopping test -- runs on the device; the host only reads back the
 * scalars it logs.
 */
#ifndef ACD_ALM_H
#define ACD_ALM_H
#include <stddef.h>
#include <stdio.h>
#include "spmv/spmv.h"
#include "approx/approx.h"
#include "thread_pool/thread_pool.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alm alm_t;

/* Copies every vector, borrows constraints (nrhs x nvars; must outlive the
 * object).  linear NULL -> zeros; lower/upper NULL -> -/+HUGE_VAL;
 * lambda_lower/lambda_upper (bounds on the multipliers, e.g. sign
 * constraints for rows that are inequalities with explicit slacks) NULL ->
 * -/+HUGE_VAL.  (alm.h:43-47) */
alm_t * alm_make(sparse_matrix_t * constraints,
                 size_t nrhs, const double * rhs,
                 size_t nvars, const double * linear,
                 const double * lower, const double * upper,
                 const double * lambda_lower, const double * lambda_upper);
/* Frees the object, not the matrix.  Returns 0.  (alm.h:53) */
int alm_free(alm_t *);

/* Accessors: mutable host vectors (edits are picked up by the next
 * alm_solve).  (alm.h:59-67) */
sparse_matrix_t * alm_matrix(alm_t *);
size_t alm_nrhs(alm_t *);
double * alm_rhs(alm_t *);
size_t alm_nvars(alm_t *);
double * alm_linear(alm_t *);
double * alm_lower(alm_t *);
double * alm_upper(alm_t *);
double * alm_lambda_lower(alm_t *);
double * alm_lambda_upper(alm_t *);

/* Up to niter outer iterations from the primal x (nvars doubles) and the
 * multipliers lambda (nconstraints doubles), both updated in place.  Stops
 * when the max constraint violation is below 1e-5 and the inner solver
 * either reached a projected-gradient norm below 1e-5 or stalled.
 * Returns 0 on that stop, 1 when niter ran out.  OUT_diagnosis (NULL or
 * double[3]): primal objective, max violation, projected-gradient norm.
 * Log line per outer iteration:
 *   k: max_violation 2norm_violation pg lagrangian_bound primal_objective
 * followed by " -stalled-" or the mean row weight.  pool is accepted and
 * ignored.  (alm.h:99-103) */
int alm_solve(alm_t *, size_t niter,
              double * x, size_t nvars,
              double * lambda, size_t nconstraints,
              FILE * log, double * OUT_diagnosis,
              thread_pool_t * pool);

/* Text instance: a matrix as for sparse_matrix_read, then rhs (nrhs),
 * linear (nvars), lower (nvars, "-inf" allowed), upper (nvars, "inf"),
 * lambda_lower (nrhs), lambda_upper (nrhs).  Returns NULL on malformed
 * input.  (alm.h:117) */
alm_t * alm_read(FILE * stream);

/* Reads n whitespace-separated doubles; exported by the original library
 * as well (alm.c:324). */
void read_doubles(FILE * stream, double * out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
