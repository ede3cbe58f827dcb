/* ora_internal.h -- private layout shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H
#include "oracle.h"

/* approx/approx.c:10-24 */
struct ora_approx {
        size_t nrhs, nvars;
        ora_matrix_t * matrix;
        double * rhs, * weight, * linear, * lower, * upper;
        uint32_t * beta;
        double * v, * inv_v;
        struct ora_approx * permuted;
};

/* Called at every checkpoint of ora_approx_solve (every 100 iterations,
 * after x, its violation w (Ax - b) and its gradient are formed, all in
 * the working copy's order) unless NULL; a nonzero return ends the solve
 * as converged (reason 2).  ora_alm_solve sets it for its KKT target, for
 * the calling thread only. */
typedef int (*ora_ckpt_hook)(void * ctx, const struct ora_approx * a,
                             const double * x, const double * viol,
                             const double * g);
void ora_approx_set_ckpt_hook(ora_ckpt_hook fn, void * ctx);

#endif
