/* acd_approx.h -- private layout of approx_t / alm_t and shared helpers of
 * the host layer. */
#ifndef ACD_HOST_APPROX_PRIVATE_H
#define ACD_HOST_APPROX_PRIVATE_H
#include <stdio.h>
#include "acd_host.h"

/* approx/approx.c:10-24 of the reference, re-laid for the GPU. */
struct approx {
        size_t nrhs, nvars;
        sparse_matrix_t * matrix;      /* borrowed */
        sparse_matrix_t * permuted;    /* owned working copy */
        double * rhs, * weight, * linear, * lower, * upper;  /* host */
        struct acd_dev_lp * dev;
};

/* alm.c:8-15 */
struct alm {
        size_t nrhs, nvars;
        sparse_matrix_t * matrix;
        double * linear;
        approx_t * approx;
        double * lambda_lower, * lambda_upper;
};

struct acd_log_ctx { FILE * log; int restart; };
/* Caller vectors of an ALM object to and from its device LP (aborts on
 * failure, like the solve entry points). */
void acd_alm_upload(alm_t * alm, const double * x, const double * lambda);
void acd_alm_download(alm_t * alm, double * x, double * lambda, size_t niter);
void acd_log_sink(void * ctx, const struct acd_event * e);
int acd_approx_upload(approx_t * a);
void acd_record_stats(const struct acd_solve_result * r);
void acd_check_result(const struct acd_solve_result * r, int rc,
                      const char * who);
#endif
