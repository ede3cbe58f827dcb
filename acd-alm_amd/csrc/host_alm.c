/* host_alm.c -- the alm.h entry points.
 *
 * Mirrors alm.c:8-364 of pkhuong/acd-alm.  The host object keeps the
 * caller-visible vectors; alm_solve uploads them, runs the whole outer
 * loop on the device (penalised objective, APPROX inner solve, residual,
 * multiplier / weight / precision updates, stopping test), and brings back
 * x, lambda and the updated row weights.  The inner solver is APPROX, as
 * alm.h:82-89 documents (the original's alm.c:255 calls cd_solve).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_approx.h"

static double * copy_default(const double * x, size_t n, double missing)
{
        double * out = malloc(sizeof(double) * (n ? n : 1));
        if (out == NULL) return NULL;
        for (size_t i = 0; i < n; i++) out[i] = x ? x[i] : missing;
        return out;
}

/* alm.c:59-83 */
alm_t * alm_make(sparse_matrix_t * constraints, size_t nrhs,
                 const double * rhs, size_t nvars, const double * linear,
                 const double * lower, const double * upper,
                 const double * lambda_lower, const double * lambda_upper)
{
        alm_t * alm = calloc(1, sizeof(*alm));
        if (alm == NULL) return NULL;
        alm->nrhs = nrhs;
        alm->nvars = nvars;
        alm->matrix = constraints;
        alm->linear = copy_default(linear, nvars, 0);
        alm->approx = approx_make(constraints, nrhs, rhs, NULL, nvars, NULL,
                                  lower, upper);
        alm->lambda_lower = copy_default(lambda_lower, nrhs, -HUGE_VAL);
        alm->lambda_upper = copy_default(lambda_upper, nrhs, HUGE_VAL);
        if (!alm->linear || !alm->approx || !alm->lambda_lower
            || !alm->lambda_upper) {
                alm_free(alm);
                return NULL;
        }
        return alm;
}

/* alm.c:85-96 */
int alm_free(alm_t * alm)
{
        if (alm == NULL) return 0;
        free(alm->linear);
        free(alm->lambda_lower);
        free(alm->lambda_upper);
        approx_free(alm->approx);
        memset(alm, 0, sizeof(*alm));
        free(alm);
        return 0;
}

/* alm.c:23-43 */
sparse_matrix_t * alm_matrix(alm_t * alm) { return alm->matrix; }
size_t alm_nrhs(alm_t * alm) { return alm->nrhs; }
size_t alm_nvars(alm_t * alm) { return alm->nvars; }
double * alm_linear(alm_t * alm) { return alm->linear; }
double * alm_lambda_lower(alm_t * alm) { return alm->lambda_lower; }
double * alm_lambda_upper(alm_t * alm) { return alm->lambda_upper; }
double * alm_rhs(alm_t * alm) { return approx_rhs(alm->approx); }
double * alm_lower(alm_t * alm) { return approx_lower(alm->approx); }
double * alm_upper(alm_t * alm) { return approx_upper(alm->approx); }

int acd_alm_reset(alm_t * alm)
{
        approx_t * a = alm->approx;
        for (size_t i = 0; i < a->nrhs; i++) a->weight[i] = 1;
        for (size_t i = 0; i < a->nvars; i++) a->linear[i] = 0;
        return 0;
}

void acd_alm_default_options(struct acd_alm_options * o)
{
        o->stop_violation = 1e-5;
        o->stop_pg = 1e-5;
        o->precision_floor = 1e-7;
        o->min_delta = 1e-11;
        o->precision = ACD_FP64;
        o->engine = ACD_ENGINE_AUTO;
        o->stop_kkt = 0;
}

int alm_solve_ex(alm_t * alm, size_t niter, double * x, size_t nvars,
                 double * lambda, size_t nconstraints, FILE * log,
                 double * OUT_diagnosis, const struct acd_alm_options * opt)
{
        struct acd_alm_options o;
        if (opt) o = *opt;
        else acd_alm_default_options(&o);
        ACD_REQUIRE(nvars == alm->nvars, "alm_solve: nvars mismatch");
        ACD_REQUIRE(nconstraints == alm->nrhs, "alm_solve: nconstraints mismatch");
        approx_t * a = alm->approx;
        struct acd_dev_lp * dev = a->dev;
        size_t m = alm->nrhs, n = alm->nvars;
        if (acd_approx_upload(a)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_ALM_LINEAR), alm->linear, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAM_LO), alm->lambda_lower, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAM_HI), alm->lambda_upper, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_X), x, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAMBDA), lambda, m * 8))
                acd_fail("alm_solve: upload");
        struct acd_solve_request q;
        memset(&q, 0, sizeof q);
        q.mode = ACD_MODE_ALM;
        q.log = log != NULL;
        q.niter = niter;
        q.stop_violation = o.stop_violation;
        q.stop_pg = o.stop_pg;
        q.precision_floor = o.precision_floor;
        q.inner_min_delta = o.min_delta;
        q.precision = o.precision;
        q.stop_kkt = o.stop_kkt;
        struct acd_solve_result r;
        struct acd_log_ctx ctx = {log, 0};
        int small = m + n <= 4000;
        int rc;
        /* The engine depends on the LP's size only, never on whether a
         * log is written: both engines sum in their own fixed orders, so
         * logging must not change which one runs. */
        if (o.engine == ACD_ENGINE_WORKGROUP
            || (o.engine == ACD_ENGINE_AUTO && small)) {
                rc = acd_dev_solve_batch(&dev, 1, &q, &r, NULL, NULL,
                                         acd_log_sink, &ctx);
        } else {
                rc = acd_dev_solve(dev, &q, &r, acd_log_sink, &ctx);
        }
        acd_check_result(&r, rc, "alm_solve");
        acd_record_stats(&r);
        /* x, lambda, and the state the original leaves in its approx
         * object: the updated row weights and the last penalised linear
         * term (alm.c:106-119, :194-212). */
        if (acd_dev_d2h(x, acd_dev_lp_vec(dev, ACD_V_X), n * 8)
            || acd_dev_d2h(lambda, acd_dev_lp_vec(dev, ACD_V_LAMBDA), m * 8)
            || acd_dev_d2h(a->weight, acd_dev_lp_vec(dev, ACD_V_WEIGHT), m * 8)
            || (niter && acd_dev_d2h(a->linear,
                                     acd_dev_lp_vec(dev, ACD_V_LINEAR), n * 8)))
                acd_fail("alm_solve: download");
        if (OUT_diagnosis != NULL)
                for (int i = 0; i < 3; i++) OUT_diagnosis[i] = r.diag[i];
        if (log) fflush(log);
        return r.ret;
}

void acd_alm_upload(alm_t * alm, const double * x, const double * lambda)
{
        approx_t * a = alm->approx;
        struct acd_dev_lp * dev = a->dev;
        size_t m = alm->nrhs, n = alm->nvars;
        if (acd_approx_upload(a)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_ALM_LINEAR), alm->linear, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAM_LO), alm->lambda_lower, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAM_HI), alm->lambda_upper, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_X), x, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(dev, ACD_V_LAMBDA), lambda, m * 8))
                acd_fail("alm_solve: upload");
}

void acd_alm_download(alm_t * alm, double * x, double * lambda,
                         size_t niter)
{
        approx_t * a = alm->approx;
        struct acd_dev_lp * dev = a->dev;
        size_t m = alm->nrhs, n = alm->nvars;
        if (acd_dev_d2h(x, acd_dev_lp_vec(dev, ACD_V_X), n * 8)
            || acd_dev_d2h(lambda, acd_dev_lp_vec(dev, ACD_V_LAMBDA), m * 8)
            || acd_dev_d2h(a->weight, acd_dev_lp_vec(dev, ACD_V_WEIGHT), m * 8)
            || (niter && acd_dev_d2h(a->linear,
                                     acd_dev_lp_vec(dev, ACD_V_LINEAR), n * 8)))
                acd_fail("alm_solve: download");
}

int alm_solve_batch(size_t count, alm_t * const * alms, size_t niter,
                    double * const * x, double * const * lambda,
                    double * const * OUT_diagnosis, int * OUT_ret,
                    const struct acd_alm_options * opt,
                    double * OUT_done_ms, uint64_t * OUT_inner,
                    struct acd_batch_info * OUT_info)
{
        struct acd_alm_options o;
        if (opt) o = *opt;
        else acd_alm_default_options(&o);
        if (count == 0) return 0;
        struct acd_dev_lp ** lps = calloc(count, sizeof(*lps));
        struct acd_solve_result * res = calloc(count, sizeof(*res));
        if (lps == NULL || res == NULL) acd_fail("alm_solve_batch: malloc");
        for (size_t i = 0; i < count; i++) {
                acd_alm_upload(alms[i], x[i], lambda[i]);
                lps[i] = alms[i]->approx->dev;
        }
        struct acd_solve_request q;
        memset(&q, 0, sizeof q);
        q.mode = ACD_MODE_ALM;
        q.niter = niter;
        q.stop_violation = o.stop_violation;
        q.stop_pg = o.stop_pg;
        q.precision_floor = o.precision_floor;
        q.inner_min_delta = o.min_delta;
        q.precision = o.precision;
        q.stop_kkt = o.stop_kkt;
        struct acd_batch_stats st;
        memset(&st, 0, sizeof st);
        int rc = acd_dev_solve_batch(lps, count, &q, res, OUT_done_ms, &st,
                                     NULL, NULL);
        if (rc) acd_fail("alm_solve_batch: device solve failed");
        for (size_t i = 0; i < count; i++) {
                /* A failed LP (non-finite iterate, positive model
                 * estimate) is reported in its own return code; the
                 * other LPs of the batch are unaffected. */
                acd_alm_download(alms[i], x[i], lambda[i], niter);
                if (OUT_diagnosis != NULL && OUT_diagnosis[i] != NULL)
                        for (int k = 0; k < 3; k++)
                                OUT_diagnosis[i][k] = res[i].diag[k];
                if (OUT_ret != NULL)
                        OUT_ret[i] = res[i].error ? -res[i].error : res[i].ret;
                if (OUT_inner != NULL) OUT_inner[i] = res[i].inner_iterations;
        }
        if (OUT_info != NULL) {
                OUT_info->device_ms = st.device_ms;
                OUT_info->launches = st.launches;
                OUT_info->inner_iterations = st.inner_iterations;
                OUT_info->alg_bytes = st.alg_bytes;
                OUT_info->kernel_ms = st.kernel_ms;
                OUT_info->lp_ops_ms = st.lp_ops_ms;
                OUT_info->lp_ctl_ms = st.lp_ctl_ms;
                for (int k = 0; k < ACD_BATCH_CLASSES; k++)
                        OUT_info->waves_per_lp[k] = st.waves_per_lp[k];
        }
        free(lps);
        free(res);
        return 0;
}

/* alm.c:289-322 */
int alm_solve(alm_t * alm, size_t niter, double * x, size_t nvars,
              double * lambda, size_t nconstraints, FILE * log,
              double * OUT_diagnosis, thread_pool_t * pool)
{
        (void)pool;
        return alm_solve_ex(alm, niter, x, nvars, lambda, nconstraints, log,
                            OUT_diagnosis, NULL);
}

/* alm.c:324-328 (reads as many values as are present, zero-filling the
 * rest; the original asserts on short input) */
void read_doubles(FILE * stream, double * out, size_t n)
{
        for (size_t i = 0; i < n; i++)
                if (fscanf(stream, " %lf", out + i) != 1) out[i] = 0;
}

static int read_exact(FILE * stream, double * out, size_t n)
{
        for (size_t i = 0; i < n; i++)
                if (fscanf(stream, " %lf", out + i) != 1) return -1;
        return 0;
}

/* alm.c:330-364 */
alm_t * alm_read(FILE * stream)
{
        sparse_matrix_t * m = sparse_matrix_read(stream, 0);
        if (m == NULL) return NULL;
        size_t nvars = sparse_matrix_ncolumns(m), nrhs = sparse_matrix_nrows(m);
        double * rhs = calloc(nrhs + 1, sizeof(double));
        double * linear = calloc(nvars + 1, sizeof(double));
        double * lower = calloc(nvars + 1, sizeof(double));
        double * upper = calloc(nvars + 1, sizeof(double));
        double * llo = calloc(nrhs + 1, sizeof(double));
        double * lhi = calloc(nrhs + 1, sizeof(double));
        alm_t * alm = NULL;
        if (rhs && linear && lower && upper && llo && lhi
            && !read_exact(stream, rhs, nrhs)
            && !read_exact(stream, linear, nvars)
            && !read_exact(stream, lower, nvars)
            && !read_exact(stream, upper, nvars)
            && !read_exact(stream, llo, nrhs)
            && !read_exact(stream, lhi, nrhs))
                alm = alm_make(m, nrhs, rhs, nvars, linear, lower, upper,
                               llo, lhi);
        if (alm == NULL) sparse_matrix_free(m);
        free(rhs);
        free(linear);
        free(lower);
        free(upper);
        free(llo);
        free(lhi);
        return alm;
}
