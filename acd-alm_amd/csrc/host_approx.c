/* host_approx.c -- the approx.h entry points.
 *
 * Mirrors approx/approx.c:45-216 and :517-708 of pkhuong/acd-alm: the
 * object owns host copies of its vectors (returned, mutable, by the
 * accessors) and a permuted working copy of the matrix; every solver
 * iteration runs on the device (dev_solver.hip).  approx_update() is where
 * host edits reach the device.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_host.h"
#include "acd_approx.h"

static double * copy_default(const double * x, size_t n, double missing)
{
        double * out = malloc(sizeof(double) * (n ? n : 1));
        if (out == NULL) return NULL;
        for (size_t i = 0; i < n; i++) out[i] = x ? x[i] : missing;
        return out;
}

/* approx/approx.c:45-92 */
approx_t * approx_make(sparse_matrix_t * constraints, size_t nrhs,
                       const double * rhs, const double * weight,
                       size_t nvars, const double * linear,
                       const double * lower, const double * upper)
{
        ACD_REQUIRE(nrhs == sparse_matrix_nrows(constraints),
                    "approx_make: nrhs does not match the matrix");
        ACD_REQUIRE(nvars == sparse_matrix_ncolumns(constraints),
                    "approx_make: nvars does not match the matrix");
        approx_t * a = calloc(1, sizeof(*a));
        if (a == NULL) return NULL;
        a->nrhs = nrhs;
        a->nvars = nvars;
        a->matrix = constraints;
        a->rhs = copy_default(rhs, nrhs, 0);
        a->weight = copy_default(weight, nrhs, 1);
        a->linear = copy_default(linear, nvars, 0);
        a->lower = copy_default(lower, nvars, -HUGE_VAL);
        a->upper = copy_default(upper, nvars, HUGE_VAL);
        a->permuted = sparse_matrix_copy(constraints, 1);
        if (!a->rhs || !a->weight || !a->linear || !a->lower || !a->upper
            || !a->permuted) {
                approx_free(a);
                return NULL;
        }
        a->dev = acd_dev_lp_create(&constraints->dev, &a->permuted->dev);
        if (a->dev == NULL) {
                approx_free(a);
                return NULL;
        }
        approx_update(a);
        return a;
}

/* approx/approx.c:111-133 */
int approx_free(approx_t * a)
{
        if (a == NULL) return 0;
        acd_dev_lp_free(a->dev);
        sparse_matrix_free(a->permuted);
        free(a->rhs);
        free(a->weight);
        free(a->linear);
        free(a->lower);
        free(a->upper);
        memset(a, 0, sizeof(*a));
        free(a);
        return 0;
}

/* approx/approx.c:94-109 */
sparse_matrix_t * approx_matrix(approx_t * a) { return a->matrix; }
size_t approx_nrhs(approx_t * a) { return a->nrhs; }
double * approx_rhs(approx_t * a) { return a->rhs; }
double * approx_weight(approx_t * a) { return a->weight; }
size_t approx_nvars(approx_t * a) { return a->nvars; }
double * approx_linear(approx_t * a) { return a->linear; }
double * approx_lower(approx_t * a) { return a->lower; }
double * approx_upper(approx_t * a) { return a->upper; }

int acd_approx_upload(approx_t * a)
{
        size_t m = a->nrhs, n = a->nvars;
        if (acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_RHS), a->rhs, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_WEIGHT), a->weight, m * 8)
            || acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_LINEAR), a->linear, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_LOWER), a->lower, n * 8)
            || acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_UPPER), a->upper, n * 8))
                return -1;
        return 0;
}

/* approx/approx.c:188-216 */
int approx_update(approx_t * a)
{
        if (acd_approx_upload(a) || acd_dev_approx_update(a->dev))
                acd_fail("approx_update");
        return 0;
}

/* ---- log replay (approx/approx.c:504-515, :550-553, :618-627) ---- */
static void print_log(FILE * log, size_t k, double value, double ng,
                      double pg, double step, double diff)
{
        if (log == NULL) return;
        if (diff < HUGE_VAL)
                fprintf(log, "\t%10zu %12g %12g %12g %8g %12g\n",
                        k, value, ng, pg, step, diff);
        else
                fprintf(log, "\t%10zu %12g %12g %12g %8g\n",
                        k, value, ng, pg, step);
}

void acd_log_sink(void * p, const struct acd_event * e)
{
        struct acd_log_ctx * ctx = p;
        FILE * log = ctx->log;
        switch (e->type) {
        case ACD_EV_APPROX_BEGIN:
                ctx->restart = 0;
                break;
        case ACD_EV_RESTART:
                if (!ctx->restart) {
                        ctx->restart = 1;
                        if (log) fprintf(log, "\t\t ");
                }
                if (log) {
                        fprintf(log, "R");
                        fflush(log);
                }
                break;
        case ACD_EV_NEWLINE:
                /* the original writes this newline to stdout */
                ctx->restart = 0;
                printf("\n");
                break;
        case ACD_EV_PRINT:
                print_log(log, (size_t)e->k, e->a[0], e->a[1], e->a[2],
                          e->a[3], e->a[4]);
                break;
        case ACD_EV_MPG:
                if (log) fprintf(log, "%10zu: %12g %12g\n", (size_t)e->k,
                                 e->a[0], e->a[1]);
                break;
        case ACD_EV_ALM_ITER:
                if (log) fprintf(log, "%4zu: %12g %12g %12g %.18g %.18g",
                                 (size_t)e->k, e->a[0], e->a[1], e->a[2],
                                 e->a[3], e->a[4]);
                break;
        case ACD_EV_ALM_TAIL:
                if (log) {
                        if (e->a[0] != 0) fprintf(log, " -stalled-\n");
                        else fprintf(log, " (%8g)\n", e->a[1]);
                }
                break;
        default:
                break;
        }
}

static struct acd_solve_stats g_stats;

void acd_record_stats(const struct acd_solve_result * r)
{
        g_stats.inner_iterations = r->inner_iterations;
        g_stats.outer_iterations = r->outer_iterations;
        g_stats.launches = r->launches;
        g_stats.device_ms = r->device_ms;
}

void acd_last_solve_stats(struct acd_solve_stats * out)
{
        if (out) *out = g_stats;
}

void acd_check_result(const struct acd_solve_result * r, int rc,
                      const char * who)
{
        static char msg[160];
        if (rc) {
                snprintf(msg, sizeof msg, "%s: device solve failed", who);
                acd_fail(msg);
        }
        switch (r->error) {
        case ACD_ERR_NONE: return;
        case ACD_ERR_UNBOUNDED:
                snprintf(msg, sizeof msg,
                         "%s: max |z| is not finite (unbounded problem)", who);
                break;
        case ACD_ERR_ESTIMATE:
                snprintf(msg, sizeof msg,
                         "%s: expected improvement > 0", who);
                break;
        case ACD_ERR_BARRIER:
                snprintf(msg, sizeof msg, "%s: grid barrier timed out", who);
                break;
        default:
                snprintf(msg, sizeof msg, "%s: launch failure", who);
                break;
        }
        acd_fail(msg);
}

static int run_inner(int mode, double * x, size_t n, approx_t * a,
                     size_t niter, double max_pg, double max_value,
                     double min_delta, FILE * log, size_t period,
                     double * OUT_diagnosis, double offset, const char * who)
{
        ACD_REQUIRE(n == a->nvars, "solve: n does not match the instance");
        double * dx = acd_dev_lp_vec(a->dev, ACD_V_X);
        if (acd_dev_h2d(dx, x, n * sizeof(double))) acd_fail(who);
        struct acd_solve_request q;
        memset(&q, 0, sizeof q);
        q.mode = mode;
        q.log = log != NULL;
        q.niter = niter;
        q.max_pg = max_pg;
        q.max_value = max_value;
        q.min_delta = min_delta;
        q.offset = offset;
        q.period = period;
        struct acd_solve_result r;
        struct acd_log_ctx ctx = {log, 0};
        int rc = acd_dev_solve(a->dev, &q, &r, acd_log_sink, &ctx);
        acd_check_result(&r, rc, who);
        acd_record_stats(&r);
        if (acd_dev_d2h(x, dx, n * sizeof(double))) acd_fail(who);
        if (OUT_diagnosis != NULL)
                for (int i = 0; i < 5; i++) OUT_diagnosis[i] = r.diag[i];
        if (log) fflush(log);
        return r.ret;
}

/* approx/approx.c:576-708 */
int approx_solve(double * x, size_t n, approx_t * a, size_t niter,
                 double max_pg, double max_value, double min_delta,
                 FILE * log, size_t period, double * OUT_diagnosis,
                 double offset, thread_pool_t * pool)
{
        (void)pool;
        return run_inner(ACD_MODE_APPROX, x, n, a, niter, max_pg, max_value,
                         min_delta, log, period, OUT_diagnosis, offset,
                         "approx_solve");
}

/* approx/approx.c:517-573 */
int mpg_solve(double * x, size_t n, approx_t * a, size_t niter,
              double max_pg, double max_value, double min_delta,
              FILE * log, size_t period, double * OUT_diagnosis,
              double offset, thread_pool_t * pool)
{
        (void)pool;
        double diag[5];
        run_inner(ACD_MODE_MPG, x, n, a, niter, max_pg, max_value, min_delta,
                  log, period, diag, offset, "mpg_solve");
        /* The original fills OUT_diagnosis like approx_solve. */
        if (OUT_diagnosis != NULL)
                for (int i = 0; i < 5; i++) OUT_diagnosis[i] = diag[i];
        return 2;
}
