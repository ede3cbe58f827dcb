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
        g_stats.kkt = r->kkt;
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

/* Parallel randomised coordinate descent with ESO steps (no counterpart
 * in the reference, whose cd_solve (approx/cd.c:332-385) takes one
 * coordinate at a time): each step updates a random sample of the
 * coordinates, each with probability `sample`, by
 * x_j <- clamp(x_j - g_j / v'_j) with v'_j = sum_i w_i a_ij^2
 * (1 + sample (omega_i - 1)), omega_i the nonzeros of row i (the expected
 * separable overapproximation of Richtarik and Takac for independent
 * sampling, with per-row factors).  An epoch is ceil(1 / sample) steps;
 * niter counts epochs; the stopping tests and diagnosis are cd_solve's. */
int acd_cd_solve_parallel(double * x, size_t n, approx_t * a, size_t niter,
                          double max_pg, double max_value, double offset,
                          double sample, uint64_t seed,
                          double * OUT_diagnosis)
{
        ACD_REQUIRE(n == a->nvars, "solve: n does not match the instance");
        if (!(sample > 0) || sample > 1) return -1;
        double * dx = acd_dev_lp_vec(a->dev, ACD_V_X);
        if (acd_dev_h2d(dx, x, n * sizeof(double)))
                acd_fail("acd_cd_solve_parallel");
        struct acd_solve_request q;
        memset(&q, 0, sizeof q);
        q.mode = ACD_MODE_PCD;
        q.niter = niter;
        q.max_pg = max_pg;
        q.max_value = max_value;
        q.offset = offset;
        q.pcd_p = sample;
        q.pcd_seed = seed;
        q.pcd_steps = (uint32_t)ceil(1.0 / sample);
        struct acd_solve_result r;
        int rc = acd_dev_solve(a->dev, &q, &r, NULL, NULL);
        acd_check_result(&r, rc, "acd_cd_solve_parallel");
        acd_record_stats(&r);
        if (acd_dev_d2h(x, dx, n * sizeof(double)))
                acd_fail("acd_cd_solve_parallel");
        if (OUT_diagnosis != NULL)
                for (int i = 0; i < 5; i++) OUT_diagnosis[i] = r.diag[i];
        return r.ret;
}

/* ---- randomised coordinate descent (approx/cd.c) ---- */

/* approx/cd.c:27-40: the column list shuffled with the process's random(),
 * as the original does (it may read one past the end when random()
 * returns RAND_MAX; the list keeps a zeroed slot there). */
static void shuffle_columns(uint32_t * data, size_t nmemb)
{
        for (size_t i = 0; i < nmemb; i++) {
                size_t j = (1.0 * random() / RAND_MAX) * (nmemb - i);
                j += i;
                uint32_t t = data[i];
                data[i] = data[j];
                data[j] = t;
        }
}

static void cd_fail(struct acd_dev_cd * cd, uint32_t * cols, double * prob)
{
        acd_dev_cd_free(cd);
        free(cols);
        free(prob);
        acd_fail("cd_solve");
}

/* approx/cd.c:284-330; 0 on success with the value in *out. */
static int cd_major_step(struct acd_dev_cd * cd, approx_t * a,
                         const double * prob, uint32_t * cols, double offset,
                         FILE * log, double scale, double * out)
{
        size_t n = a->nvars;
        double value, pg, nv;
        if (acd_dev_cd_clamp(cd) || acd_dev_cd_gradient(cd, &value))
                return -1;
        if (log != NULL) {
                if (acd_dev_cd_norms(cd, &pg, NULL, &nv)) return -1;
                fprintf(log, "\tEntering: %12g %12g %12g\n", value + offset,
                        pg, nv);
        }
        size_t cnt = 0, maxn = 11 * n;
        double total = 0;
        for (size_t i = 0; i < n; i++) total += prob[i];
        for (size_t i = 0; i < n; i++) {
                size_t count = ceil(10.0 * n * prob[i] / total);
                ACD_REQUIRE(cnt + count <= maxn,
                            "cd_solve: sampled column list overflow");
                for (size_t j = 0; j < count; j++) cols[cnt++] = (uint32_t)i;
        }
        memset(cols + cnt, 0, (maxn + 1 - cnt) * sizeof(uint32_t));
        shuffle_columns(cols, cnt);
        if (acd_dev_cd_steps(cd, cols, cnt, scale) || acd_dev_cd_clamp(cd)
            || acd_dev_cd_gradient(cd, &value))
                return -1;
        if (log != NULL) {
                if (acd_dev_cd_norms(cd, &pg, NULL, &nv)) return -1;
                /* the original prints this line to stdout */
                printf("\t\tLeaving:  %12g %12g %12g\n", value + offset, pg,
                       nv);
        }
        *out = value + offset;
        return 0;
}

/* approx/cd.c:332-385 */
int cd_solve(double * x, size_t n, approx_t * a, size_t niter,
             double max_pg, double max_value, double min_delta,
             FILE * log, size_t period, double * OUT_diagnosis,
             double offset, thread_pool_t * pool)
{
        (void)min_delta;
        (void)pool;
        period /= 10;
        ACD_REQUIRE(n == a->nvars, "cd_solve: n does not match the instance");
        if (acd_approx_upload(a)) acd_fail("cd_solve");
        struct acd_dev_cd * cd = acd_dev_cd_create(&a->matrix->dev, a->dev);
        double * prob = malloc(sizeof(double) * (n + 1));
        uint32_t * cols = calloc(11 * n + 1, sizeof(uint32_t));
        if (cd == NULL || prob == NULL || cols == NULL
            || acd_dev_cd_setup(cd, prob)
            || acd_dev_h2d(acd_dev_lp_vec(a->dev, ACD_V_X), x,
                           n * sizeof(double)))
                cd_fail(cd, cols, prob);
        int reason = 0;
        size_t i;
        for (i = 0; i < niter; i++) {
                int print = (i == 0) || (period && ((i + 1) % period == 0))
                        || (i + 1 == niter);
                if (print && log != NULL) fprintf(log, "\t%6zu ", i + 1);
                double z = 0, pg = 0;
                if (cd_major_step(cd, a, prob, cols, offset,
                                  print ? log : NULL, .5, &z)
                    || acd_dev_cd_norms(cd, &pg, NULL, NULL))
                        cd_fail(cd, cols, prob);
                if (z < max_value) {
                        reason = 1;
                        break;
                }
                if (pg < max_pg) {
                        reason = 2;
                        break;
                }
        }
        if (OUT_diagnosis != NULL) {
                double value, pg, ng;
                if (acd_dev_cd_gradient(cd, &value)
                    || acd_dev_cd_norms(cd, &pg, &ng, NULL))
                        cd_fail(cd, cols, prob);
                OUT_diagnosis[0] = value + offset;
                OUT_diagnosis[1] = ng;
                OUT_diagnosis[2] = pg;
                OUT_diagnosis[3] = 0;
                OUT_diagnosis[4] = i + 1;
        }
        if (acd_dev_d2h(x, acd_dev_lp_vec(a->dev, ACD_V_X),
                        n * sizeof(double)))
                cd_fail(cd, cols, prob);
        acd_dev_cd_free(cd);
        free(cols);
        free(prob);
        if (log) fflush(log);
        return reason;
}
