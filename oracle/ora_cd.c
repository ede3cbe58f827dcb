/* ora_cd.c -- oracle restatement of the reference's randomised coordinate
 * descent (the as-shipped ALM inner solver).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows approx/cd.c:27-385.  Randomness is glibc random(), consumed in
 * the same order as the reference, so a process-level srandom(s) before
 * both reproduces the same trajectory.  Differences that do not change
 * any output: the reference callocs 11*n*n uint32 for the sampled column
 * list (cd.c:298-299) but only uses the first 11*n; the oracle allocates
 * 11*n.  The reference's fisher_yates can index one past the end when
 * random() returns RAND_MAX (cd.c:31); the oracle keeps that expression.
 */
#include "ora_internal.h"
#include <assert.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* approx/cd.c:27-40 */
static void shuffle_u32(uint32_t * data, size_t nmemb)
{
        for (size_t i = 0; i < nmemb; i++) {
                size_t j = (1.0 * random() / RAND_MAX) * (nmemb - i);
                j += i;
                uint32_t t = data[i];
                data[i] = data[j];
                data[j] = t;
        }
}

/* approx/cd.c:42-52 */
static void clamp_box(double * x, const ora_approx_t * a)
{
        for (size_t i = 0; i < a->nvars; i++)
                x[i] = fmin(fmax(a->lower[i], x[i]), a->upper[i]);
}

static double seq_norm2(const double * x, size_t n)
{
        double acc = 0;
        for (size_t i = 0; i < n; i++) acc += x[i] * x[i];
        return sqrt(acc);
}

/* approx/cd.c:64-78 */
static double norm_pg(const double * g, size_t n, const double * x,
                      const double * l, const double * u)
{
        double acc = 0;
        for (size_t i = 0; i < n; i++) {
                double xi = x[i];
                double xp = xi - g[i];
                if (xp < l[i]) xp = l[i];
                if (xp > u[i]) xp = u[i];
                double d = xp - xi;
                acc += d * d;
        }
        return sqrt(acc);
}

/* approx/cd.c:80-109: value = 1/2 sum w v^2 + c'x, g = A'(w v) + c. */
static double full_gradient(double * g, double * viol, const double * x,
                            const ora_approx_t * a)
{
        double value = 0;
        double * scaled = calloc(a->nrhs + 1, sizeof(double));
        ora_sparse_matrix_multiply(viol, a->nrhs, a->matrix, x, a->nvars, 0);
        for (size_t i = 0; i < a->nrhs; i++) {
                double v = viol[i] -= a->rhs[i];
                double w = a->weight[i];
                scaled[i] = w * v;
                value += w * v * v;
        }
        value *= .5;
        ora_sparse_matrix_multiply(g, a->nvars, a->matrix, scaled, a->nrhs, 1);
        for (size_t i = 0; i < a->nvars; i++) {
                double li = a->linear[i];
                g[i] += li;
                value += li * x[i];
        }
        free(scaled);
        return value;
}

struct spvec { size_t nnz; uint32_t * idx; double * val; };

/* approx/cd.c:128-146: keep entries with |x| >= 1e-10. */
static void sparsify(struct spvec * out, const double * x, size_t n)
{
        out->idx = calloc(n + 1, sizeof(uint32_t));
        out->val = calloc(n + 1, sizeof(double));
        out->nnz = 0;
        for (size_t i = 0; i < n; i++) {
                if (fabs(x[i]) < 1e-10) continue;
                out->idx[out->nnz] = (uint32_t)i;
                out->val[out->nnz] = x[i];
                out->nnz++;
        }
}

struct column {
        double probability, inv_step, lo, hi;
        struct spvec q;     /* column i of A' W A */
};

/* approx/cd.c:164-204 */
static void column_info(struct column * ci, const ora_approx_t * a,
                        uint32_t col)
{
        size_t n = a->nvars, m = a->nrhs;
        double * dense = calloc(m + 1, sizeof(double));
        const uint32_t * ptr = ora_csr_row_ptr(a->matrix, 1);
        const uint32_t * rows = ora_csr_columns(a->matrix, 1);
        const double * vals = ora_csr_values(a->matrix, 1);
        for (uint32_t k = ptr[col]; k < ptr[col + 1]; k++)
                dense[rows[k]] += vals[k];
        double norm = 0;
        for (size_t i = 0; i < m; i++) {
                double d = dense[i], w = a->weight[i];
                dense[i] = w * d;
                norm += w * d * d;
        }
        double * q = calloc(n + 1, sizeof(double));
        ora_sparse_matrix_multiply(q, n, a->matrix, dense, m, 1);
        sparsify(&ci->q, q, n);
        ci->probability = (norm * norm) + 1e-2;
        ci->inv_step = 1.0 / norm;
        assert(ci->inv_step > 0);
        ci->lo = a->lower[col];
        ci->hi = a->upper[col];
        free(q);
        free(dense);
}

struct cd {
        double * x, * g, * viol;
        size_t ncol;
        struct column * info;
};

/* approx/cd.c:222-239: every column shares the smallest 1/norm. */
static void cd_init(struct cd * s, const ora_approx_t * a)
{
        s->x = calloc(a->nvars + 1, sizeof(double));
        s->g = calloc(a->nvars + 1, sizeof(double));
        s->viol = calloc(a->nrhs + 1, sizeof(double));
        s->info = calloc(a->nvars + 1, sizeof(struct column));
        s->ncol = a->nvars;
        double min_inv = HUGE_VAL;
        for (size_t i = 0; i < a->nvars; i++) {
                column_info(s->info + i, a, (uint32_t)i);
                min_inv = fmin(min_inv, s->info[i].inv_step);
        }
        for (size_t i = 0; i < a->nvars; i++) s->info[i].inv_step = min_inv;
}

static void cd_free(struct cd * s)
{
        for (size_t i = 0; i < s->ncol; i++) {
                free(s->info[i].q.idx);
                free(s->info[i].q.val);
        }
        free(s->info);
        free(s->viol);
        free(s->g);
        free(s->x);
}

/* approx/cd.c:252-282 */
static void one_step(struct cd * s, uint32_t col, double scale)
{
        double xi = s->x[col], gi = s->g[col];
        if (gi == 0) return;
        struct column * ci = s->info + col;
        double xp = xi - scale * (gi * ci->inv_step);
        if (xp < ci->lo) xp = ci->lo;
        if (xp > ci->hi) xp = ci->hi;
        double st = xp - xi;
        if (st == 0) return;
        s->x[col] = xp;
        for (size_t k = 0; k < ci->q.nnz; k++)
                s->g[ci->q.idx[k]] += st * ci->q.val[k];
}

/* approx/cd.c:284-330 */
static double major_step(struct cd * s, ora_approx_t * a, double offset,
                         FILE * log, double scale)
{
        clamp_box(s->x, a);
        double value = full_gradient(s->g, s->viol, s->x, a);
        if (log != NULL)
                fprintf(log, "\tEntering: %12g %12g %12g\n", value + offset,
                        norm_pg(s->g, s->ncol, s->x, a->lower, a->upper),
                        seq_norm2(s->viol, a->nrhs));
        size_t cnt = 0, maxn = 11 * s->ncol;
        uint32_t * cols = calloc(maxn + 1, sizeof(uint32_t));
        double total = 0;
        for (size_t i = 0; i < s->ncol; i++) total += s->info[i].probability;
        for (size_t i = 0; i < s->ncol; i++) {
                double p = s->info[i].probability;
                size_t count = ceil(10.0 * s->ncol * p / total);
                for (size_t j = 0; j < count; j++) {
                        assert(cnt < maxn);
                        cols[cnt++] = (uint32_t)i;
                }
        }
        shuffle_u32(cols, cnt);
        for (size_t i = 0; i < cnt; i++) one_step(s, cols[i], scale);
        free(cols);
        clamp_box(s->x, a);
        value = full_gradient(s->g, s->viol, s->x, a);
        if (log != NULL)
                printf("\t\tLeaving:  %12g %12g %12g\n", value + offset,
                       norm_pg(s->g, s->ncol, s->x, a->lower, a->upper),
                       seq_norm2(s->viol, a->nrhs));
        return value + offset;
}

/* approx/cd.c:332-385 (uses the caller's, unpermuted, matrix). */
int ora_cd_solve(double * x, size_t n, ora_approx_t * a, size_t niter,
                 double max_pg, double max_value, double min_delta,
                 FILE * log, size_t period, double * diag, double offset)
{
        (void)min_delta;
        period /= 10;
        assert(n == a->nvars);
        struct cd s;
        cd_init(&s, a);
        memcpy(s.x, x, n * sizeof(double));
        int reason = 0;
        size_t i;
        for (i = 0; i < niter; i++) {
                int print = (i == 0) || (period && ((i + 1) % period == 0))
                        || (i + 1 == niter);
                if (print && log != NULL) fprintf(log, "\t%6zu ", i + 1);
                double z = major_step(&s, a, offset, print ? log : NULL, .5);
                double pg = norm_pg(s.g, s.ncol, s.x, a->lower, a->upper);
                if (z < max_value) { reason = 1; break; }
                if (pg < max_pg) { reason = 2; break; }
        }
        if (diag != NULL) {
                diag[0] = full_gradient(s.g, s.viol, s.x, a) + offset;
                diag[1] = seq_norm2(s.g, s.ncol);
                diag[2] = norm_pg(s.g, s.ncol, s.x, a->lower, a->upper);
                diag[3] = 0;
                diag[4] = i + 1;
        }
        memcpy(x, s.x, n * sizeof(double));
        cd_free(&s);
        return reason;
}
