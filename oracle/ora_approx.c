/* ora_approx.c -- oracle restatement of the reference's APPROX solver.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows approx/approx.c:45-708, approx/vector.inc:1-62 and :209-335,
 * approx/gradient_value.inc, approx/step_project.inc and
 * approx/mpg.inc:155-199 in serial mode.  The reference's v2d loops are
 * written here as explicit even/odd lanes: element i goes to lane i & 1,
 * each lane accumulates sequentially, and the lanes are combined as
 * lane0 + lane1, which is what the SSE2 code computes.  SSE2 maxpd(a, b)
 * is (a > b ? a : b) and minpd(a, b) is (a < b ? a : b).
 */
#include "ora_internal.h"
#include <assert.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>


/* Test knob: when set, every reduction below visits its elements in
 * reverse order (each lane still accumulates sequentially).  The result
 * is another valid summation order, used by the fixtures to measure how
 * sensitive a trajectory is to the order of its sums; 0 (the default)
 * is the reference's order. */
static int g_reverse;
void ora_set_reverse_sums(int on) { g_reverse = on; }
#define FOR_IDX(i, n) for (size_t i##_ = 0, i = g_reverse ? (n) - 1 : 0; \
                           i##_ < (n); i##_++, i = g_reverse ? (n) - 1 - i##_ : i##_)

static inline double sse_max(double a, double b) { return a > b ? a : b; }
static inline double sse_min(double a, double b) { return a < b ? a : b; }

/* huge_alloc/huge_alloc.default.inc:6-27: zeroed with slop, so that the
 * pairwise loops may touch element n when n is odd. */
static double * zalloc(size_t n)
{
        return calloc(n + 2, sizeof(double));
}

static double * copy_default(const double * x, size_t n, double missing)
{
        double * out = zalloc(n);
        for (size_t i = 0; i < n; i++) out[i] = x ? x[i] : missing;
        return out;
}

/* approx/approx.c:45-70 */
static ora_approx_t * make_1(ora_matrix_t * m, size_t nrhs, const double * rhs,
                             const double * weight, size_t nvars,
                             const double * linear, const double * lower,
                             const double * upper)
{
        assert(nrhs == ora_sparse_matrix_nrows(m));
        assert(nvars == ora_sparse_matrix_ncolumns(m));
        ora_approx_t * a = calloc(1, sizeof(*a));
        a->nrhs = nrhs;
        a->nvars = nvars;
        a->matrix = m;
        a->rhs = copy_default(rhs, nrhs, 0);
        a->weight = copy_default(weight, nrhs, 1);
        a->linear = copy_default(linear, nvars, 0);
        a->lower = copy_default(lower, nvars, -HUGE_VAL);
        a->upper = copy_default(upper, nvars, HUGE_VAL);
        a->beta = calloc(nrhs + 2, sizeof(uint32_t));
        a->v = zalloc(nvars);
        a->inv_v = zalloc(nvars);
        return a;
}

/* approx/approx.c:135-186: beta_j = entries in row j; v_i accumulates
 * ((w_j * beta_j) * a) * a over the Z-ordered triplets. */
static void update_step_size(ora_approx_t * a)
{
        ora_matrix_t * m = a->matrix;
        size_t nnz = ora_sparse_matrix_nnz(m);
        const uint32_t * rows = ora_sparse_matrix_rows(m),
                * cols = ora_sparse_matrix_columns(m);
        const double * vals = ora_sparse_matrix_values(m);
        memset(a->beta, 0, a->nrhs * sizeof(uint32_t));
        memset(a->v, 0, a->nvars * sizeof(double));
        for (size_t k = 0; k < nnz; k++) a->beta[rows[k]]++;
        for (size_t k = 0; k < nnz; k++) {
                uint32_t r = rows[k];
                double value = vals[k];
                double w = a->weight[r];
                a->v[cols[k]] += w * a->beta[r] * value * value;
        }
        for (size_t i = 0; i < a->nvars; i++) {
                double vi = a->v[i];
                if (vi == 0 && a->linear[i] == 0) a->inv_v[i] = 1;
                else a->inv_v[i] = 1.0 / vi;
        }
}

/* approx/approx.c:188-216 */
int ora_approx_update(ora_approx_t * a)
{
        update_step_size(a);
        if (a->permuted) {
                ora_approx_t * p = a->permuted;
                ora_matrix_t * pm = p->matrix;
                assert(!ora_sparse_matrix_row_permute(pm, p->rhs, a->nrhs, a->rhs, 1));
                assert(!ora_sparse_matrix_row_permute(pm, p->weight, a->nrhs, a->weight, 1));
                assert(!ora_sparse_matrix_col_permute(pm, p->linear, a->nvars, a->linear, 1));
                assert(!ora_sparse_matrix_col_permute(pm, p->lower, a->nvars, a->lower, 1));
                assert(!ora_sparse_matrix_col_permute(pm, p->upper, a->nvars, a->upper, 1));
                update_step_size(p);
        }
        return 0;
}

/* approx/approx.c:72-92 */
ora_approx_t * ora_approx_make(ora_matrix_t * m, size_t nrhs,
                               const double * rhs, const double * weight,
                               size_t nvars, const double * linear,
                               const double * lower, const double * upper)
{
        ora_approx_t * a = make_1(m, nrhs, rhs, weight, nvars, linear,
                                  lower, upper);
        a->permuted = make_1(ora_sparse_matrix_copy(m, 1), nrhs, rhs, weight,
                             nvars, linear, lower, upper);
        ora_approx_update(a);
        return a;
}

static void free_1(ora_approx_t * a)
{
        free(a->rhs); free(a->weight); free(a->linear);
        free(a->lower); free(a->upper);
        free(a->beta); free(a->v); free(a->inv_v);
        free(a);
}

/* approx/approx.c:111-133 */
int ora_approx_free(ora_approx_t * a)
{
        if (a == NULL) return 0;
        if (a->permuted) {
                ora_sparse_matrix_free(a->permuted->matrix);
                free_1(a->permuted);
        }
        free_1(a);
        return 0;
}

ora_matrix_t * ora_approx_matrix(ora_approx_t * a) { return a->matrix; }
double * ora_approx_rhs(ora_approx_t * a) { return a->rhs; }
double * ora_approx_weight(ora_approx_t * a) { return a->weight; }
double * ora_approx_linear(ora_approx_t * a) { return a->linear; }
double * ora_approx_lower(ora_approx_t * a) { return a->lower; }
double * ora_approx_upper(ora_approx_t * a) { return a->upper; }
const double * ora_approx_step_v(ora_approx_t * a, int p)
{ return p ? a->permuted->v : a->v; }
const double * ora_approx_step_inv_v(ora_approx_t * a, int p)
{ return p ? a->permuted->inv_v : a->inv_v; }

/* ---- vectors with a cached violation (approx/vector.inc:1-62) ---- */
struct vec {
        double * x;
        double * viol;      /* weight (A x - b) */
        size_t n, m;
        int violp;
        double quad;        /* 1/2 sum weight (Ax-b)^2, NaN = stale */
        double value;       /* linear'x + quad, NaN = stale */
};

static void vec_init(struct vec * v, size_t n, size_t m)
{
        v->n = n;
        v->m = m;
        v->x = zalloc(n);
        v->viol = zalloc(m);
        v->violp = 0;
        v->quad = v->value = nan("");
}

static void vec_free(struct vec * v)
{
        free(v->x);
        free(v->viol);
        memset(v, 0, sizeof(*v));
}

static void vec_copy(struct vec * dst, const struct vec * src)
{
        memcpy(dst->x, src->x, src->n * sizeof(double));
        if (src->violp) {
                memcpy(dst->viol, src->viol, src->m * sizeof(double));
                dst->violp = 1;
        } else {
                dst->violp = 0;
        }
        dst->quad = src->quad;
        dst->value = src->value;
}

/* ---- kernels ---- */

/* approx/gradient_value.inc:15-81: viol = w (A x - b), quad = 1/2 sum
 * viol*(Ax-b) in even/odd row lanes. */
static void compute_violation(struct vec * xv, ora_approx_t * a)
{
        size_t m = a->nrhs;
        ora_sparse_matrix_multiply(xv->viol, m, a->matrix, xv->x, a->nvars, 0);
        double acc[2] = {0, 0};
        FOR_IDX(i, m) {
                double v = xv->viol[i] - a->rhs[i];
                double s = a->weight[i] * v;
                xv->viol[i] = s;
                acc[i & 1] += s * v;
        }
        xv->quad = .5 * (acc[0] + acc[1]);
        xv->violp = 1;
}

/* approx/vector.inc:209-234 */
static double dot_pair(const double * x, const double * y, size_t n)
{
        double acc[2] = {0, 0};
        FOR_IDX(i, n) acc[i & 1] += x[i] * y[i];
        return acc[0] + acc[1];
}

/* approx/gradient_value.inc:111-133 */
static double compute_value(ora_approx_t * a, struct vec * xv)
{
        if (!isnan(xv->value)) return xv->value;
        double lin = dot_pair(a->linear, xv->x, a->nvars);
        if (isnan(xv->quad)) compute_violation(xv, a);
        return xv->value = xv->quad + lin;
}

/* approx/gradient_value.inc:158-181: g = A' viol + c */
static void gradient(struct vec * g, ora_approx_t * a, struct vec * xv)
{
        if (!xv->violp) compute_violation(xv, a);
        ora_sparse_matrix_multiply(g->x, a->nvars, a->matrix, xv->viol,
                                   a->nrhs, 1);
        for (size_t i = 0; i < a->nvars; i++) g->x[i] += a->linear[i];
}

/* approx/gradient_value.inc:237-268: both gradients in one pass. */
static void gradient2(struct vec * g0, struct vec * g1, ora_approx_t * a,
                      struct vec * x0, struct vec * x1)
{
        if (!x0->violp) compute_violation(x0, a);
        if (!x1->violp) compute_violation(x1, a);
        double * out[2] = {g0->x, g1->x};
        const double * in[2] = {x0->viol, x1->viol};
        ora_sparse_matrix_multiply_2(out, a->nvars, a->matrix, in, a->nrhs, 1);
        for (size_t i = 0; i < a->nvars; i++) {
                double c = a->linear[i];
                g0->x[i] += c;
                g1->x[i] += c;
        }
}

/* approx/step_project.inc:11-25 */
static void project(struct vec * xv, const double * l, const double * u)
{
        for (size_t i = 0; i < xv->n; i++) {
                double lo = sse_max(l[i], xv->x[i]);
                xv->x[i] = sse_min(lo, u[i]);
        }
        xv->violp = 0;
        xv->quad = xv->value = nan("");
}

/* approx/step_project.inc:34-94: zp = clamp(z - g*((1-1e-6)/theta*inv_v)),
 * returns max |zp|. */
static double step(struct vec * zp, double theta, const struct vec * g,
                   const struct vec * z, ora_approx_t * a)
{
        double itheta = (1 - 1e-6) / theta;
        double mx[2] = {0, 0};
        for (size_t i = 0; i < zp->n; i++) {
                double s = itheta * a->inv_v[i];
                double t = z->x[i] - g->x[i] * s;
                t = sse_max(a->lower[i], t);
                t = sse_min(a->upper[i], t);
                zp->x[i] = t;
                mx[i & 1] = sse_max(mx[i & 1], fabs(t));
        }
        double r = sse_max(mx[0], mx[1]);
        assert(r < HUGE_VAL);
        zp->violp = 0;
        zp->value = zp->quad = nan("");
        return r;
}

/* approx/step_project.inc:103-131 and the fused forms at :179-236 and
 * :320-375: the long step with length len, and the model decrease
 * sum g*delta + 1/2 sum delta^2 * (theta/len) * v. */
static double long_step(struct vec * zp, double theta, double len,
                        const struct vec * g, const struct vec * z,
                        ora_approx_t * a)
{
        assert(len >= 1);
        double itheta = (len - 1e-6) / theta;
        double theta2 = theta / len;
        double lin[2] = {0, 0}, quad[2] = {0, 0};
        FOR_IDX(i, zp->n) {
                double gi = g->x[i], zi = z->x[i];
                double s = itheta * a->inv_v[i];
                double t = zi - gi * s;
                t = sse_max(a->lower[i], t);
                t = sse_min(a->upper[i], t);
                zp->x[i] = t;
                double d = t - zi;
                lin[i & 1] += gi * d;
                quad[i & 1] += d * d * theta2 * a->v[i];
        }
        zp->violp = 0;
        zp->value = zp->quad = nan("");
        return (lin[0] + lin[1]) + .5 * (quad[0] + quad[1]);
}

/* approx/approx.c:227-256 */
static double dot_diff(const struct vec * g, const struct vec * z,
                       const struct vec * zp)
{
        double acc[2] = {0, 0};
        FOR_IDX(i, g->n)
                acc[i & 1] += g->x[i] * (zp->x[i] - z->x[i]);
        return acc[0] + acc[1];
}

/* approx/approx.c:258-278 */
static double pg_norm(const struct vec * g, const struct vec * x,
                      const double * l, const double * u)
{
        double acc[2] = {0, 0};
        FOR_IDX(i, g->n) {
                double xi = x->x[i];
                double xp = xi - g->x[i];
                xp = sse_max(l[i], xp);
                xp = sse_min(u[i], xp);
                double d = xi - xp;
                acc[i & 1] += d * d;
        }
        return sqrt(acc[0] + acc[1]);
}

/* approx/approx.c:480-502 */
static double diff_norm(const double * x, const double * y, size_t n)
{
        double acc[2] = {0, 0};
        FOR_IDX(i, n) {
                double d = x[i] - y[i];
                acc[i & 1] += d * d;
        }
        return sqrt(acc[0] + acc[1]);
}

static double norm2(const struct vec * v)
{
        double acc[2] = {0, 0};
        FOR_IDX(i, v->n) acc[i & 1] += v->x[i] * v->x[i];
        return sqrt(acc[0] + acc[1]);
}

/* approx/vector.inc:236-335: x <- (1-t)x + t zp (point and violation),
 * then y.viol <- (1-t')x.viol + t' zp.viol. */
static void linterp_xy(struct vec * y, struct vec * x, const struct vec * z,
                       double theta, double next_theta)
{
        assert(theta >= 0 && theta <= 1);
        assert(next_theta >= 0 && next_theta <= 1);
        assert(x->violp && z->violp);
        double s = 1 - theta, ns = 1 - next_theta;
        for (size_t i = 0; i < x->n; i++)
                x->x[i] = s * x->x[i] + theta * z->x[i];
        for (size_t i = 0; i < x->m; i++) {
                double xv = s * x->viol[i] + theta * z->viol[i];
                y->viol[i] = ns * xv + next_theta * z->viol[i];
                x->viol[i] = xv;
        }
        y->value = x->value = nan("");
        y->quad = x->quad = nan("");
}

/* ---- driver (approx/approx.c:280-478) ---- */
struct state {
        struct vec y, z, zp, x;
        size_t iteration;
        double theta;
        struct vec g, g2, u;
        double step_length;
};

static void state_init(struct state * s, size_t n, size_t m)
{
        vec_init(&s->y, n, m); vec_init(&s->z, n, m);
        vec_init(&s->zp, n, m); vec_init(&s->x, n, m);
        s->iteration = 0;
        s->theta = 1;
        vec_init(&s->g, n, m); vec_init(&s->g2, n, m); vec_init(&s->u, n, m);
        s->step_length = 1;
}

static void state_free(struct state * s)
{
        vec_free(&s->y); vec_free(&s->z); vec_free(&s->zp); vec_free(&s->x);
        vec_free(&s->g); vec_free(&s->g2); vec_free(&s->u);
}

/* approx/approx.c:328-339 (default: theta_k = 2/(k+2)) */
static double next_theta(struct state * s)
{
        size_t k = ++s->iteration;
        return s->theta = 2.0 / (k + 2);
}

/* approx/approx.c:341-392 */
static int try_long_step(ora_approx_t * a, struct state * s, double len,
                         int compute_zg)
{
        int safe = (len <= 1 + 1e-6);
        if (safe) len = 1;
        assert(s->z.violp && s->y.violp);
        if (compute_zg) gradient2(&s->g, &s->g2, a, &s->z, &s->y);
        else gradient(&s->g2, a, &s->y);
        double expected = long_step(&s->zp, s->theta, len, &s->g2, &s->z, a);
        double initial = compute_value(a, &s->z);
        double now = compute_value(a, &s->zp);
        assert(expected <= 0);
        if (now <= initial + expected) {
                s->step_length = len * 1.01;
                return 1;
        }
        s->step_length = .9 * len;
        if (!safe) step(&s->zp, s->theta, &s->g2, &s->z, a);
        return 0;
}

/* approx/approx.c:394-415 */
static int short_step(ora_approx_t * a, struct state * s, double len)
{
        assert(s->z.violp && s->y.violp);
        gradient2(&s->g, &s->g2, a, &s->z, &s->y);
        step(&s->zp, s->theta, &s->g2, &s->z, a);
        s->step_length = len * 1.01;
        if (s->step_length > 1) s->step_length = 1 + 1e-6;
        return 0;
}

/* approx/approx.c:417-478; returns the centre (old z, or x on restart). */
static const struct vec * iter(ora_approx_t * a, struct state * s,
                               double * OUT_pg)
{
        int descent, got_gradient;
        double len = s->step_length;
        if (len <= 1) {
                descent = short_step(a, s, len);
                got_gradient = 1;
        } else {
                got_gradient = (OUT_pg != NULL) || (len <= 1 + 1e-6);
                descent = try_long_step(a, s, len, got_gradient);
        }
        if (OUT_pg != NULL)
                *OUT_pg = pg_norm(&s->g, &s->z, a->lower, a->upper);
        if (!descent) {
                if (!got_gradient) gradient(&s->g, a, &s->z);
                if (dot_diff(&s->g, &s->z, &s->zp) > 0) {
                        vec_copy(&s->x, &s->z);
                        vec_copy(&s->y, &s->z);
                        s->iteration = 0;
                        s->theta = 1;
                        return &s->x;
                }
        }
        if (!s->zp.violp) compute_violation(&s->zp, a);
        double theta = s->theta;
        double next = next_theta(s);
        linterp_xy(&s->y, &s->x, &s->zp, theta, next);
        struct vec t = s->z;
        s->z = s->zp;
        s->zp = t;
        return &s->zp;
}

/* approx/approx.c:504-515 */
static void print_log(FILE * log, size_t k, double value, double ng,
                      double pg, double step_len, double diff)
{
        if (log == NULL) return;
        if (diff < HUGE_VAL)
                fprintf(log, "\t%10zu %12g %12g %12g %8g %12g\n",
                        k, value, ng, pg, step_len, diff);
        else
                fprintf(log, "\t%10zu %12g %12g %12g %8g\n",
                        k, value, ng, pg, step_len);
}

/* approx/approx.c:576-708 */
/* Per thread: bench's CPU baseline solves several LPs at once, one per
 * thread. */
static __thread ora_ckpt_hook ckpt_hook;
static __thread void * ckpt_ctx;

void ora_approx_set_ckpt_hook(ora_ckpt_hook fn, void * ctx)
{
        ckpt_hook = fn;
        ckpt_ctx = ctx;
}

int ora_approx_solve(double * xio, size_t n, ora_approx_t * a, size_t niter,
                     double max_pg, double max_value, double min_delta,
                     FILE * log, size_t period, double * diag, double offset)
{
        assert(n == a->nvars);
        if (a->permuted) a = a->permuted;
        struct state s;
        state_init(&s, a->nvars, a->nrhs);
        ora_sparse_matrix_col_permute(a->matrix, s.x.x, n, xio, 1);
        project(&s.x, a->lower, a->upper);
        compute_violation(&s.x, a);
        compute_value(a, &s.x);
        vec_copy(&s.z, &s.x);
        vec_copy(&s.y, &s.x);
        double * prev_x = zalloc(n);
        memcpy(prev_x, s.x.x, n * sizeof(double));

        const struct vec * center = &s.x;
        double value = compute_value(a, (struct vec *)center) + offset;
        double ng = HUGE_VAL, pg = HUGE_VAL, delta = HUGE_VAL;
        size_t i;
        int restart = 0, reason = 0;
        for (i = 0; i < niter; i++) {
                delta = HUGE_VAL;
                pg = HUGE_VAL;
                center = iter(a, &s, ((i == 0) || ((i + 1) % 10 == 0))
                              ? &pg : NULL);
                if (center == &s.x) {
                        if (!restart) {
                                restart = 1;
                                if (log != NULL) fprintf(log, "\t\t ");
                        }
                        if (log != NULL) {
                                fprintf(log, "R");
                                fflush(log);
                        }
                }
                value = compute_value(a, (struct vec *)center) + offset;
                if (value < max_value) { reason = 1; break; }
                if (pg < max_pg) { reason = 2; break; }
                if ((i + 1) % 100 == 0) {
                        center = &s.x;
                        compute_violation(&s.x, a);
                        value = compute_value(a, &s.x) + offset;
                        gradient(&s.g, a, &s.x);
                        pg = pg_norm(&s.g, &s.x, a->lower, a->upper);
                        delta = diff_norm(prev_x, s.x.x, n)
                                / (norm2(&s.x) + 1e-10);
                        if (value < max_value) { reason = 1; break; }
                        if (pg < max_pg) { reason = 2; break; }
                        if (delta < min_delta) { reason = 3; break; }
                        if (ckpt_hook != NULL &&
                            ckpt_hook(ckpt_ctx, a, s.x.x, s.x.viol, s.g.x)) {
                                reason = 2;
                                break;
                        }
                        memcpy(prev_x, s.x.x, n * sizeof(double));
                }
                if ((i == 0) || (period && ((i + 1) % period == 0))) {
                        if (restart) {
                                restart = 0;
                                printf("\n");
                        }
                        ng = norm2(&s.g);
                        print_log(log, i + 1, value, ng, pg, s.step_length,
                                  delta);
                }
        }
        if (restart) printf("\n");

        delta = diff_norm(prev_x, center->x, n) / (norm2(center) + 1e-10);
        value = compute_value(a, (struct vec *)center) + offset;
        gradient(&s.g, a, (struct vec *)center);
        ng = norm2(&s.g);
        pg = pg_norm(&s.g, center, a->lower, a->upper);
        print_log(log, i + 1, value, ng, pg, s.step_length, delta);
        ora_sparse_matrix_col_permute(a->matrix, xio, n, center->x, -1);
        if (diag != NULL) {
                diag[0] = value;
                diag[1] = ng;
                diag[2] = pg;
                diag[3] = delta;
                diag[4] = i + 1;
        }
        free(prev_x);
        state_free(&s);
        return reason;
}

/* approx/mpg.inc:155-199: one spectral projected-gradient step. */
static int spg_iter(ora_approx_t * a, struct state * s, double * OUT_pg)
{
        size_t n = a->nvars;
        project(&s->x, a->lower, a->upper);
        gradient(&s->g, a, &s->x);
        if (OUT_pg != NULL)
                *OUT_pg = pg_norm(&s->g, &s->x, a->lower, a->upper);
        for (size_t i = 0; i < n; i++) {
                s->y.x[i] = s->x.x[i] - s->z.x[i];
                s->u.x[i] = s->g.x[i] - s->g2.x[i];
        }
        double st = dot_pair(s->y.x, s->y.x, n) / dot_pair(s->y.x, s->u.x, n);
        if (isnan(st)) st = 1;
        if (st > 1e8) return 1;
        for (size_t i = 0; i < n; i++) s->z.x[i] = s->x.x[i] - st * s->g.x[i];
        project(&s->z, a->lower, a->upper);
        struct vec t = s->z; s->z = s->x; s->x = t;
        t = s->g2; s->g2 = s->g; s->g = t;
        return 0;
}

/* approx/approx.c:517-573 */
int ora_mpg_solve(double * xio, size_t n, ora_approx_t * a, size_t niter,
                  double max_pg, double max_value, double min_delta,
                  FILE * log, size_t period, double * diag, double offset)
{
        (void)min_delta;
        (void)diag;
        assert(n == a->nvars);
        if (a->permuted) a = a->permuted;
        struct state s;
        state_init(&s, a->nvars, a->nrhs);
        ora_sparse_matrix_col_permute(a->matrix, s.x.x, n, xio, 1);
        project(&s.x, a->lower, a->upper);
        vec_copy(&s.z, &s.x);
        gradient(&s.g, a, &s.x);
        vec_copy(&s.g2, &s.g);
        int reason = 0;
        size_t i;
        for (i = 0; (i < niter) && !reason; i++) {
                double pg;
                spg_iter(a, &s, &pg);
                double value = offset + compute_value(a, &s.x);
                if (pg < max_pg) reason = 2;
                if (value < max_value) reason = 1;
                if (log && ((period && (i % period == 0)) || (i == 0) || reason))
                        fprintf(log, "%10zu: %12g %12g\n", i + 1, pg, value);
        }
        ora_sparse_matrix_col_permute(a->matrix, xio, n, s.x.x, -1);
        if (diag != NULL) {
                gradient(&s.g, a, &s.x);
                diag[0] = offset + compute_value(a, &s.x);
                diag[1] = norm2(&s.g);
                diag[2] = pg_norm(&s.g, &s.x, a->lower, a->upper);
                diag[3] = HUGE_VAL;
                diag[4] = i + 1;
        }
        state_free(&s);
        return 2;
}

