/* ora_alm.c -- oracle restatement of the reference's augmented Lagrangian
 * outer loop.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows alm.c:8-322.  The reference calls cd_solve as its inner solver
 * (alm.c:255); the inner solver here is a parameter, APPROX by default,
 * which is what the reference computes when alm.c is compiled with
 * -Dcd_solve=approx_solve.  The stopping constants are parameters whose
 * defaults are the reference's.
 */
#include "ora_internal.h"
#include <assert.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct ora_alm {
        size_t nrhs, nvars;
        ora_matrix_t * matrix;
        double * linear;
        ora_approx_t * approx;
        double * lambda_lower, * lambda_upper;
};

static double * copy_default(const double * x, size_t n, double missing)
{
        double * out = calloc(n + 1, sizeof(double));
        for (size_t i = 0; i < n; i++) out[i] = x ? x[i] : missing;
        return out;
}

void ora_alm_default_options(struct ora_alm_options * o)
{
        o->stop_violation = 1e-5;
        o->stop_pg = 1e-5;
        o->precision_floor = 1e-7;
        o->min_delta = 1e-11;
        o->inner = ORA_INNER_APPROX;
        o->stop_kkt = 0;
}

/* alm.c:59-83 */
ora_alm_t * ora_alm_make(ora_matrix_t * m, size_t nrhs, const double * rhs,
                         size_t nvars, const double * linear,
                         const double * lower, const double * upper,
                         const double * lambda_lower,
                         const double * lambda_upper)
{
        ora_alm_t * a = calloc(1, sizeof(*a));
        a->nrhs = nrhs;
        a->nvars = nvars;
        a->matrix = m;
        a->linear = copy_default(linear, nvars, 0);
        a->approx = ora_approx_make(m, nrhs, rhs, NULL, nvars, NULL,
                                    lower, upper);
        a->lambda_lower = copy_default(lambda_lower, nrhs, -HUGE_VAL);
        a->lambda_upper = copy_default(lambda_upper, nrhs, HUGE_VAL);
        return a;
}

int ora_alm_free(ora_alm_t * a)
{
        if (a == NULL) return 0;
        free(a->linear);
        free(a->lambda_lower);
        free(a->lambda_upper);
        ora_approx_free(a->approx);
        free(a);
        return 0;
}

/* alm.c:23-43 */
double * ora_alm_linear(ora_alm_t * a) { return a->linear; }
ora_approx_t * ora_alm_approx(ora_alm_t * a) { return a->approx; }

/* alm.c:98-104: plain sequential sum. */
static double seq_dot(const double * x, const double * y, size_t n)
{
        double acc = 0;
        for (size_t i = 0; i < n; i++) acc += x[i] * y[i];
        return acc;
}

/* alm.c:106-119: c' = c - A' lambda; returns lambda . b */
static double penalise_linear(ora_alm_t * a, const double * lambda)
{
        double * c = a->approx->linear;
        ora_sparse_matrix_multiply(c, a->nvars, a->matrix, lambda, a->nrhs, 1);
        for (size_t i = 0; i < a->nvars; i++) c[i] = a->linear[i] - c[i];
        return seq_dot(lambda, a->approx->rhs, a->nrhs);
}

/* alm.c:121-132 */
static void violation(double * r, ora_alm_t * a, const double * x)
{
        ora_sparse_matrix_multiply(r, a->nrhs, a->matrix, x, a->nvars, 0);
        for (size_t i = 0; i < a->nrhs; i++) r[i] -= a->approx->rhs[i];
}

/* alm.c:134-169 */
static double seq_norm2(const double * x, size_t n)
{
        double acc = 0;
        for (size_t i = 0; i < n; i++) acc += x[i] * x[i];
        return sqrt(acc);
}

static double seq_norm1(const double * x, size_t n)
{
        double acc = 0;
        for (size_t i = 0; i < n; i++) acc += fabs(x[i]);
        return acc;
}

static double norm_inf(const double * x, size_t n, size_t * OUT_i)
{
        double mx = -HUGE_VAL;
        size_t at = (size_t)-1;
        for (size_t i = 0; i < n; i++) {
                double v = fabs(x[i]);
                if (v > mx) { mx = v; at = i; }
        }
        if (OUT_i != NULL) *OUT_i = at;
        return mx;
}

/* alm.c:171-192 */
static void project_multipliers(double * l, ora_alm_t * a)
{
        for (size_t i = 0; i < a->nrhs; i++)
                l[i] = fmax(a->lambda_lower[i], fmin(l[i], a->lambda_upper[i]));
}

static void update_multipliers(double * l, ora_alm_t * a, const double * r)
{
        const double * w = a->approx->weight;
        for (size_t i = 0; i < a->nrhs; i++) l[i] -= w[i] * r[i];
        project_multipliers(l, a);
}

/* alm.c:194-212 */
static double update_weights(ora_alm_t * a, const double * r, double prev)
{
        double norm = seq_norm2(r, a->nrhs);
        if (norm < .5 * prev) return norm;
        double * w = a->approx->weight;
        double scale = 2 * fmin(1, norm / prev);
        for (size_t i = 0; i < a->nrhs; i++) w[i] *= scale;
        size_t at;
        double mx = norm_inf(r, a->nrhs, &at);
        if (mx > .5 * norm) w[at] *= 10;
        return norm;
}

struct alm_state { double * viol; double prev_norm, precision; };

/* alm.c:235-245 */
static void update_precision(struct alm_state * s, ora_alm_t * a,
                             const double * r, double floor_)
{
        double viol = norm_inf(r, a->nrhs, NULL);
        double w = a->nrhs / seq_norm1(a->approx->weight, a->nrhs);
        s->precision = fmin(s->precision, fmin(pow(viol, 1), pow(w, 1.5)));
        s->precision = fmax(s->precision, floor_);
}

/* alm.c:247-287 */
static int iter(struct alm_state * s, ora_alm_t * a, double * x,
                double * lambda, FILE * log, size_t k, double * OUT_pg,
                double * OUT_max_viol, const struct ora_alm_options * o)
{
        double offset = penalise_linear(a, lambda);
        double diag[5];
        ora_approx_update(a->approx);
        int reason;
        if (o->inner == ORA_INNER_CD)
                reason = ora_cd_solve(x, a->nvars, a->approx, -1u,
                                      fmin(1, s->precision), -HUGE_VAL,
                                      o->min_delta, log, 10000, diag, offset);
        else
                reason = ora_approx_solve(x, a->nvars, a->approx, -1u,
                                          fmin(1, s->precision), -HUGE_VAL,
                                          o->min_delta, log, 10000, diag,
                                          offset);
        violation(s->viol, a, x);
        double max_viol = norm_inf(s->viol, a->nrhs, NULL);
        if (log != NULL)
                fprintf(log, "%4zu: %12g %12g %12g %.18g %.18g", k, max_viol,
                        seq_norm2(s->viol, a->nrhs), diag[2], diag[0],
                        seq_dot(x, a->linear, a->nvars));
        if (OUT_pg != NULL) *OUT_pg = diag[2];
        if (OUT_max_viol != NULL) *OUT_max_viol = max_viol;
        update_multipliers(lambda, a, s->viol);
        if (reason != 2) {
                if (log != NULL) fprintf(log, " -stalled-\n");
                update_precision(s, a, s->viol, o->precision_floor);
                return 1;
        }
        s->prev_norm = update_weights(a, s->viol, s->prev_norm);
        update_precision(s, a, s->viol, o->precision_floor);
        if (log != NULL)
                fprintf(log, " (%8g)\n",
                        seq_norm1(a->approx->weight, a->nrhs) / a->nrhs);
        return 0;
}

/* The device library's relative KKT residual (oracle.h), in plain
 * sequential sums. */
double ora_alm_kkt(ora_alm_t * a, const double * x, const double * lambda)
{
        const size_t m = a->nrhs, n = a->nvars;
        const double * b = a->approx->rhs, * c = a->linear;
        const double * lo = a->approx->lower, * up = a->approx->upper;
        double * r = calloc(m + 1, sizeof(double));
        double * atl = calloc(n + 1, sizeof(double));
        ora_sparse_matrix_multiply(r, m, a->matrix, x, n, 0);
        ora_sparse_matrix_multiply(atl, n, a->matrix, lambda, m, 1);
        double rr = 0, bb = 0, bl = 0;
        for (size_t i = 0; i < m; i++) {
                const double ri = r[i] - b[i];
                rr += ri * ri;
                bb += b[i] * b[i];
                bl += b[i] * lambda[i];
        }
        double px = 0, inf2 = 0, bnd = 0, cc = 0, cx = 0;
        for (size_t j = 0; j < n; j++) {
                const double d = c[j] - atl[j];
                const double pr = fmin(fmax(x[j] - d, lo[j]), up[j]);
                const double dp = fmax(d, 0), dn = fmax(-d, 0);
                const int lo_inf = !isfinite(lo[j]), up_inf = !isfinite(up[j]);
                const double inf = (lo_inf ? dp : 0) + (up_inf ? dn : 0);
                px += (x[j] - pr) * (x[j] - pr);
                inf2 += inf * inf;
                bnd += (lo_inf ? 0 : lo[j] * dp) - (up_inf ? 0 : up[j] * dn);
                cc += c[j] * c[j];
                cx += c[j] * x[j];
        }
        free(r);
        free(atl);
        const double dual = bl + bnd;
        const double rp = sqrt(rr) / (1 + sqrt(bb));
        const double rd = (sqrt(px) + sqrt(inf2)) / (1 + sqrt(cc));
        const double rg = fabs(cx - dual) / (1 + fabs(cx) + fabs(dual));
        return fmax(rp, fmax(rd, rg));
}

/* The device library's inner-checkpoint test for its KKT target (the
 * OP_KKT_CK op and its continuation): the relative KKT residual of
 * (x, lambda'), lambda' = lambda - w (Ax - b), from the checkpoint's
 * violation and gradient g = c - A' lambda' (no product needed), valid
 * while lambda' needs no projection onto the multiplier bounds.  Sums of
 * c^2 and b^2 are taken once per solve, in the caller's order. */
struct kkt_ckpt {
        ora_alm_t * alm;
        const double * lambda;      /* caller order, current */
        double stop, cc, bb;
        double * c, * lam, * lam_lo, * lam_hi;  /* working order */
};

static int kkt_checkpoint(void * ctx, const struct ora_approx * ap,
                          const double * x, const double * viol,
                          const double * g)
{
        struct kkt_ckpt * k = ctx;
        const size_t n = ap->nvars, m = ap->nrhs;
        ora_sparse_matrix_row_permute(ap->matrix, k->lam, m, k->lambda, 1);
        double px = 0, inf2 = 0, bnd = 0, cx = 0;
        for (size_t j = 0; j < n; j++) {
                const double d = g[j], lo = ap->lower[j], up = ap->upper[j];
                const double pr = fmin(fmax(x[j] - d, lo), up);
                const double dp = fmax(d, 0), dn = fmax(-d, 0);
                const int lo_inf = !isfinite(lo), up_inf = !isfinite(up);
                const double inf = (lo_inf ? dp : 0) + (up_inf ? dn : 0);
                px += (x[j] - pr) * (x[j] - pr);
                inf2 += inf * inf;
                bnd += (lo_inf ? 0 : lo * dp) - (up_inf ? 0 : up * dn);
                cx += k->c[j] * x[j];
        }
        double bl = 0, rr = 0, proj = 0;
        for (size_t i = 0; i < m; i++) {
                const double r = viol[i] / ap->weight[i];
                const double l1 = k->lam[i] - viol[i];
                const double l2 = fmax(k->lam_lo[i], fmin(l1, k->lam_hi[i]));
                bl += ap->rhs[i] * l2;
                rr += r * r;
                proj = fmax(proj, fabs(l2 - l1));
        }
        const double dual = bl + bnd;
        const double rp = sqrt(rr) / (1 + sqrt(k->bb));
        const double rd = (sqrt(px) + sqrt(inf2)) / (1 + sqrt(k->cc));
        const double rg = fabs(cx - dual) / (1 + fabs(cx) + fabs(dual));
        return proj == 0 && fmax(rp, fmax(rd, rg)) <= k->stop;
}

/* alm.c:289-322 */
int ora_alm_solve(ora_alm_t * a, size_t niter, double * x, size_t nvars,
                  double * lambda, size_t ncons, FILE * log, double * diag,
                  const struct ora_alm_options * opt)
{
        struct ora_alm_options o;
        if (opt) o = *opt; else ora_alm_default_options(&o);
        assert(nvars == a->nvars);
        assert(ncons == a->nrhs);
        struct alm_state s;
        s.viol = calloc(a->nrhs + 1, sizeof(double));
        s.prev_norm = HUGE_VAL;
        s.precision = 1;
        project_multipliers(lambda, a);
        double pg = HUGE_VAL, max_viol = HUGE_VAL;
        int ret = 1;
        struct kkt_ckpt ck = {a, lambda, o.stop_kkt, 0, 0, NULL, NULL, NULL,
                              NULL};
        if (o.stop_kkt > 0) {
                const ora_approx_t * ap = a->approx->permuted
                        ? a->approx->permuted : a->approx;
                const size_t n = a->nvars, m = a->nrhs;
                for (size_t j = 0; j < n; j++)
                        ck.cc += a->linear[j] * a->linear[j];
                for (size_t i = 0; i < m; i++)
                        ck.bb += a->approx->rhs[i] * a->approx->rhs[i];
                ck.c = calloc(n + 1, sizeof(double));
                ck.lam = calloc(m + 1, sizeof(double));
                ck.lam_lo = calloc(m + 1, sizeof(double));
                ck.lam_hi = calloc(m + 1, sizeof(double));
                ora_sparse_matrix_col_permute(ap->matrix, ck.c, n, a->linear, 1);
                ora_sparse_matrix_row_permute(ap->matrix, ck.lam_lo, m,
                                              a->lambda_lower, 1);
                ora_sparse_matrix_row_permute(ap->matrix, ck.lam_hi, m,
                                              a->lambda_upper, 1);
                ora_approx_set_ckpt_hook(kkt_checkpoint, &ck);
        }
        for (size_t i = 0; i < niter; i++) {
                int status = iter(&s, a, x, lambda, log, i + 1, &pg,
                                  &max_viol, &o);
                if (o.stop_kkt > 0) {
                        if (ora_alm_kkt(a, x, lambda) <= o.stop_kkt) {
                                ret = 0;
                                break;
                        }
                        continue;
                }
                if ((max_viol < o.stop_violation) && (status || (pg < o.stop_pg))) {
                        ret = 0;
                        break;
                }
        }
        if (diag != NULL) {
                diag[0] = seq_dot(x, a->linear, a->nvars);
                diag[1] = max_viol;
                diag[2] = pg;
        }
        if (o.stop_kkt > 0) {
                ora_approx_set_ckpt_hook(NULL, NULL);
                free(ck.c);
                free(ck.lam);
                free(ck.lam_lo);
                free(ck.lam_hi);
        }
        free(s.viol);
        return ret;
}
