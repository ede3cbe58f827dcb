/* drop_in.c -- a C program written against the reference's interface
 * (alm.h, approx/approx.h, spmv/spmv.h) and linked with -lacdalm in place
 * of the reference's objects.  Its three checks follow the reference's own
 * test mains, with assertions where those only print:
 *
 *   spmv    TEST_SPMV   (spmv/spmv.c:261-386): every shape in [0,20)^2,
 *           random dense +-1 matrices at density 1/2, A x and A'x (and the
 *           two-vector form) against the dense products.  The reference
 *           accepts a 2-norm error below 1e-4; the sums here are in the
 *           reference's order, so they must agree exactly.
 *   approx  TEST_APPROX (approx/approx.c:710-798): rows, columns in
 *           [10,20], random matrix at density 1/2 plus a diagonal,
 *           rhs = A sol, NULL weight / cost / bounds; cd_solve with
 *           max_value 1e-13 must return > 0 with |Ax - b| < 1e-6 on the
 *           shapes whose row and column counts differ by at least 3 (the
 *           square-ish ones can be ill-conditioned enough to take
 *           minutes); approx_solve must do the same on a fifth of the
 *           shapes with the diagonal made dominant.
 *   alm     TEST_ALM    (alm.c:366-395): an LP read with alm_read from the
 *           text format, solved by alm_solve with a log; the result must
 *           be the known optimum; and a random feasible LP whose KKT
 *           residual must be small.
 *
 * Exit status 0 and "drop_in: ok" when every check passes. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "alm.h"
#include "approx/approx.h"
#include "spmv/spmv.h"

static int failures;

#define CHECK(cond, ...)                                               \
        do {                                                           \
                if (!(cond)) {                                         \
                        fprintf(stderr, "FAIL %s:%d: ", __FILE__,      \
                                __LINE__);                             \
                        fprintf(stderr, __VA_ARGS__);                  \
                        fputc('\n', stderr);                           \
                        failures++;                                    \
                }                                                      \
        } while (0)

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static uint64_t next_u64(void)
{
        uint64_t x = rng_state;
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return rng_state = x;
}

static double uniform(double lo, double hi)
{
        return lo + (hi - lo) * ((next_u64() >> 11) * 0x1.0p-53);
}

/* COO of a dense row-major matrix's nonzeros, in row-major order. */
static size_t to_coo(const double * dense, size_t nrows, size_t ncols,
                     uint32_t * rows, uint32_t * cols, double * vals)
{
        size_t nnz = 0;
        for (size_t i = 0; i < nrows; i++)
                for (size_t j = 0; j < ncols; j++)
                        if (dense[i * ncols + j] != 0) {
                                rows[nnz] = (uint32_t)i;
                                cols[nnz] = (uint32_t)j;
                                vals[nnz] = dense[i * ncols + j];
                                nnz++;
                        }
        return nnz;
}

/* y = A x (or A'x) summed in ascending index order from +0.0. */
static void dense_mult(double * y, const double * dense, size_t nrows,
                       size_t ncols, const double * x, int transpose)
{
        size_t ny = transpose ? ncols : nrows, nx = transpose ? nrows : ncols;
        for (size_t i = 0; i < ny; i++) {
                double acc = 0.0;
                for (size_t j = 0; j < nx; j++) {
                        double a = transpose ? dense[j * ncols + i]
                                             : dense[i * ncols + j];
                        if (a != 0) acc = acc + a * x[j];
                }
                y[i] = acc;
        }
}

static void check_spmv(void)
{
        enum { N = 20 };
        double dense[N * N], x[N], x2[N], y[N], y2[N], ref[N], ref2[N];
        uint32_t rows[N * N], cols[N * N];
        double vals[N * N];
        size_t products = 0;
        for (size_t nrows = 0; nrows < N; nrows++)
                for (size_t ncols = 0; ncols < N; ncols++) {
                        for (size_t k = 0; k < nrows * ncols; k++)
                                dense[k] = (next_u64() & 1)
                                                   ? ((next_u64() & 1) ? 1.0 : -1.0)
                                                   : 0.0;
                        size_t nnz = to_coo(dense, nrows, ncols, rows, cols, vals);
                        sparse_matrix_t * m = sparse_matrix_make(
                                ncols, nrows, nnz, rows, cols, vals, 0);
                        CHECK(m != NULL, "sparse_matrix_make %zux%zu", nrows, ncols);
                        if (m == NULL) continue;
                        CHECK(sparse_matrix_nrows(m) == nrows
                              && sparse_matrix_ncolumns(m) == ncols
                              && sparse_matrix_nnz(m) == nnz, "shape %zux%zu",
                              nrows, ncols);
                        for (int rep = 0; rep < 3; rep++)
                                for (int tr = 0; tr < 2; tr++) {
                                        size_t ny = tr ? ncols : nrows;
                                        size_t nx = tr ? nrows : ncols;
                                        for (size_t j = 0; j < nx; j++) {
                                                x[j] = uniform(-1, 1);
                                                x2[j] = uniform(-1, 1);
                                        }
                                        dense_mult(ref, dense, nrows, ncols, x, tr);
                                        dense_mult(ref2, dense, nrows, ncols, x2, tr);
                                        CHECK(sparse_matrix_multiply(y, ny, m, x, nx,
                                                                     tr, NULL) == 0,
                                              "multiply");
                                        CHECK(memcmp(y, ref, ny * sizeof(double)) == 0,
                                              "A%sx %zux%zu differs", tr ? "'" : "",
                                              nrows, ncols);
                                        double * ys[2] = {y, y2};
                                        const double * xs[2] = {x, x2};
                                        CHECK(sparse_matrix_multiply_2(ys, ny, m, xs, nx,
                                                                       tr, NULL) == 0,
                                              "multiply_2");
                                        CHECK(memcmp(y, ref, ny * sizeof(double)) == 0
                                              && memcmp(y2, ref2, ny * sizeof(double)) == 0,
                                              "two-vector A%sx %zux%zu differs",
                                              tr ? "'" : "", nrows, ncols);
                                        products++;
                                }
                        sparse_matrix_free(m);
                }
        printf("spmv: %zu products checked\n", products);
}

/* |b - A x| with the library's own product, as the reference test does */
static double residual(sparse_matrix_t * m, size_t nrows, size_t ncols,
                       const double * x, const double * b)
{
        double ax[64], acc = 0;
        CHECK(sparse_matrix_multiply(ax, nrows, m, x, ncols, 0, NULL) == 0,
              "multiply");
        for (size_t i = 0; i < nrows; i++)
                acc += (ax[i] - b[i]) * (ax[i] - b[i]);
        return sqrt(acc);
}

/* Density 1/2 plus the diagonal, values U[-1, 1] (the diagonal 10 when
 * `dominant`), as TEST_APPROX's random_matrix. */
static sparse_matrix_t * random_system(size_t nrows, size_t ncols,
                                       int dominant)
{
        static uint32_t rows[64 * 64], cols[64 * 64];
        static double vals[64 * 64];
        size_t nnz = 0;
        for (size_t i = 0; i < nrows; i++)
                for (size_t j = 0; j < ncols; j++) {
                        if ((next_u64() & 1) && i != j) continue;
                        rows[nnz] = (uint32_t)i;
                        cols[nnz] = (uint32_t)j;
                        vals[nnz] = dominant && i == j ? 10.0 : uniform(-1, 1);
                        nnz++;
                }
        return sparse_matrix_make(ncols, nrows, nnz, rows, cols, vals, 0);
}

static void check_approx(void)
{
        double sol[64], rhs[64], x[64], diag[5];
        int systems = 0;
        for (size_t nrows = 10; nrows <= 20; nrows++)
                for (size_t ncols = 10; ncols <= 20; ncols++)
                        for (int dominant = 0; dominant < 2; dominant++) {
                                /* approx_solve on every fifth shape; cd_solve
                                 * where the row and column counts differ by 3
                                 * or more (square-ish random systems can be
                                 * so ill-conditioned that the reference build
                                 * itself takes minutes: 187 s on one 20x20) */
                                if (dominant && (nrows + ncols) % 5) continue;
                                if (!dominant && (nrows > ncols ? nrows - ncols
                                                                : ncols - nrows) < 3)
                                        continue;
                                sparse_matrix_t * m = random_system(nrows, ncols,
                                                                    dominant);
                                for (size_t j = 0; j < ncols; j++) {
                                        sol[j] = uniform(-1, 1);
                                        x[j] = 0;
                                }
                                CHECK(sparse_matrix_multiply(rhs, nrows, m, sol, ncols,
                                                             0, NULL) == 0, "rhs");
                                /* NULL weight, cost and bounds: 1, 0, (-inf, inf) */
                                approx_t * ap = approx_make(m, nrows, rhs, NULL, ncols,
                                                            NULL, NULL, NULL);
                                CHECK(ap != NULL, "approx_make");
                                if (ap == NULL) continue;
                                int r = dominant
                                        ? approx_solve(x, ncols, ap, 1000000, 0, 1e-13,
                                                       0, NULL, 0, diag, 0, NULL)
                                        : cd_solve(x, ncols, ap, -1U, 0, 1e-13, 0,
                                                   NULL, 0, diag, 0, NULL);
                                double res = residual(m, nrows, ncols, x, rhs);
                                CHECK(r > 0, "%s %zux%zu returned %d",
                                      dominant ? "approx_solve" : "cd_solve", nrows,
                                      ncols, r);
                                CHECK(res < 1e-6, "%s %zux%zu residual %g",
                                      dominant ? "approx_solve" : "cd_solve", nrows,
                                      ncols, res);
                                approx_free(ap);
                                sparse_matrix_free(m);
                                systems++;
                        }
        printf("approx: %d systems solved\n", systems);
}

/* Relative KKT residual of (x, lambda) for min c'x, Ax = b, l <= x <= u
 * with free multipliers: primal, dual (reduced costs against the bounds)
 * and gap terms, each over 1 + the norm of its data. */
static double kkt(const double * dense, size_t m, size_t n, const double * b,
                  const double * c, const double * l, const double * u,
                  const double * x, const double * lam)
{
        double rp = 0, nb = 0, rd = 0, nc = 0, pobj = 0, dobj = 0;
        for (size_t i = 0; i < m; i++) {
                double r = -b[i];
                for (size_t j = 0; j < n; j++) r += dense[i * n + j] * x[j];
                rp += r * r;
                nb += b[i] * b[i];
                dobj += b[i] * lam[i];
        }
        for (size_t j = 0; j < n; j++) {
                double d = c[j];
                for (size_t i = 0; i < m; i++) d -= dense[i * n + j] * lam[i];
                /* d splits into a part for the lower and one for the upper
                 * bound; what cannot be attributed is dual infeasibility */
                double viol = 0;
                if (d > 0) {
                        if (isinf(l[j])) viol = d;
                        else dobj += d * l[j];
                } else if (d < 0) {
                        if (isinf(u[j])) viol = -d;
                        else dobj += d * u[j];
                }
                rd += viol * viol;
                nc += c[j] * c[j];
                pobj += c[j] * x[j];
        }
        double p = sqrt(rp) / (1 + sqrt(nb)), q = sqrt(rd) / (1 + sqrt(nc));
        double g = fabs(pobj - dobj) / (1 + fabs(pobj) + fabs(dobj));
        return fmax(p, fmax(q, g));
}

static void check_alm(void)
{
        /* min x1 + 2 x2  s.t.  x1 + x2 = 1,  0 <= x <= 10:  x* = (1, 0) */
        FILE * f = tmpfile();
        CHECK(f != NULL, "tmpfile");
        if (f == NULL) return;
        fprintf(f, "1 2 2\n0 0 1\n0 1 1\n1\n1 2\n0 0\n10 10\n-inf\ninf\n");
        rewind(f);
        alm_t * alm = alm_read(f);
        fclose(f);
        CHECK(alm != NULL, "alm_read");
        if (alm != NULL) {
                double x[2] = {0, 0}, lam[1] = {0}, diag[3];
                FILE * log = tmpfile();
                int r = alm_solve(alm, 1000, x, 2, lam, 1, log, diag, NULL);
                long logged = log ? ftell(log) : 0;
                if (log) fclose(log);
                CHECK(r == 0, "alm_solve returned %d", r);
                /* the reference's own build stops at x1 = 0.999996 */
                CHECK(fabs(x[0] - 1) < 1e-4 && fabs(x[1]) < 1e-4,
                      "x = (%g, %g)", x[0], x[1]);
                CHECK(fabs(diag[0] - 1) < 1e-4, "objective %g", diag[0]);
                CHECK(logged > 0, "no log lines");
                alm_free(alm);
        }
        /* a random feasible LP: 30 x 80, density 1/4, b = A x0 */
        enum { M = 30, N = 80 };
        static double dense[M * N];
        static uint32_t rows[M * N], cols[M * N];
        static double vals[M * N];
        double b[M], c[N], l[N], u[N], llo[M], lhi[M], x0[N], x[N], lam[M];
        for (size_t k = 0; k < M * N; k++)
                dense[k] = (next_u64() & 3) == 0 ? uniform(0.5, 2) *
                                (next_u64() & 1 ? 1 : -1) : 0;
        for (size_t j = 0; j < N; j++) {
                x0[j] = uniform(0, 10);
                c[j] = uniform(-1, 1);
                l[j] = 0;
                u[j] = 10;
                x[j] = 0;
        }
        for (size_t i = 0; i < M; i++) {
                double acc = 0;
                for (size_t j = 0; j < N; j++) acc += dense[i * N + j] * x0[j];
                b[i] = acc;
                llo[i] = -HUGE_VAL;
                lhi[i] = HUGE_VAL;
                lam[i] = 0;
        }
        size_t nnz = to_coo(dense, M, N, rows, cols, vals);
        sparse_matrix_t * m = sparse_matrix_make(N, M, nnz, rows, cols, vals, 0);
        alm = alm_make(m, M, b, N, c, l, u, llo, lhi);
        CHECK(alm != NULL, "alm_make");
        if (alm != NULL) {
                double diag[3];
                int r = alm_solve(alm, 1000, x, N, lam, M, NULL, diag, NULL);
                double k = kkt(dense, M, N, b, c, l, u, x, lam);
                CHECK(r == 0, "random LP: alm_solve returned %d", r);
                CHECK(k < 1e-5, "random LP: KKT residual %g", k);
                printf("alm: random LP objective %.9g, KKT %.3g\n", diag[0], k);
                alm_free(alm);
        }
        sparse_matrix_free(m);
}

int main(int argc, char ** argv)
{
        const char * only = argc > 1 ? argv[1] : "";
        sparse_matrix_init();
        if (!*only || !strcmp(only, "spmv")) check_spmv();
        if (!*only || !strcmp(only, "approx")) check_approx();
        if (!*only || !strcmp(only, "alm")) check_alm();
        if (failures) {
                printf("drop_in: %d failures\n", failures);
                return 1;
        }
        printf("drop_in: ok\n");
        return 0;
}
