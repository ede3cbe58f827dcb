/* ref_harness.c -- command-line driver around a library that exports the
 * original acd-alm C interface (alm.h / spmv.h / thread_pool.h).
 *
 * TEST INFRASTRUCTURE ONLY.  Linked against the reference compiled from
 * its own sources (oracle/Makefile -> oracle/_ref/ref_bench) it times the
 * reference's thread_pool CPU path for the cpu_baseline figure of
 * bench.py.  It calls the public interface only.
 *
 * usage: ref_bench INSTANCE.lpb THREADS MODE [REPEAT]
 *   MODE alm  : alm_solve(niter=1000) from x = 0, lambda = 0
 *   MODE spmv : REPEAT x (A x, then A' y)
 *   MODE approx : REPEAT APPROX iterations of the first ALM subproblem
 *                 (lambda = 0, unit weights, the instance's cost), from
 *                 x = 0, with no stopping test; the bounded CPU sample
 * Prints one JSON line.  The instance format (little endian):
 *   "ACDLP001", u64 nrows, u64 ncolumns, u64 nnz,
 *   u32 rows[nnz], u32 columns[nnz], f64 values[nnz],
 *   f64 rhs[nrows], linear[ncolumns], lower[ncolumns], upper[ncolumns],
 *   lambda_lower[nrows], lambda_upper[nrows]
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <math.h>
#include "alm.h"
#include "approx/approx.h"

static double now_s(void)
{
        struct timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void * read_n(FILE * f, size_t n, size_t size)
{
        void * p = calloc(n + 2, size);
        if (n && fread(p, size, n, f) != n) {
                fprintf(stderr, "short read\n");
                exit(2);
        }
        return p;
}

int main(int argc, char ** argv)
{
        if (argc < 4) {
                fprintf(stderr, "usage: %s INSTANCE THREADS MODE [REPEAT]\n",
                        argv[0]);
                return 2;
        }
        FILE * f = fopen(argv[1], "rb");
        if (f == NULL) { perror("open"); return 2; }
        char magic[8];
        uint64_t dims[3];
        if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "ACDLP001", 8)
            || fread(dims, 8, 3, f) != 3) {
                fprintf(stderr, "bad instance header\n");
                return 2;
        }
        size_t m = dims[0], n = dims[1], nnz = dims[2];
        uint32_t * rows = read_n(f, nnz, 4), * cols = read_n(f, nnz, 4);
        double * vals = read_n(f, nnz, 8);
        double * rhs = read_n(f, m, 8), * lin = read_n(f, n, 8);
        double * lo = read_n(f, n, 8), * hi = read_n(f, n, 8);
        double * llo = read_n(f, m, 8), * lhi = read_n(f, m, 8);
        fclose(f);

        int threads = atoi(argv[2]);
        const char * mode = argv[3];
        int repeat = argc > 4 ? atoi(argv[4]) : 10;
        sparse_matrix_init();
        thread_pool_t * pool = threads > 1 ? thread_pool_init(threads) : NULL;

        double t0 = now_s();
        sparse_matrix_t * a = sparse_matrix_make(n, m, nnz, rows, cols, vals, 0);
        double t_make = now_s() - t0;

        if (!strcmp(mode, "spmv")) {
                double * x = calloc(n + 2, 8), * y = calloc(m + 2, 8);
                for (size_t i = 0; i < n; i++) x[i] = 1.0 / (1 + i % 7);
                for (size_t i = 0; i < m; i++) y[i] = 1.0 / (1 + i % 5);
                double * ox = calloc(m + 2, 8), * oy = calloc(n + 2, 8);
                sparse_matrix_multiply(ox, m, a, x, n, 0, pool);
                sparse_matrix_multiply(oy, n, a, y, m, 1, pool);
                double ta = 0, tt = 0;
                for (int r = 0; r < repeat; r++) {
                        double s = now_s();
                        sparse_matrix_multiply(ox, m, a, x, n, 0, pool);
                        ta += now_s() - s;
                        s = now_s();
                        sparse_matrix_multiply(oy, n, a, y, m, 1, pool);
                        tt += now_s() - s;
                }
                printf("{\"mode\": \"spmv\", \"threads\": %d, \"repeat\": %d, "
                       "\"s_per_a\": %.9g, \"s_per_at\": %.9g, "
                       "\"make_s\": %.6g, \"checksum\": %.17g}\n",
                       threads, repeat, ta / repeat, tt / repeat, t_make,
                       ox[0] + oy[0]);
        } else if (!strcmp(mode, "approx")) {
                double * w = calloc(m + 2, 8), * x = calloc(n + 2, 8);
                for (size_t i = 0; i < m; i++) w[i] = 1;
                approx_t * ap = approx_make(a, m, rhs, w, n, lin, lo, hi);
                double s = now_s();
                approx_update(ap);
                double t_update = now_s() - s;
                double diag[5];
                s = now_s();
                int ret = approx_solve(x, n, ap, (size_t)repeat, 0.0,
                                       -HUGE_VAL, 0.0, NULL, 0, diag, 0.0,
                                       pool);
                double t = now_s() - s;
                printf("{\"mode\": \"approx\", \"threads\": %d, "
                       "\"ret\": %d, \"iterations\": %d, "
                       "\"s_per_iteration\": %.9g, \"update_s\": %.6g, "
                       "\"make_s\": %.6g, \"objective\": %.17g}\n",
                       threads, ret, repeat, t / repeat, t_update, t_make,
                       diag[0]);
                approx_free(ap);
        } else {
                alm_t * alm = alm_make(a, m, rhs, n, lin, lo, hi, llo, lhi);
                double * x = calloc(n + 2, 8), * lam = calloc(m + 2, 8);
                double diag[3];
                double s = now_s();
                int ret = alm_solve(alm, 1000, x, n, lam, m, NULL, diag, pool);
                double t = now_s() - s;
                printf("{\"mode\": \"alm\", \"threads\": %d, \"ret\": %d, "
                       "\"solve_s\": %.9g, \"make_s\": %.6g, "
                       "\"objective\": %.17g, \"max_viol\": %.17g, "
                       "\"pg\": %.17g}\n",
                       threads, ret, t, t_make, diag[0], diag[1], diag[2]);
                alm_free(alm);
        }
        if (pool) thread_pool_free(pool);
        sparse_matrix_free(a);
        return 0;
}
