/* ora_bench.c -- times the oracle (this repository's C restatement of the
 * reference's CPU path) on host cores, for the cpu_baseline figures of
 * bench.py.  TEST INFRASTRUCTURE ONLY (see oracle.h): it measures the
 * checker; the product never runs it.
 *
 * usage:
 *   ora_bench alm THREADS STOP_KKT FILE...
 *       LP-parallel ALM solves (APPROX inner solver, niter = 1000, from
 *       x = 0, lambda = 0), one LP per worker at a time, THREADS pthreads
 *       taking the files in order from a shared counter.  The reference's
 *       thread pool runs one job at a time (thread_pool/thread_pool.c:
 *       102-103), so a batch of independent LPs is spread over cores by
 *       solving several LPs at once with pool = NULL each.  STOP_KKT > 0
 *       stops on the relative KKT residual (ora_alm_kkt), 0 on the
 *       reference's absolute rule (alm.c:303-319).
 *   ora_bench spmv THREADS REPEAT FILE
 *       y = A x and y = A' x over the oracle's CSR arrays, rows split in
 *       chunks of 16 handed out dynamically to THREADS pthreads (the
 *       granularity of spmv/spmv.c:159-193's thread_pool_for); each row is
 *       summed sequentially in column order as in the reference
 *       (spmv/spmv_block.c:160-194).  Mean seconds per product over REPEAT
 *       products after one warm-up product.
 *   ora_bench approx ITERS FILE
 *       ITERS serial APPROX iterations (approx/approx.c:576-708) of the
 *       first ALM subproblem (lambda = 0, unit weights, the instance's
 *       cost) from x = 0 with no stopping test: seconds per iteration.
 * Prints one JSON line.  Instance files are "ACDLP001" records (the
 * format of acdalm.instances.LP.write_binary).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

static double now_s(void)
{
        struct timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

struct lp {
        uint64_t m, n, nnz;
        uint32_t * rows, * cols;
        double * vals, * rhs, * linear, * lower, * upper, * llo, * lhi;
};

static void * read_n(FILE * f, size_t n, size_t size)
{
        void * p = calloc(n + 2, size);
        if (p == NULL || (n && fread(p, size, n, f) != n)) {
                fprintf(stderr, "ora_bench: short read\n");
                exit(2);
        }
        return p;
}

static void read_lp(const char * path, struct lp * lp)
{
        FILE * f = fopen(path, "rb");
        char magic[8];
        uint64_t dims[3];
        if (f == NULL || fread(magic, 1, 8, f) != 8
            || memcmp(magic, "ACDLP001", 8) || fread(dims, 8, 3, f) != 3) {
                fprintf(stderr, "ora_bench: bad instance %s\n", path);
                exit(2);
        }
        lp->m = dims[0];
        lp->n = dims[1];
        lp->nnz = dims[2];
        lp->rows = read_n(f, lp->nnz, 4);
        lp->cols = read_n(f, lp->nnz, 4);
        lp->vals = read_n(f, lp->nnz, 8);
        lp->rhs = read_n(f, lp->m, 8);
        lp->linear = read_n(f, lp->n, 8);
        lp->lower = read_n(f, lp->n, 8);
        lp->upper = read_n(f, lp->n, 8);
        lp->llo = read_n(f, lp->m, 8);
        lp->lhi = read_n(f, lp->m, 8);
        fclose(f);
}

static void free_lp(struct lp * lp)
{
        free(lp->rows); free(lp->cols); free(lp->vals); free(lp->rhs);
        free(lp->linear); free(lp->lower); free(lp->upper);
        free(lp->llo); free(lp->lhi);
}

/* ---- alm ---- */
struct alm_job {
        char ** files;
        size_t nfiles;
        double stop_kkt;
        size_t next;                    /* shared counter */
        double * seconds, * kkt, * objective;
        int * ret;
};

static void * alm_worker(void * p)
{
        struct alm_job * job = p;
        for (;;) {
                size_t i = __atomic_fetch_add(&job->next, 1, __ATOMIC_RELAXED);
                if (i >= job->nfiles) break;
                struct lp lp;
                read_lp(job->files[i], &lp);
                double t = now_s();
                ora_matrix_t * m = ora_sparse_matrix_make(
                        lp.n, lp.m, lp.nnz, lp.rows, lp.cols, lp.vals, 0);
                ora_alm_t * a = ora_alm_make(m, lp.m, lp.rhs, lp.n, lp.linear,
                                             lp.lower, lp.upper, lp.llo,
                                             lp.lhi);
                double * x = calloc(lp.n + 1, sizeof(double));
                double * lam = calloc(lp.m + 1, sizeof(double));
                double diag[3];
                struct ora_alm_options o;
                ora_alm_default_options(&o);
                o.stop_kkt = job->stop_kkt;
                job->ret[i] = ora_alm_solve(a, 1000, x, lp.n, lam, lp.m, NULL,
                                            diag, &o);
                job->seconds[i] = now_s() - t;
                job->kkt[i] = ora_alm_kkt(a, x, lam);
                job->objective[i] = diag[0];
                free(x);
                free(lam);
                ora_alm_free(a);
                ora_sparse_matrix_free(m);
                free_lp(&lp);
        }
        return NULL;
}

static int run_alm(int threads, double stop_kkt, char ** files, size_t n)
{
        struct alm_job job = {files, n, stop_kkt, 0,
                              calloc(n, 8), calloc(n, 8), calloc(n, 8),
                              calloc(n, sizeof(int))};
        pthread_t * th = calloc(threads, sizeof(pthread_t));
        double t = now_s();
        for (int k = 0; k < threads; k++)
                pthread_create(&th[k], NULL, alm_worker, &job);
        for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
        double wall = now_s() - t;
        printf("{\"mode\": \"alm\", \"threads\": %d, \"lps\": %zu, "
               "\"wall_s\": %.9g, \"seconds\": [", threads, n, wall);
        for (size_t i = 0; i < n; i++)
                printf("%s%.9g", i ? ", " : "", job.seconds[i]);
        printf("], \"kkt\": [");
        for (size_t i = 0; i < n; i++)
                printf("%s%.6g", i ? ", " : "", job.kkt[i]);
        printf("], \"objective\": [");
        for (size_t i = 0; i < n; i++)
                printf("%s%.17g", i ? ", " : "", job.objective[i]);
        printf("], \"ret\": [");
        for (size_t i = 0; i < n; i++)
                printf("%s%d", i ? ", " : "", job.ret[i]);
        printf("]}\n");
        free(th);
        return 0;
}

/* ---- spmv ---- */
struct spmv_job {
        const uint32_t * ptr, * col;
        const double * val;
        const double * x;
        double * y;
        size_t rows;
        size_t next;
};

static void * spmv_worker(void * p)
{
        struct spmv_job * job = p;
        for (;;) {
                size_t r0 = __atomic_fetch_add(&job->next, 16, __ATOMIC_RELAXED);
                if (r0 >= job->rows) break;
                size_t r1 = r0 + 16 < job->rows ? r0 + 16 : job->rows;
                for (size_t r = r0; r < r1; r++) {
                        double acc = 0;
                        for (uint32_t e = job->ptr[r]; e < job->ptr[r + 1]; e++)
                                acc = acc + job->val[e] * job->x[job->col[e]];
                        job->y[r] = acc;
                }
        }
        return NULL;
}

static double spmv_once(struct spmv_job * job, int threads, pthread_t * th)
{
        job->next = 0;
        double t = now_s();
        for (int k = 1; k < threads; k++)
                pthread_create(&th[k], NULL, spmv_worker, job);
        spmv_worker(job);
        for (int k = 1; k < threads; k++) pthread_join(th[k], NULL);
        return now_s() - t;
}

static int run_spmv(int threads, int repeat, const char * file)
{
        struct lp lp;
        read_lp(file, &lp);
        ora_matrix_t * m = ora_sparse_matrix_make(lp.n, lp.m, lp.nnz, lp.rows,
                                                  lp.cols, lp.vals, 0);
        pthread_t * th = calloc(threads, sizeof(pthread_t));
        double sec[2];
        for (int tr = 0; tr < 2; tr++) {
                size_t rows = tr ? lp.n : lp.m, cols = tr ? lp.m : lp.n;
                double * x = calloc(cols + 1, sizeof(double));
                double * y = calloc(rows + 1, sizeof(double));
                for (size_t i = 0; i < cols; i++) x[i] = 1.0 / (1.0 + i % 7);
                struct spmv_job job = {ora_csr_row_ptr(m, tr),
                                       ora_csr_columns(m, tr),
                                       ora_csr_values(m, tr), x, y, rows, 0};
                spmv_once(&job, threads, th);
                double total = 0;
                for (int k = 0; k < repeat; k++)
                        total += spmv_once(&job, threads, th);
                sec[tr] = total / repeat;
                free(x);
                free(y);
        }
        printf("{\"mode\": \"spmv\", \"threads\": %d, \"repeat\": %d, "
               "\"s_per_a\": %.9g, \"s_per_at\": %.9g}\n",
               threads, repeat, sec[0], sec[1]);
        free(th);
        ora_sparse_matrix_free(m);
        free_lp(&lp);
        return 0;
}

static int run_approx(int iters, const char * file)
{
        struct lp lp;
        read_lp(file, &lp);
        ora_matrix_t * m = ora_sparse_matrix_make(lp.n, lp.m, lp.nnz, lp.rows,
                                                  lp.cols, lp.vals, 0);
        ora_approx_t * a = ora_approx_make(m, lp.m, lp.rhs, NULL, lp.n,
                                           lp.linear, lp.lower, lp.upper);
        ora_approx_update(a);
        double * x = calloc(lp.n + 1, sizeof(double));
        double diag[5];
        double t = now_s();
        ora_approx_solve(x, lp.n, a, iters, 0, -HUGE_VAL, 0, NULL, 0, diag,
                         0);
        double sec = now_s() - t;
        printf("{\"mode\": \"approx\", \"iterations\": %d, "
               "\"s_per_iteration\": %.9g}\n", iters, sec / iters);
        free(x);
        ora_approx_free(a);
        ora_sparse_matrix_free(m);
        free_lp(&lp);
        return 0;
}

int main(int argc, char ** argv)
{
        if (argc == 4 && !strcmp(argv[1], "approx"))
                return run_approx(atoi(argv[2]), argv[3]);
        if (argc >= 5 && !strcmp(argv[1], "alm"))
                return run_alm(atoi(argv[2]), atof(argv[3]), argv + 4,
                               (size_t)(argc - 4));
        if (argc == 5 && !strcmp(argv[1], "spmv"))
                return run_spmv(atoi(argv[2]), atoi(argv[3]), argv[4]);
        fprintf(stderr, "usage: %s alm THREADS STOP_KKT FILE... | "
                "spmv THREADS REPEAT FILE | approx ITERS FILE\n", argv[0]);
        return 2;
}
