/* oracle.h -- CPU restatement of pkhuong/acd-alm's LP path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the acd-alm_amd product links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function follows the reference file:line named in its comment, in
 * serial mode (pool == NULL), and keeps the reference's floating-point
 * evaluation order (two-lane even/odd accumulators, per-row sequential
 * SpMV sums, no fused multiply-add), so its outputs are bit-identical to
 * the reference built with gcc -O2 -ffp-contract=off.  That identity is
 * checked against the reference compiled from its own sources
 * (oracle/Makefile -> oracle/_ref/) and against the golden fixtures under
 * tests/golden/.
 *
 * Symbols carry an ora_ prefix so the oracle, the reference and the GPU
 * library can share one process.
 */
#ifndef ORA_ORACLE_H
#define ORA_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_matrix ora_matrix_t;
typedef struct ora_approx ora_approx_t;
typedef struct ora_alm ora_alm_t;

/* ---- sparse matrix (spmv/spmv.c, permutation.c, spmv_csr.c,
 *      spmv.swizzle.inc) ---- */
ora_matrix_t * ora_sparse_matrix_make(size_t ncolumns, size_t nrows,
                                      size_t nnz, const uint32_t * rows,
                                      const uint32_t * columns,
                                      const double * values, int permute);
ora_matrix_t * ora_sparse_matrix_copy(const ora_matrix_t *, int permute);
int ora_sparse_matrix_free(ora_matrix_t *);
size_t ora_sparse_matrix_ncolumns(const ora_matrix_t *);
size_t ora_sparse_matrix_nrows(const ora_matrix_t *);
size_t ora_sparse_matrix_nnz(const ora_matrix_t *);
const uint32_t * ora_sparse_matrix_rows(const ora_matrix_t *);
const uint32_t * ora_sparse_matrix_columns(const ora_matrix_t *);
const double * ora_sparse_matrix_values(const ora_matrix_t *);
int ora_sparse_matrix_row_permute(const ora_matrix_t *, double * dest,
                                  size_t n, const double * src, int dir);
int ora_sparse_matrix_col_permute(const ora_matrix_t *, double * dest,
                                  size_t n, const double * src, int dir);
int ora_sparse_matrix_multiply(double * y, size_t ny, const ora_matrix_t *,
                               const double * x, size_t nx, int transpose);
int ora_sparse_matrix_multiply_2(double ** y, size_t ny,
                                 const ora_matrix_t *, const double ** x,
                                 size_t nx, int transpose);
/* Internal arrays, for tests: CSR of A (transpose=0) or of A'. */
const uint32_t * ora_csr_row_ptr(const ora_matrix_t *, int transpose);
const uint32_t * ora_csr_columns(const ora_matrix_t *, int transpose);
const double * ora_csr_values(const ora_matrix_t *, int transpose);
/* Permutation arrays: row (which=1) or column (which=0); idx or ridx. */
const uint32_t * ora_perm_idx(const ora_matrix_t *, int row);
const uint32_t * ora_perm_ridx(const ora_matrix_t *, int row);
/* Z-order key of (column, row): column bits at even positions. */
uint64_t ora_morton_key(uint32_t column, uint32_t row);

/* ---- APPROX (approx/approx.c, vector.inc, gradient_value.inc,
 *      step_project.inc, mpg.inc) ---- */
ora_approx_t * ora_approx_make(ora_matrix_t * constraints, size_t nrhs,
                               const double * rhs, const double * weight,
                               size_t nvars, const double * linear,
                               const double * lower, const double * upper);
int ora_approx_free(ora_approx_t *);
ora_matrix_t * ora_approx_matrix(ora_approx_t *);
double * ora_approx_rhs(ora_approx_t *);
double * ora_approx_weight(ora_approx_t *);
double * ora_approx_linear(ora_approx_t *);
double * ora_approx_lower(ora_approx_t *);
double * ora_approx_upper(ora_approx_t *);
int ora_approx_update(ora_approx_t *);
/* Step sizes of the permuted working copy (v and 1/v), for tests. */
const double * ora_approx_step_v(ora_approx_t *, int permuted);
const double * ora_approx_step_inv_v(ora_approx_t *, int permuted);
int ora_approx_solve(double * x, size_t n, ora_approx_t *, size_t niter,
                     double max_pg, double max_value, double min_delta,
                     FILE * log, size_t period, double * diag, double offset);
int ora_mpg_solve(double * x, size_t n, ora_approx_t *, size_t niter,
                  double max_pg, double max_value, double min_delta,
                  FILE * log, size_t period, double * diag, double offset);
int ora_cd_solve(double * x, size_t n, ora_approx_t *, size_t niter,
                 double max_pg, double max_value, double min_delta,
                 FILE * log, size_t period, double * diag, double offset);

/* Test knob: sum every APPROX/SPG reduction in reverse element order
 * (another valid summation order); 0 restores the reference's order. */
void ora_set_reverse_sums(int on);

/* ---- ALM (alm.c) ---- */
/* Inner solver used by the outer loop. */
enum { ORA_INNER_APPROX = 0, ORA_INNER_CD = 1 };
/* Stopping constants of the outer loop; the reference's are
 * {1e-5, 1e-5, 1e-7, 1e-11} (alm.c:306-308, :243-244, :256). */
struct ora_alm_options {
        double stop_violation;
        double stop_pg;
        double precision_floor;
        double min_delta;
        int inner;
        /* > 0: stop once the relative KKT residual of (x, lambda) is at
         * most this after an outer iteration, instead of the constants
         * above (the stopping rule of the device library's stop_kkt). */
        double stop_kkt;
};
void ora_alm_default_options(struct ora_alm_options *);
/* Relative KKT residual max(r_p, r_d, r_gap) of (x, lambda) for the LP of
 * an ALM object, with d = c - A'lambda:
 *   r_p = |Ax - b| / (1 + |b|),
 *   r_d = (|x - clamp(x - d, l, u)| + |dual infeasibility|) / (1 + |c|),
 *   r_gap = |c'x - dual| / (1 + |c'x| + |dual|),
 *   dual = b'lambda + sum(l d+) - sum(u d-) over finite bounds,
 * where a reduced cost pushing against an infinite bound is dual
 * infeasibility.  The reference defines no such measure (its rule is
 * absolute, alm.c:303-319); this is the one the device library stops on. */
double ora_alm_kkt(ora_alm_t *, const double * x, const double * lambda);
ora_alm_t * ora_alm_make(ora_matrix_t * constraints, size_t nrhs,
                         const double * rhs, size_t nvars,
                         const double * linear, const double * lower,
                         const double * upper, const double * lambda_lower,
                         const double * lambda_upper);
int ora_alm_free(ora_alm_t *);
/* alm.c:23-43 accessors: the LP cost (mutable) and the approx object that
 * holds rhs, lower and upper (mutable through ora_approx_*). */
double * ora_alm_linear(ora_alm_t *);
ora_approx_t * ora_alm_approx(ora_alm_t *);
int ora_alm_solve(ora_alm_t *, size_t niter, double * x, size_t nvars,
                  double * lambda, size_t nconstraints, FILE * log,
                  double * diag, const struct ora_alm_options *);

#ifdef __cplusplus
}
#endif
#endif
