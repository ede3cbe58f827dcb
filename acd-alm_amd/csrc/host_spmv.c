/* host_spmv.c -- the spmv.h entry points.
 *
 * Mirrors spmv/spmv.c:20-259 of pkhuong/acd-alm.  The matrix build, the
 * products and the permutation gathers run on the device (dev_*.hip);
 * this file keeps the host-visible triplets and marshals host arrays.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_host.h"

void acd_fail(const char * what)
{
        char buf[256];
        acd_dev_last_error(buf, sizeof buf);
        fprintf(stderr, "acd-alm-amd: %s%s%s\n", what, buf[0] ? ": " : "", buf);
        abort();
}

/* spmv/spmv.c:20-25 */
void sparse_matrix_init(void)
{
        if (acd_dev_init())
                acd_fail("sparse_matrix_init: no usable HIP device");
        /* device address of the allocation-free product's staging array */
        if (acd_dev_scratch_setup())
                acd_fail("sparse_matrix_init: staging array");
}

/* spmv/spmv.c:27-92 */
sparse_matrix_t * sparse_matrix_make(size_t ncolumns, size_t nrows,
                                     size_t nnz, const uint32_t * rows,
                                     const uint32_t * columns,
                                     const double * values, int permute)
{
        if (nrows > UINT32_MAX || ncolumns > UINT32_MAX || nnz > UINT32_MAX)
                return NULL;
        if (acd_dev_init()) return NULL;
        sparse_matrix_t * m = calloc(1, sizeof(*m));
        if (m == NULL) return NULL;
        m->ncolumns = ncolumns;
        m->nrows = nrows;
        m->nnz = nnz;
        m->rows = malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
        m->columns = malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
        m->values = malloc(sizeof(double) * (nnz ? nnz : 1));
        m->row_idx = malloc(sizeof(uint32_t) * (nrows ? nrows : 1));
        m->row_ridx = malloc(sizeof(uint32_t) * (nrows ? nrows : 1));
        m->col_idx = malloc(sizeof(uint32_t) * (ncolumns ? ncolumns : 1));
        m->col_ridx = malloc(sizeof(uint32_t) * (ncolumns ? ncolumns : 1));
        if (!m->rows || !m->columns || !m->values || !m->row_idx
            || !m->row_ridx || !m->col_idx || !m->col_ridx) {
                sparse_matrix_free(m);
                return NULL;
        }
        if (nnz) {
                memcpy(m->rows, rows, nnz * sizeof(uint32_t));
                memcpy(m->columns, columns, nnz * sizeof(uint32_t));
                memcpy(m->values, values, nnz * sizeof(double));
        }
        for (size_t k = 0; k < nnz; k++)
                ACD_REQUIRE(m->rows[k] < nrows && m->columns[k] < ncolumns,
                            "sparse_matrix_make: index out of range");
        if (acd_dev_matrix_build(&m->dev, (uint32_t)nrows, (uint32_t)ncolumns,
                                 nnz, m->rows, m->columns, m->values,
                                 permute, m->row_idx, m->row_ridx,
                                 m->col_idx, m->col_ridx)) {
                sparse_matrix_free(m);
                return NULL;
        }
        return m;
}

/* spmv/spmv.c:94-100 */
sparse_matrix_t * sparse_matrix_copy(const sparse_matrix_t * src, int permute)
{
        return sparse_matrix_make(src->ncolumns, src->nrows, src->nnz,
                                  src->rows, src->columns, src->values,
                                  permute);
}

/* spmv/spmv.c:102-124 */
int sparse_matrix_free(sparse_matrix_t * m)
{
        if (m == NULL) return 0;
        acd_dev_matrix_free(&m->dev);
        acd_dev_free(m->stage_x);
        acd_dev_free(m->stage_y);
        acd_dev_free(m->stage_x1);
        acd_dev_free(m->stage_y1);
        free(m->rows);
        free(m->columns);
        free(m->values);
        free(m->row_idx);
        free(m->row_ridx);
        free(m->col_idx);
        free(m->col_ridx);
        memset(m, 0, sizeof(*m));
        free(m);
        return 0;
}

/* spmv/spmv.c:126-139 */
size_t sparse_matrix_ncolumns(sparse_matrix_t * m) { return m->ncolumns; }
size_t sparse_matrix_nrows(sparse_matrix_t * m) { return m->nrows; }
size_t sparse_matrix_nnz(sparse_matrix_t * m) { return m->nnz; }
const uint32_t * sparse_matrix_rows(sparse_matrix_t * m) { return m->rows; }
const uint32_t * sparse_matrix_columns(sparse_matrix_t * m) { return m->columns; }
const double * sparse_matrix_values(sparse_matrix_t * m) { return m->values; }

int acd_matrix_stage(struct sparse_matrix * m)
{
        size_t len = (m->nrows > m->ncolumns ? m->nrows : m->ncolumns) + 1;
        if (m->stage_x) return 0;
        m->stage_x = acd_dev_malloc(len * sizeof(double));
        m->stage_y = acd_dev_malloc(len * sizeof(double));
        m->stage_x1 = acd_dev_malloc(len * sizeof(double));
        m->stage_y1 = acd_dev_malloc(len * sizeof(double));
        if (!m->stage_x || !m->stage_y || !m->stage_x1 || !m->stage_y1)
                return -1;
        return 0;
}

/* spmv/permutation.c:136-148, gathered on the device. */
static int permute(struct sparse_matrix * m, double * dest, size_t n,
                   const double * src, const uint32_t * fwd,
                   const uint32_t * inv, size_t expect)
{
        if (n != expect) return 1;
        if (n == 0) return 0;
        if (acd_matrix_stage(m)) acd_fail("permute: device staging");
        const uint32_t * p = fwd ? fwd : inv;
        if (acd_dev_h2d(m->stage_x, src, n * sizeof(double))
            || acd_dev_gather(m->stage_y, m->stage_x, p, (uint32_t)n)
            || acd_dev_d2h(dest, m->stage_y, n * sizeof(double)))
                acd_fail("permute");
        return 0;
}

int sparse_matrix_row_permute(sparse_matrix_t * m, double * dest, size_t n,
                              const double * src, int direction)
{
        return permute(m, dest, n, src,
                       direction >= 0 ? m->dev.row_idx : NULL,
                       m->dev.row_ridx, m->nrows);
}

int sparse_matrix_col_permute(sparse_matrix_t * m, double * dest, size_t n,
                              const double * src, int direction)
{
        return permute(m, dest, n, src,
                       direction >= 0 ? m->dev.col_idx : NULL,
                       m->dev.col_ridx, m->ncolumns);
}

/* spmv/spmv.c:159-193 */
int sparse_matrix_multiply(double * OUT_y, size_t ny,
                           const sparse_matrix_t * a, const double * x,
                           size_t nx, int transpose, thread_pool_t * pool)
{
        (void)pool;
        struct sparse_matrix * m = (struct sparse_matrix *)a;
        size_t rows = transpose ? m->ncolumns : m->nrows;
        size_t cols = transpose ? m->nrows : m->ncolumns;
        ACD_REQUIRE(ny == rows, "sparse_matrix_multiply: ny mismatch");
        ACD_REQUIRE(nx == cols, "sparse_matrix_multiply: nx mismatch");
        if (rows == 0) return 0;
        if (acd_matrix_stage(m)) acd_fail("multiply: device staging");
        if (acd_dev_h2d(m->stage_x, x, nx * sizeof(double))
            || acd_dev_spmv(&m->dev, transpose, m->stage_y, m->stage_x)
            || acd_dev_d2h(OUT_y, m->stage_y, ny * sizeof(double)))
                acd_fail("sparse_matrix_multiply");
        return 0;
}

int acd_sparse_matrix_multiply_noalloc(double * OUT_y, size_t ny,
                                       const sparse_matrix_t * a,
                                       const double * x, size_t nx,
                                       int transpose)
{
        if (a == NULL || (ny && OUT_y == NULL) || (nx && x == NULL)) {
                acd_dev_record_error(-1, "allocation-free product: NULL "
                                     "argument");
                return -1;
        }
        return acd_dev_spmv_noalloc(&a->dev, transpose, OUT_y, ny, x, nx);
}

/* spmv/spmv.c:195-233 */
int sparse_matrix_multiply_2(double ** OUT_y, size_t ny,
                             const sparse_matrix_t * a, const double ** x,
                             size_t nx, int transpose, thread_pool_t * pool)
{
        (void)pool;
        struct sparse_matrix * m = (struct sparse_matrix *)a;
        size_t rows = transpose ? m->ncolumns : m->nrows;
        size_t cols = transpose ? m->nrows : m->ncolumns;
        ACD_REQUIRE(ny == rows, "sparse_matrix_multiply_2: ny mismatch");
        ACD_REQUIRE(nx == cols, "sparse_matrix_multiply_2: nx mismatch");
        if (rows == 0) return 0;
        if (acd_matrix_stage(m)) acd_fail("multiply_2: device staging");
        if (acd_dev_h2d(m->stage_x, x[0], nx * sizeof(double))
            || acd_dev_h2d(m->stage_x1, x[1], nx * sizeof(double))
            || acd_dev_spmv2(&m->dev, transpose, m->stage_y, m->stage_y1,
                             m->stage_x, m->stage_x1)
            || acd_dev_d2h(OUT_y[0], m->stage_y, ny * sizeof(double))
            || acd_dev_d2h(OUT_y[1], m->stage_y1, ny * sizeof(double)))
                acd_fail("sparse_matrix_multiply_2");
        return 0;
}

/* spmv/spmv.c:235-259 */
sparse_matrix_t * sparse_matrix_read(FILE * stream, int permute)
{
        size_t nrows, ncolumns, nnz;
        if (stream == NULL
            || fscanf(stream, " %zu %zu %zu", &nrows, &ncolumns, &nnz) != 3)
                return NULL;
        uint32_t * rows = calloc(nnz + 1, sizeof(uint32_t));
        uint32_t * cols = calloc(nnz + 1, sizeof(uint32_t));
        double * vals = calloc(nnz + 1, sizeof(double));
        sparse_matrix_t * m = NULL;
        if (rows && cols && vals) {
                size_t i;
                for (i = 0; i < nnz; i++)
                        if (fscanf(stream, " %u %u %lf", rows + i, cols + i,
                                   vals + i) != 3)
                                break;
                if (i == nnz)
                        m = sparse_matrix_make(ncolumns, nrows, nnz, rows,
                                               cols, vals, permute);
        }
        free(rows);
        free(cols);
        free(vals);
        return m;
}
