/* ora_spmv.c -- oracle restatement of the reference's sparse-matrix layer.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows spmv/spmv.c:27-233 (build, copy, multiply), spmv/permutation.c
 * (windowed count sort), spmv/spmv_csr.c:27-80 (CSR by sorted
 * (row << 32 | column) keys) and spmv/spmv.swizzle.inc:1-52 (Z-order).
 * The reference multiplies through 4-row padded slabs
 * (spmv/spmv_block.c:160-254); each output there is the sequential sum,
 * in CSR order, of value*x products starting from +0.0, the padding
 * adding +-0.0 terms that leave the sum unchanged for finite x.  The
 * oracle sums the CSR row directly, which yields the same bits.
 */
#include "oracle.h"
#include <assert.h>
#include <stdlib.h>
#include <string.h>

struct ora_csr {
        size_t nrows;
        uint32_t * ptr;    /* nrows + 1 */
        uint32_t * col;
        double * val;
};

struct ora_matrix {
        size_t ncolumns, nrows, nnz;
        uint32_t * rows, * columns;
        double * values;
        struct ora_csr a, at;
        uint32_t * row_idx, * row_ridx;
        uint32_t * col_idx, * col_ridx;
};

struct keyed { uint64_t key; uint32_t row, col; double val; };

static int cmp_keyed(const void * a, const void * b)
{
        uint64_t x = ((const struct keyed *)a)->key,
                y = ((const struct keyed *)b)->key;
        return (x > y) - (x < y);
}

/* spmv/spmv.swizzle.inc:1-12: bit i of the column goes to bit 2i, bit i
 * of the row to bit 2i+1. */
uint64_t ora_morton_key(uint32_t column, uint32_t row)
{
        uint64_t key = 0;
        for (unsigned b = 0; b < 32; b++) {
                key |= (uint64_t)((column >> b) & 1u) << (2 * b);
                key |= (uint64_t)((row >> b) & 1u) << (2 * b + 1);
        }
        return key;
}

/* spmv/permutation.c:63-125 with PERMUTATION_SORTING_SCOPE = 256 and
 * neither PERMUTATION_FIRST_NZ nor PERMUTATION_SHUFFLE: the tie-break key
 * "least" is the constant 0xFFFFFFFF, so windows of 256 consecutive
 * indices are sorted by (entry count, original index). */
struct count_index { uint32_t count, index; };

static int cmp_count_index(const void * a, const void * b)
{
        const struct count_index * x = a, * y = b;
        if (x->count != y->count) return (x->count > y->count) ? 1 : -1;
        return (x->index > y->index) - (x->index < y->index);
}

static void make_permutation(const uint32_t * keys, size_t nnz, size_t n,
                             uint32_t ** OUT_idx, uint32_t ** OUT_ridx)
{
        struct count_index * ci = calloc(n ? n : 1, sizeof(*ci));
        for (size_t i = 0; i < n; i++) ci[i].index = (uint32_t)i;
        for (size_t k = 0; k < nnz; k++) {
                assert(keys[k] < n);
                ci[keys[k]].count++;
        }
        for (size_t w = 0; w < n; w += 256) {
                size_t len = (n - w < 256) ? n - w : 256;
                qsort(ci + w, len, sizeof(*ci), cmp_count_index);
        }
        uint32_t * idx = calloc(n ? n : 1, sizeof(uint32_t));
        uint32_t * ridx = calloc(n ? n : 1, sizeof(uint32_t));
        for (size_t i = 0; i < n; i++) {
                idx[i] = ci[i].index;
                ridx[ci[i].index] = (uint32_t)i;
        }
        free(ci);
        *OUT_idx = idx;
        *OUT_ridx = ridx;
}

static void identity_permutation(size_t n, uint32_t ** idx, uint32_t ** ridx)
{
        *idx = calloc(n ? n : 1, sizeof(uint32_t));
        *ridx = calloc(n ? n : 1, sizeof(uint32_t));
        for (size_t i = 0; i < n; i++) (*idx)[i] = (*ridx)[i] = (uint32_t)i;
}

/* spmv/spmv_csr.c:27-80: sort by (major << 32 | minor); the row pointer
 * array is the running maximum of "last position + 1" per major index. */
static void build_csr(struct ora_csr * csr, const ora_matrix_t * m,
                      int transpose)
{
        size_t nnz = m->nnz;
        size_t nmajor = transpose ? m->ncolumns : m->nrows;
        struct keyed * e = calloc(nnz ? nnz : 1, sizeof(*e));
        for (size_t k = 0; k < nnz; k++) {
                uint32_t major = transpose ? m->columns[k] : m->rows[k];
                uint32_t minor = transpose ? m->rows[k] : m->columns[k];
                e[k].key = ((uint64_t)major << 32) | minor;
                e[k].row = major;
                e[k].col = minor;
                e[k].val = m->values[k];
        }
        qsort(e, nnz, sizeof(*e), cmp_keyed);
        csr->nrows = nmajor;
        csr->ptr = calloc(nmajor + 1, sizeof(uint32_t));
        csr->col = calloc(nnz ? nnz : 1, sizeof(uint32_t));
        csr->val = calloc(nnz ? nnz : 1, sizeof(double));
        for (size_t k = 0; k < nnz; k++) {
                csr->ptr[e[k].row + 1] = (uint32_t)(k + 1);
                csr->col[k] = e[k].col;
                csr->val[k] = e[k].val;
        }
        uint32_t run = 0;
        for (size_t i = 0; i <= nmajor; i++) {
                if (csr->ptr[i] > run) run = csr->ptr[i];
                csr->ptr[i] = run;
        }
        free(e);
}

static void free_csr(struct ora_csr * csr)
{
        free(csr->ptr);
        free(csr->col);
        free(csr->val);
        memset(csr, 0, sizeof(*csr));
}

/* spmv/spmv.swizzle.inc:14-52 */
static void morton_sort(ora_matrix_t * m)
{
        size_t nnz = m->nnz;
        struct keyed * e = calloc(nnz ? nnz : 1, sizeof(*e));
        for (size_t k = 0; k < nnz; k++) {
                e[k].key = ora_morton_key(m->columns[k], m->rows[k]);
                e[k].row = m->rows[k];
                e[k].col = m->columns[k];
                e[k].val = m->values[k];
        }
        qsort(e, nnz, sizeof(*e), cmp_keyed);
        for (size_t k = 0; k < nnz; k++) {
                m->rows[k] = e[k].row;
                m->columns[k] = e[k].col;
                m->values[k] = e[k].val;
        }
        free(e);
}

/* spmv/spmv.c:27-92 */
ora_matrix_t * ora_sparse_matrix_make(size_t ncolumns, size_t nrows,
                                      size_t nnz, const uint32_t * rows,
                                      const uint32_t * columns,
                                      const double * values, int permute)
{
        ora_matrix_t * m = calloc(1, sizeof(*m));
        m->ncolumns = ncolumns;
        m->nrows = nrows;
        m->nnz = nnz;
        m->rows = calloc(nnz ? nnz : 1, sizeof(uint32_t));
        m->columns = calloc(nnz ? nnz : 1, sizeof(uint32_t));
        m->values = calloc(nnz ? nnz : 1, sizeof(double));
        if (nnz) {
                memcpy(m->rows, rows, nnz * sizeof(uint32_t));
                memcpy(m->columns, columns, nnz * sizeof(uint32_t));
                memcpy(m->values, values, nnz * sizeof(double));
        }
        if (permute) {
                make_permutation(m->rows, nnz, nrows,
                                 &m->row_idx, &m->row_ridx);
                make_permutation(m->columns, nnz, ncolumns,
                                 &m->col_idx, &m->col_ridx);
                for (size_t k = 0; k < nnz; k++) {
                        m->rows[k] = m->row_ridx[m->rows[k]];
                        m->columns[k] = m->col_ridx[m->columns[k]];
                }
        } else {
                identity_permutation(nrows, &m->row_idx, &m->row_ridx);
                identity_permutation(ncolumns, &m->col_idx, &m->col_ridx);
        }
        build_csr(&m->a, m, 0);
        build_csr(&m->at, m, 1);
        morton_sort(m);
        return m;
}

ora_matrix_t * ora_sparse_matrix_copy(const ora_matrix_t * src, int permute)
{
        return ora_sparse_matrix_make(src->ncolumns, src->nrows, src->nnz,
                                      src->rows, src->columns, src->values,
                                      permute);
}

int ora_sparse_matrix_free(ora_matrix_t * m)
{
        if (m == NULL) return 0;
        free(m->rows);
        free(m->columns);
        free(m->values);
        free_csr(&m->a);
        free_csr(&m->at);
        free(m->row_idx);
        free(m->row_ridx);
        free(m->col_idx);
        free(m->col_ridx);
        free(m);
        return 0;
}

size_t ora_sparse_matrix_ncolumns(const ora_matrix_t * m) { return m->ncolumns; }
size_t ora_sparse_matrix_nrows(const ora_matrix_t * m) { return m->nrows; }
size_t ora_sparse_matrix_nnz(const ora_matrix_t * m) { return m->nnz; }
const uint32_t * ora_sparse_matrix_rows(const ora_matrix_t * m) { return m->rows; }
const uint32_t * ora_sparse_matrix_columns(const ora_matrix_t * m) { return m->columns; }
const double * ora_sparse_matrix_values(const ora_matrix_t * m) { return m->values; }

const uint32_t * ora_csr_row_ptr(const ora_matrix_t * m, int t) { return t ? m->at.ptr : m->a.ptr; }
const uint32_t * ora_csr_columns(const ora_matrix_t * m, int t) { return t ? m->at.col : m->a.col; }
const double * ora_csr_values(const ora_matrix_t * m, int t) { return t ? m->at.val : m->a.val; }
const uint32_t * ora_perm_idx(const ora_matrix_t * m, int row) { return row ? m->row_idx : m->col_idx; }
const uint32_t * ora_perm_ridx(const ora_matrix_t * m, int row) { return row ? m->row_ridx : m->col_ridx; }

/* spmv/permutation.c:136-148 */
static int permute_vector(double * dest, size_t n, size_t expect,
                          const uint32_t * idx, const uint32_t * ridx,
                          const double * src, int direction)
{
        if (n != expect) return 1;
        const uint32_t * p = (direction >= 0) ? idx : ridx;
        for (size_t i = 0; i < n; i++) dest[i] = src[p[i]];
        return 0;
}

int ora_sparse_matrix_row_permute(const ora_matrix_t * m, double * dest,
                                  size_t n, const double * src, int dir)
{
        return permute_vector(dest, n, m->nrows, m->row_idx, m->row_ridx,
                              src, dir);
}

int ora_sparse_matrix_col_permute(const ora_matrix_t * m, double * dest,
                                  size_t n, const double * src, int dir)
{
        return permute_vector(dest, n, m->ncolumns, m->col_idx, m->col_ridx,
                              src, dir);
}

/* One output of spmv/spmv_block.c:160-194: acc = +0.0, then acc += v*x
 * over the row in ascending column order. */
static void csr_mult(double * y, const struct ora_csr * c, const double * x)
{
        for (size_t r = 0; r < c->nrows; r++) {
                double acc = 0.0;
                for (uint32_t k = c->ptr[r]; k < c->ptr[r + 1]; k++)
                        acc += c->val[k] * x[c->col[k]];
                y[r] = acc;
        }
}

/* spmv/spmv.c:159-193 */
int ora_sparse_matrix_multiply(double * y, size_t ny, const ora_matrix_t * m,
                               const double * x, size_t nx, int transpose)
{
        const struct ora_csr * c = transpose ? &m->at : &m->a;
        assert(ny == (transpose ? m->ncolumns : m->nrows));
        assert(nx == (transpose ? m->nrows : m->ncolumns));
        csr_mult(y, c, x);
        return 0;
}

/* spmv/spmv.c:195-233: each right-hand side accumulates independently. */
int ora_sparse_matrix_multiply_2(double ** y, size_t ny, const ora_matrix_t * m,
                                 const double ** x, size_t nx, int transpose)
{
        const struct ora_csr * c = transpose ? &m->at : &m->a;
        assert(ny == (transpose ? m->ncolumns : m->nrows));
        assert(nx == (transpose ? m->nrows : m->ncolumns));
        csr_mult(y[0], c, x[0]);
        csr_mult(y[1], c, x[1]);
        return 0;
}
