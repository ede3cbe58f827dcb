/* spmv.h -- sparse matrices resident in MI355X HBM, and y = op(A) x.
 *
 * Drop-in for spmv/spmv.h of pkhuong/acd-alm (same prototypes, same
 * return codes).  A matrix built here keeps, in device memory, a row-major
 * and a transposed sliced-ELL copy (64 rows per slice, one wavefront per
 * slice), the row/column permutations, and per-row entry counts.  The
 * host keeps the coordinate triplets so the read-only accessors below
 * return the same arrays, in the same order, as the original library:
 * rows/columns renumbered by the permutation when one was requested, then
 * sorted in Z-order (interleaved column/row bits).
 */
#ifndef ACD_SPMV_H
#define ACD_SPMV_H
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include "../thread_pool/thread_pool.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sparse_matrix sparse_matrix_t;

/* Call once before building matrices.  Selects the HIP device (device 0
 * unless acd_set_device() chose another) and creates the library stream.
 * (spmv/spmv.h:13) */
void sparse_matrix_init(void);

/* Build from coordinate triplets (any order, no repeated (row, column)).
 * Note the argument order: ncolumns before nrows.  With permute != 0 the
 * rows and the columns are renumbered: within consecutive windows of 256
 * indices, indices are ordered by ascending entry count, ties by index.
 * All preprocessing (permutation, CSR/transposed CSR, slicing, Z-order
 * sort) runs on the device and is bit-identical to the original library.
 * Returns NULL if the device cannot hold the matrix.  (spmv/spmv.h:19-24)
 */
sparse_matrix_t * sparse_matrix_make(size_t ncolumns, size_t nrows,
                                     size_t nnz,
                                     const uint32_t * rows,
                                     const uint32_t * columns,
                                     const double * values,
                                     int permute);
/* Rebuild from another matrix's (already renumbered) triplets, optionally
 * renumbering again.  (spmv/spmv.h:25-26) */
sparse_matrix_t * sparse_matrix_copy(const sparse_matrix_t * matrix,
                                     int permute);
/* Release host and device storage; NULL is ignored.  Returns 0.
 * (spmv/spmv.h:29) */
int sparse_matrix_free(sparse_matrix_t *);

/* Read-only accessors.  rows/columns/values are host arrays of nnz
 * entries in Z-order.  (spmv/spmv.h:34-39) */
size_t sparse_matrix_ncolumns(sparse_matrix_t *);
size_t sparse_matrix_nrows(sparse_matrix_t *);
size_t sparse_matrix_nnz(sparse_matrix_t *);
const uint32_t * sparse_matrix_rows(sparse_matrix_t *);
const uint32_t * sparse_matrix_columns(sparse_matrix_t *);
const double * sparse_matrix_values(sparse_matrix_t *);

/* dest[i] = src[perm[i]] with perm the forward row (column) permutation
 * when direction >= 0, the inverse one otherwise.  Host arrays; the gather
 * runs on the device.  Returns 1 if n is not the permutation's length,
 * 0 otherwise.  (spmv/spmv.h:41-44) */
int sparse_matrix_row_permute(sparse_matrix_t *, double * dest, size_t n,
                              const double * src, int direction);
int sparse_matrix_col_permute(sparse_matrix_t *, double * dest, size_t n,
                              const double * src, int direction);

/* OUT_y <- A x (transpose == 0) or A' x (transpose != 0).  Host arrays of
 * ny and nx doubles; ny/nx must equal the operator's dimensions (checked,
 * the call aborts otherwise, as the original's assertions do).  Each
 * output is the sequential FP64 sum, in ascending index order, of its
 * row's products, without fused multiply-add: bit-identical to the
 * original library.  pool is accepted and ignored.  Returns 0.
 * (spmv/spmv.h:58-62) */
int sparse_matrix_multiply(double * OUT_y, size_t ny,
                           const sparse_matrix_t * a,
                           const double * x, size_t nx,
                           int transpose,
                           thread_pool_t * pool);
/* Two right-hand sides sharing one traversal of the matrix.
 * (spmv/spmv.h:70-74) */
int sparse_matrix_multiply_2(double ** OUT_y, size_t ny,
                             const sparse_matrix_t * a,
                             const double ** x, size_t nx,
                             int transpose,
                             thread_pool_t * pool);

/* Text format: "nrows ncolumns nnz" then nnz "row column value" triplets
 * (0-based).  Returns NULL on malformed input.  (spmv/spmv.h:83) */
sparse_matrix_t * sparse_matrix_read(FILE * stream, int permute);

#ifdef __cplusplus
}
#endif
#endif
