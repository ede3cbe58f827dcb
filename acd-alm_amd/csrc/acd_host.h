/* acd_host.h -- private host-side object layouts of the C library. */
#ifndef ACD_HOST_H
#define ACD_HOST_H
#include <stdint.h>
#include <stddef.h>
#include "spmv/spmv.h"
#include "approx/approx.h"
#include "alm.h"
#include "acd_dev.h"
#include "acd_gpu.h"

/* spmv/spmv_internal.h:25-41 of the reference, re-laid for the GPU: host
 * triplets for the accessors, device operators for the arithmetic. */
struct sparse_matrix {
        size_t ncolumns, nrows, nnz;
        uint32_t * rows, * columns;    /* host, renumbered, Z-order */
        double * values;
        uint32_t * row_idx, * row_ridx, * col_idx, * col_ridx;  /* host */
        struct acd_dev_matrix dev;
        double * stage_x, * stage_y;   /* device staging for host calls */
        double * stage_x1, * stage_y1;
};

/* Abort with a message, as the reference's assertions do. */
void acd_fail(const char * what);
#define ACD_REQUIRE(cond, what) do { if (!(cond)) acd_fail(what); } while (0)

/* Device staging of a matrix (ensures stage buffers exist). */
int acd_matrix_stage(struct sparse_matrix *);
#endif
