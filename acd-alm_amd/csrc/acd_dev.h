/* acd_dev.h -- the thin C ABI between the host C layer (host_*.c) and the
 * HIP device layer (dev_*.hip).
 *
 * Host C code never includes HIP headers: every device allocation, copy,
 * launch and synchronisation goes through the functions below, which the
 * .hip translation units implement with extern "C" linkage.  All device
 * work is ordered on one library stream (acd_dev_stream()).
 */
#ifndef ACD_DEV_H
#define ACD_DEV_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime (dev_runtime.hip) ---- */
int acd_dev_init(void);                 /* idempotent; 0 on success */
int acd_dev_set_device(int device);
int acd_dev_device(void);
void * acd_dev_malloc(size_t bytes);    /* zero-filled; NULL on failure */
void acd_dev_free(void * p);
int acd_dev_h2d(void * dst, const void * src, size_t bytes);
int acd_dev_d2h(void * dst, const void * src, size_t bytes);
int acd_dev_d2d(void * dst, const void * src, size_t bytes);
int acd_dev_memset(void * dst, int byte, size_t bytes);
int acd_dev_sync(void);
void * acd_dev_stream(void);            /* hipStream_t */
int acd_dev_set_stream(void * stream);
int acd_dev_cu_count(void);
/* First HIP error seen by the library since the last call (0 = none);
 * text copied into buf. */
int acd_dev_last_error(char * buf, size_t len);
void acd_dev_record_error(int code, const char * what);
/* Abort the process with a message if code != 0 (mirrors the original
 * library's assertion behaviour on unrecoverable states). */
void acd_dev_check(int code, const char * what);

/* ---- sliced ELL operator (one of A, A') ---- */
#define ACD_SLICE 64u      /* rows per slice = one wavefront */

struct acd_sell {
        uint32_t nrows;        /* rows of the operator */
        uint32_t ncols;        /* columns of the operator */
        uint32_t nslices;      /* ceil(nrows / 64) */
        uint32_t max_len;      /* longest row */
        uint64_t nent;         /* stored entries, padding included */
        uint64_t * off;        /* device [nslices+1]: first entry of slice */
        uint32_t * col;        /* device [nent] */
        double * val;          /* device [nent] */
        float * val32;         /* device [nent] or NULL (mixed precision) */
        uint32_t * rowlen;     /* device [nrows]: entries per row */
};

/* ---- preprocessing (dev_preprocess.hip) ---- */
/* Device-side matrix build.  Inputs are host triplets; outputs:
 *   - row/col permutation idx/ridx (device and host copies),
 *   - A and A' in sliced ELL (device),
 *   - the renumbered triplets sorted in Z-order, copied back to the host
 *     arrays rows/cols/vals (which hold the input on entry).
 * Returns 0 on success. */
struct acd_dev_matrix {
        uint32_t nrows, ncols;
        uint64_t nnz;
        uint32_t * row_idx, * row_ridx;   /* device [nrows] */
        uint32_t * col_idx, * col_ridx;   /* device [ncols] */
        struct acd_sell a, at;
};
int acd_dev_matrix_build(struct acd_dev_matrix * out,
                         uint32_t nrows, uint32_t ncols, uint64_t nnz,
                         uint32_t * rows, uint32_t * cols, double * vals,
                         int permute,
                         uint32_t * h_row_idx, uint32_t * h_row_ridx,
                         uint32_t * h_col_idx, uint32_t * h_col_ridx);
void acd_dev_matrix_free(struct acd_dev_matrix *);
/* FP32 copy of the stored values (exact for 0/+-1 matrices). */
int acd_dev_sell_make_fp32(struct acd_sell *);

/* ---- vector kernels (dev_spmv.hip) ---- */
/* y = op x on device pointers (op = A' when transpose). */
int acd_dev_spmv(const struct acd_dev_matrix *, int transpose,
                 double * y, const double * x);
int acd_dev_spmv2(const struct acd_dev_matrix *, int transpose,
                  double * y0, double * y1, const double * x0,
                  const double * x1);
/* dest[i] = src[p[i]] over n doubles. */
int acd_dev_gather(double * dest, const double * src, const uint32_t * p,
                   uint32_t n);
/* Time `reps` back-to-back launches of the FP64 SpMV kernel with HIP
 * events on the library stream; returns mean milliseconds per launch. */
int acd_dev_spmv_time(const struct acd_dev_matrix *, int transpose,
                      double * y, const double * x, int reps,
                      double * OUT_ms);

#ifdef __cplusplus
}
#endif
#endif
