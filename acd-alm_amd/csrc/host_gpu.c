/* host_gpu.c -- acd_gpu.h entry points that only marshal to the device
 * layer (device selection, errors, device-resident products). */
#include <stdlib.h>
#include "acd_host.h"

int acd_set_device(int device) { return acd_dev_set_device(device); }

int acd_last_error(char * buf, size_t len)
{
        return acd_dev_last_error(buf, len);
}

int acd_synchronize(void) { return acd_dev_sync(); }

void * acd_device_alloc(size_t bytes) { return acd_dev_malloc(bytes); }
void acd_device_free(void * p) { acd_dev_free(p); }

int acd_copy_to_device(void * dst, const void * src, size_t bytes)
{
        return acd_dev_h2d(dst, src, bytes);
}

int acd_copy_to_host(void * dst, const void * src, size_t bytes)
{
        return acd_dev_d2h(dst, src, bytes);
}

int acd_spmv_device(const sparse_matrix_t * a, int transpose, double * d_y,
                    const double * d_x)
{
        return acd_dev_spmv(&a->dev, transpose, d_y, d_x);
}

int acd_spmv_device_time(const sparse_matrix_t * a, int transpose,
                         double * d_y, const double * d_x, int reps,
                         double * OUT_ms)
{
        return acd_dev_spmv_time(&a->dev, transpose, d_y, d_x, reps, OUT_ms);
}

uint64_t acd_matrix_stored_entries(const sparse_matrix_t * a, int transpose)
{
        return transpose ? a->dev.at.nent : a->dev.a.nent;
}
