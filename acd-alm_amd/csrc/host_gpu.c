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

int acd_device_memory_info(size_t * OUT_free, size_t * OUT_total)
{
        return acd_dev_mem_info(OUT_free, OUT_total);
}

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

int acd_copy_device(void * dst, const void * src, size_t bytes)
{
        if (bytes % 16 == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0)
                return acd_dev_copy_kernel(dst, src, bytes) || acd_dev_sync();
        return acd_dev_d2d(dst, src, bytes) || acd_dev_sync();
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

int acd_spmv_rotate_time(const sparse_matrix_t * const * mats, size_t k,
                         int transpose, double * const * d_y,
                         const double * const * d_x, int reps,
                         double * OUT_ms)
{
        if (k == 0 || k > 64) return -1;
        const struct acd_dev_matrix * dm[64];
        for (size_t i = 0; i < k; i++) dm[i] = &mats[i]->dev;
        return acd_dev_spmv_time_rotate(dm, k, transpose, d_y, d_x, reps,
                                        OUT_ms);
}

uint64_t acd_matrix_stored_entries(const sparse_matrix_t * a, int transpose)
{
        return transpose ? a->dev.at.nent : a->dev.a.nent;
}

int acd_spmv_variant_time(const sparse_matrix_t * a, int transpose,
                          double * d_y, const double * d_x, int reps,
                          int variant, double * OUT_ms)
{
        return acd_dev_spmv_time_variant(&a->dev, transpose, d_y, d_x, reps,
                                         variant, OUT_ms);
}

int acd_spmv_autotune(sparse_matrix_t * a, int reps, int * OUT_variants)
{
        return acd_dev_spmv_autotune(&a->dev, reps, OUT_variants);
}

int acd_spmv_set_variant(int variant)
{
        if (variant < 0 || variant >= ACD_SPMV_VARIANTS) return -1;
        return acd_dev_spmv_set_default_variant(variant);
}

void acd_last_solve_profile(uint64_t * counts, double * us, int n)
{
        acd_dev_last_profile(counts, us, n);
}

void acd_last_solve_bytes(double * step_us, double * bytes, int n)
{
        acd_dev_last_profile_ex(step_us, bytes, n);
}

const char * acd_pass_name(int p)
{
        static const char * names[] = {
                "none", "init_x", "violation", "pass1", "pass2", "fixup",
                "restart_copy", "checkpoint_x", "gradient_pg", "norm_g",
                "finish", "permute_out", "spg_grad0", "spg_grad", "spg_step",
                "alm_project", "alm_penalise", "permute_problem",
                "step_sizes", "alm_violation", "alm_argmax", "alm_update",
                "kkt", "pcd_step", "pcd_sizes", "kkt_checkpoint"};
        if (p < 0 || p >= (int)(sizeof names / sizeof names[0])) return "";
        return names[p];
}
