/* thread_pool.h -- host worker pool handle carried through every solver
 * signature.
 *
 * Mirrors thread_pool/thread_pool.h:13-89 of pkhuong/acd-alm.  The GPU
 * solvers accept a pool argument for signature compatibility and ignore
 * it: the device grid replaces the pool on every compute path.  The pool
 * itself is a working pthread pool (parallel-for and map-reduce over index
 * ranges) so that host code written against the original interface keeps
 * running unchanged.
 */
#ifndef ACD_THREAD_POOL_H
#define ACD_THREAD_POOL_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NULL is a valid pool everywhere and means "run serially". */
typedef struct thread_pool thread_pool_t;

/* Create a pool of nworker workers; the calling thread is worker 0, so
 * nworker-1 threads are started.  nworker == 0 behaves like 1.
 * (thread_pool/thread_pool.h:13) */
thread_pool_t * thread_pool_init(unsigned nworker);
/* Stop the workers and release the pool; NULL is ignored.
 * (thread_pool/thread_pool.h:15) */
void thread_pool_free(thread_pool_t *);

/* Worker count, caller included. (thread_pool/thread_pool.h:19) */
size_t thread_pool_count(thread_pool_t *);
/* Zeroed, cache-line separated scratch of char_per_worker bytes for each
 * worker; returns an array of worker-count pointers owned by the pool.
 * A size of 0 releases the backing store.  (thread_pool/thread_pool.h:31) */
void * const * thread_pool_worker_storage(thread_pool_t *,
                                          size_t char_per_worker);
/* Same storage as one flat block; the padded per-worker stride is written
 * to OUT_aligned_size when non-NULL.  (thread_pool/thread_pool.h:36) */
void * thread_pool_worker_storage_flat(thread_pool_t *,
                                       size_t char_per_worker,
                                       size_t * OUT_aligned_size);

/* Park idle workers on a condition variable / bring them back to spinning.
 * Running a job wakes them implicitly.  (thread_pool/thread_pool.h:46-47) */
void thread_pool_sleep(thread_pool_t *);
void thread_pool_wakeup(thread_pool_t *);

/* Parallel loop over [from, end): function receives disjoint half-open
 * ranges whose lengths are multiples of granularity (except the last),
 * granularity being scaled up to give each worker about ten chunks.
 * (thread_pool/thread_pool.h:63-67) */
typedef void (*thread_pool_function)(size_t begin, size_t end, void * info,
                                     unsigned worker_id);
void thread_pool_for(thread_pool_t *,
                     size_t from, size_t end, size_t granularity,
                     thread_pool_function function, void * info);

/* Parallel loop whose chunks each return a double, combined with reducer;
 * initial_value must be the reducer's identity.
 * (thread_pool/thread_pool.h:80-89) */
typedef double (*thread_pool_map)(size_t begin, size_t end, void * info,
                                  unsigned worker_id);
enum thread_pool_reducer{THREAD_POOL_REDUCE_SUM,
                         THREAD_POOL_REDUCE_MAX,
                         THREAD_POOL_REDUCE_MIN};
double thread_pool_map_reduce(thread_pool_t *,
                              size_t from, size_t end, size_t granularity,
                              thread_pool_map function, void * info,
                              enum thread_pool_reducer reducer,
                              double initial_value);

#ifdef __cplusplus
}
#endif
#endif
