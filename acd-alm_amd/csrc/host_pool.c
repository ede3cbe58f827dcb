/* host_pool.c -- the thread_pool.h interface.
 *
 * Mirrors thread_pool/thread_pool.h:13-89 of pkhuong/acd-alm with the same
 * observable semantics (worker 0 is the caller, chunk sizes are multiples
 * of the granularity scaled to about ten chunks per worker --
 * thread_pool/thread_pool.c:296-302 --, NULL means serial), built on a
 * mutex/condition-variable job hand-off instead of spinning.  None of the
 * GPU solvers use it; it is here so that host code that creates pools and
 * passes them to the solvers keeps compiling and running.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "thread_pool/thread_pool.h"

struct job {
        size_t next_unused;
        atomic_size_t next;
        size_t end, step;
        thread_pool_function function;
        void * info;
};

struct thread_pool {
        unsigned nworkers;            /* caller included */
        pthread_t * threads;
        pthread_mutex_t lock;
        pthread_cond_t wake, done;
        struct job * job;
        unsigned long generation;
        unsigned active;
        int stop;
        void * storage;
        void ** storage_ptrs;
        size_t storage_stride;
};

struct worker_arg { thread_pool_t * pool; unsigned id; };

static void run_chunks(struct job * job, unsigned id)
{
        for (;;) {
                size_t b = atomic_fetch_add(&job->next, job->step);
                if (b >= job->end) return;
                size_t e = b + job->step;
                if (e > job->end || e < b) e = job->end;
                job->function(b, e, job->info, id);
        }
}

static void * worker_main(void * p)
{
        struct worker_arg * arg = p;
        thread_pool_t * pool = arg->pool;
        unsigned id = arg->id;
        free(arg);
        unsigned long seen = 0;
        pthread_mutex_lock(&pool->lock);
        for (;;) {
                while (!pool->stop && pool->generation == seen)
                        pthread_cond_wait(&pool->wake, &pool->lock);
                if (pool->stop) break;
                seen = pool->generation;
                struct job * job = pool->job;
                pthread_mutex_unlock(&pool->lock);
                run_chunks(job, id);
                pthread_mutex_lock(&pool->lock);
                if (--pool->active == 0) pthread_cond_signal(&pool->done);
        }
        pthread_mutex_unlock(&pool->lock);
        return NULL;
}

thread_pool_t * thread_pool_init(unsigned nworker)
{
        if (nworker == 0) nworker = 1;
        thread_pool_t * pool = calloc(1, sizeof(*pool));
        if (pool == NULL) return NULL;
        pool->nworkers = nworker;
        pthread_mutex_init(&pool->lock, NULL);
        pthread_cond_init(&pool->wake, NULL);
        pthread_cond_init(&pool->done, NULL);
        pool->threads = calloc(nworker, sizeof(pthread_t));
        for (unsigned i = 1; i < nworker; i++) {
                struct worker_arg * arg = malloc(sizeof(*arg));
                arg->pool = pool;
                arg->id = i;
                pthread_create(&pool->threads[i], NULL, worker_main, arg);
        }
        return pool;
}

void thread_pool_free(thread_pool_t * pool)
{
        if (pool == NULL) return;
        pthread_mutex_lock(&pool->lock);
        pool->stop = 1;
        pthread_cond_broadcast(&pool->wake);
        pthread_mutex_unlock(&pool->lock);
        for (unsigned i = 1; i < pool->nworkers; i++)
                pthread_join(pool->threads[i], NULL);
        free(pool->threads);
        free(pool->storage);
        free(pool->storage_ptrs);
        pthread_mutex_destroy(&pool->lock);
        pthread_cond_destroy(&pool->wake);
        pthread_cond_destroy(&pool->done);
        free(pool);
}

size_t thread_pool_count(thread_pool_t * pool)
{
        return pool ? pool->nworkers : 1;
}

void * thread_pool_worker_storage_flat(thread_pool_t * pool, size_t bytes,
                                       size_t * OUT_aligned_size)
{
        if (pool == NULL) return NULL;
        size_t stride = (bytes + 63) & ~(size_t)63;
        free(pool->storage);
        pool->storage = NULL;
        pool->storage_stride = stride;
        if (stride) {
                if (posix_memalign(&pool->storage, 64,
                                   stride * pool->nworkers))
                        pool->storage = NULL;
                else
                        memset(pool->storage, 0, stride * pool->nworkers);
        }
        if (OUT_aligned_size) *OUT_aligned_size = stride;
        return pool->storage;
}

void * const * thread_pool_worker_storage(thread_pool_t * pool, size_t bytes)
{
        if (pool == NULL) return NULL;
        char * flat = thread_pool_worker_storage_flat(pool, bytes, NULL);
        free(pool->storage_ptrs);
        pool->storage_ptrs = NULL;
        if (flat == NULL) return NULL;
        pool->storage_ptrs = calloc(pool->nworkers, sizeof(void *));
        for (unsigned i = 0; i < pool->nworkers; i++)
                pool->storage_ptrs[i] = flat + i * pool->storage_stride;
        return pool->storage_ptrs;
}

void thread_pool_sleep(thread_pool_t * pool) { (void)pool; }
void thread_pool_wakeup(thread_pool_t * pool) { (void)pool; }

/* thread_pool/thread_pool.c:296-302 */
static size_t ideal_granularity(size_t n, size_t minimum, unsigned nthreads)
{
        if (minimum == 0) minimum = 1;
        n = (n + minimum - 1) / minimum;
        size_t nchunks = (size_t)nthreads * 10;
        if (n < nchunks) return minimum;
        return minimum * ((n + nchunks - 1) / nchunks);
}

void thread_pool_for(thread_pool_t * pool, size_t from, size_t end,
                     size_t granularity, thread_pool_function function,
                     void * info)
{
        if (pool == NULL || pool->nworkers <= 1
            || end - from <= granularity) {
                function(from, end, info, 0);
                return;
        }
        struct job job;
        atomic_init(&job.next, from);
        job.end = end;
        job.step = ideal_granularity(end - from, granularity, pool->nworkers);
        job.function = function;
        job.info = info;
        pthread_mutex_lock(&pool->lock);
        pool->job = &job;
        pool->active = pool->nworkers - 1;
        pool->generation++;
        pthread_cond_broadcast(&pool->wake);
        pthread_mutex_unlock(&pool->lock);
        run_chunks(&job, 0);
        pthread_mutex_lock(&pool->lock);
        while (pool->active) pthread_cond_wait(&pool->done, &pool->lock);
        pool->job = NULL;
        pthread_mutex_unlock(&pool->lock);
}

struct mr { thread_pool_map function; void * info;
            enum thread_pool_reducer reducer; double * acc; };

static void mr_worker(size_t b, size_t e, void * p, unsigned id)
{
        struct mr * mr = p;
        double v = mr->function(b, e, mr->info, id);
        double * a = (double *)((char *)mr->acc + id * 64);
        switch (mr->reducer) {
        case THREAD_POOL_REDUCE_SUM: *a += v; break;
        case THREAD_POOL_REDUCE_MAX: *a = fmax(*a, v); break;
        default: *a = fmin(*a, v); break;
        }
}

double thread_pool_map_reduce(thread_pool_t * pool, size_t from, size_t end,
                              size_t granularity, thread_pool_map function,
                              void * info, enum thread_pool_reducer reducer,
                              double initial_value)
{
        if (pool == NULL) return function(from, end, info, 0);
        unsigned n = pool->nworkers;
        double * acc = NULL;
        if (posix_memalign((void **)&acc, 64, 64 * (size_t)n)) abort();
        for (unsigned i = 0; i < n; i++)
                *(double *)((char *)acc + 64 * i) = initial_value;
        struct mr mr = {function, info, reducer, acc};
        thread_pool_for(pool, from, end, granularity, mr_worker, &mr);
        double r = initial_value;
        for (unsigned i = 0; i < n; i++) {
                double v = *(double *)((char *)acc + 64 * i);
                switch (reducer) {
                case THREAD_POOL_REDUCE_SUM: r += v; break;
                case THREAD_POOL_REDUCE_MAX: r = fmax(r, v); break;
                default: r = fmin(r, v); break;
                }
        }
        free(acc);
        return r;
}
