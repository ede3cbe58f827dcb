/* host_server.c -- continuous-batching request loop (acd_gpu.h, acd_server_*).
 *
 * Requests are either LPs given as host arrays (acd_server_submit_lp: the
 * matrix and the ALM object are built on the device when the request is
 * admitted, and freed when it retires, so device memory follows the
 * admitted set) or ALM objects the caller already made
 * (acd_server_submit), each with caller-owned x / lambda / diagnosis
 * buffers.  They wait in a FIFO queue until their estimated device
 * footprint fits the byte budget next to the requests already running;
 * then small LPs take a slot of the batched kernel (one workgroup each,
 * advancing in rounds of a few launches) and large LPs run the multi-
 * workgroup solver on their own stream, polled without blocking.  Work of
 * both kinds overlaps on the device; a request that arrives while others
 * run is admitted at the next round.  Each LP's arithmetic depends only on
 * that LP, so its result is the same alone and in any mix.
 */
#include <stdlib.h>
#include <string.h>
#include "acd_host.h"
#include "acd_approx.h"

/* A round: 2 batched launches of up to 128 inner iterations per LP.
 * Measured on Poisson arrivals of batched-workload LPs with 1% large LPs
 * (200 requests/s for 2 s; 2,000/s for 1 s): budget 32 gave p50 5.9 s /
 * 20.9 s, 128 gave 2.7 s / 13.6 s at the best throughput (40 / 68 solves/s);
 * 256 and 1,024 lowered the small LPs' p50 a little more but starved the
 * large lane, whose chunks are relaunched only between rounds (large p50
 * 9.6 and 10.8 s against 8.2 s). */
enum { kRoundLaunches = 2, kRoundBudget = 128 };
/* A large-lane request also holds its own stream, its solver graphs and
 * pinned state (device memory the allocator hands out in 2 MiB blocks). */
#define LARGE_LANE_BYTES ((size_t)8 << 20)

struct request {
        alm_t * alm;
        const struct acd_lp_data * data;   /* host LP, built at admission */
        sparse_matrix_t * own_matrix;      /* built here (data != NULL) */
        double * x, * lambda, * diag;
        int * ret;
        uint64_t tag;
        size_t bytes;
        int large;
};

struct acd_server {
        struct acd_alm_options opt;
        struct acd_solve_request q;
        size_t budget, used;
        size_t large_rows;
        struct acd_dev_pool * pool;
        uint32_t cap;
        struct request * slot;          /* [cap] */
        uint32_t * free_slots, nfree;
        uint32_t * active, nactive;
        struct request * queue;         /* ring */
        size_t qcap, qhead, qlen;
        struct request * large;         /* [nlarge_cap] running large LPs */
        size_t nlarge, nlarge_cap;
        uint64_t completed;
        uint64_t * dropped;             /* tags that failed to build */
        size_t ndropped, dropped_cap;
        size_t free0;                   /* device free bytes at make */
        void * build_stream;            /* device builds of admissions */
};

/* Device bytes of an LP of m rows, n columns and nnz entries, an upper
 * estimate: the caller's and the permuted matrix (A and A' in sliced ELL
 * with up to 50% padding: FP64, FP32 and BF16 values and u32 columns, 18
 * bytes per stored entry), their permutation and row-length arrays, the
 * solver's 30 vectors (14 m-side, 16 n-side, FP64), and the fixed
 * per-LP records (event ring, partial sums, state). */
size_t acd_server_footprint(size_t m, size_t n, size_t nnz)
{
        const size_t raw = (size_t)(4 * 1.5 * 18) * nnz + 24 * (m + n)
                           + 8 * (14 * m + 16 * n) + ((size_t)384 << 10);
        /* The device allocator hands out memory in 2 MiB blocks that it
         * sub-allocates; an LP's allocations may straddle one block more
         * than their size fills. */
        const size_t block = (size_t)2 << 20;
        return (raw + block - 1) / block * block + block;
}

static size_t footprint(const alm_t * alm)
{
        return acd_server_footprint(alm->nrhs, alm->nvars,
                                    sparse_matrix_nnz(alm->matrix));
}

acd_server_t * acd_server_make(size_t hbm_budget, size_t slots,
                               size_t large_rows,
                               const struct acd_alm_options * opt)
{
        acd_server_t * s = calloc(1, sizeof(*s));
        if (s == NULL) return NULL;
        if (opt) s->opt = *opt;
        else acd_alm_default_options(&s->opt);
        s->budget = hbm_budget;
        s->large_rows = large_rows;
        s->cap = (uint32_t)(slots ? slots : 2048);
        s->pool = acd_dev_pool_create(s->cap, s->opt.precision);
        s->slot = calloc(s->cap, sizeof(*s->slot));
        s->free_slots = malloc(s->cap * sizeof(uint32_t));
        s->active = malloc(s->cap * sizeof(uint32_t));
        s->qcap = 1024;
        s->queue = malloc(s->qcap * sizeof(*s->queue));
        s->nlarge_cap = 4;
        s->large = calloc(s->nlarge_cap, sizeof(*s->large));
        s->dropped_cap = 1024;
        s->dropped = malloc(s->dropped_cap * sizeof(uint64_t));
        if (acd_dev_mem_info(&s->free0, NULL)) s->free0 = 0;
        s->build_stream = acd_dev_stream_create();
        if (s->pool == NULL || !s->slot || !s->free_slots || !s->active
            || !s->queue || !s->large || !s->dropped || !s->build_stream) {
                acd_server_free(s);
                return NULL;
        }
        for (uint32_t i = 0; i < s->cap; i++)
                s->free_slots[i] = s->cap - 1 - i;
        s->nfree = s->cap;
        memset(&s->q, 0, sizeof s->q);
        s->q.mode = ACD_MODE_ALM;
        s->q.niter = 1000;
        s->q.stop_violation = s->opt.stop_violation;
        s->q.stop_pg = s->opt.stop_pg;
        s->q.precision_floor = s->opt.precision_floor;
        s->q.inner_min_delta = s->opt.min_delta;
        s->q.precision = s->opt.precision;
        s->q.stop_kkt = s->opt.stop_kkt;
        return s;
}

void acd_server_free(acd_server_t * s)
{
        if (s == NULL) return;
        acd_dev_pool_free(s->pool);
        free(s->slot);
        free(s->free_slots);
        free(s->active);
        free(s->queue);
        free(s->large);
        free(s->dropped);
        acd_dev_stream_destroy(s->build_stream);
        free(s);
}

static int enqueue(acd_server_t * s, const struct request * r)
{
        if (s->qlen == s->qcap) {
                size_t nc = 2 * s->qcap;
                struct request * nq = malloc(nc * sizeof(*nq));
                if (nq == NULL) return -1;
                for (size_t i = 0; i < s->qlen; i++)
                        nq[i] = s->queue[(s->qhead + i) % s->qcap];
                free(s->queue);
                s->queue = nq;
                s->qcap = nc;
                s->qhead = 0;
        }
        s->queue[(s->qhead + s->qlen) % s->qcap] = *r;
        s->qlen++;
        return 0;
}

int acd_server_submit(acd_server_t * s, alm_t * alm, double * x,
                      double * lambda, double * OUT_diagnosis, int * OUT_ret,
                      uint64_t tag)
{
        struct request r;
        memset(&r, 0, sizeof r);
        r.alm = alm;
        r.x = x;
        r.lambda = lambda;
        r.diag = OUT_diagnosis;
        r.ret = OUT_ret;
        r.tag = tag;
        r.large = alm->nrhs >= s->large_rows;
        r.bytes = footprint(alm) + (r.large ? LARGE_LANE_BYTES : 0);
        return enqueue(s, &r);
}

int acd_server_submit_lp(acd_server_t * s, const struct acd_lp_data * lp,
                         double * x, double * lambda, double * OUT_diagnosis,
                         int * OUT_ret, uint64_t tag)
{
        if (lp == NULL || (lp->nnz && (!lp->rows || !lp->cols || !lp->values))
            || (lp->nrows && !lp->rhs) || x == NULL || lambda == NULL)
                return -1;
        struct request r;
        memset(&r, 0, sizeof r);
        r.data = lp;
        r.x = x;
        r.lambda = lambda;
        r.diag = OUT_diagnosis;
        r.ret = OUT_ret;
        r.tag = tag;
        r.large = lp->nrows >= s->large_rows;
        r.bytes = acd_server_footprint(lp->nrows, lp->ncols, lp->nnz)
                  + (r.large ? LARGE_LANE_BYTES : 0);
        return enqueue(s, &r);
}

size_t acd_server_bytes_admitted(const acd_server_t * s) { return s->used; }

size_t acd_server_pending(const acd_server_t * s)
{
        return s->qlen + s->nactive + s->nlarge + s->ndropped;
}

/* A request whose solve failed (non-finite iterate, positive model
 * estimate) completes with a negative return code; the loop goes on. */
static void finish(struct request * r, const struct acd_solve_result * res)
{
        acd_alm_download(r->alm, r->x, r->lambda, 1);
        if (r->diag != NULL)
                for (int k = 0; k < 3; k++) r->diag[k] = res->diag[k];
        if (r->ret != NULL) *r->ret = res->error ? -res->error : res->ret;
        if (r->own_matrix != NULL) {   /* built at admission: release it */
                alm_free(r->alm);
                sparse_matrix_free(r->own_matrix);
                r->alm = NULL;
                r->own_matrix = NULL;
        }
}

/* Builds a host-data request's matrix and ALM object on the device, on
 * the server's build stream so the build overlaps the batched launches in
 * flight on the library stream (every step of the build completes before
 * it returns); -1 when the device cannot hold them. */
static int build_on(struct request * r);
static int build(acd_server_t * s, struct request * r)
{
        void * main = acd_dev_stream();
        acd_dev_set_stream(s->build_stream);
        int rc = build_on(r);
        acd_dev_set_stream(main);
        return rc;
}

static int build_on(struct request * r)
{
        const struct acd_lp_data * d = r->data;
        sparse_matrix_t * m = sparse_matrix_make(d->ncols, d->nrows, d->nnz,
                                                 d->rows, d->cols, d->values,
                                                 0);
        if (m == NULL) return -1;
        alm_t * a = alm_make(m, d->nrows, d->rhs, d->ncols, d->linear,
                             d->lower, d->upper, d->lambda_lower,
                             d->lambda_upper);
        if (a == NULL) {
                sparse_matrix_free(m);
                return -1;
        }
        r->own_matrix = m;
        r->alm = a;
        return 0;
}

/* Device bytes in use beyond what was in use when the server was made
 * (hipMemGetInfo: the allocator's real blocks, whoever holds them). */
static size_t measured_use(const acd_server_t * s)
{
        size_t f = 0;
        if (s->free0 == 0 || acd_dev_mem_info(&f, NULL)) return 0;
        return f < s->free0 ? s->free0 - f : 0;
}

/* FIFO admission while the head fits: a free slot (small) or a large lane
 * (at most nlarge_cap), and the budget: the admitted requests' estimated
 * footprints, or the device memory measurably in use if that is more,
 * plus the head's estimate (always admitted when nothing runs, so one
 * request larger than the budget still makes progress). */
static int admit(acd_server_t * s)
{
        while (s->qlen) {
                struct request * r = &s->queue[s->qhead];
                int idle = s->nactive == 0 && s->nlarge == 0;
                if (!idle) {
                        size_t in_use = measured_use(s);
                        if (in_use < s->used) in_use = s->used;
                        if (in_use + r->bytes > s->budget) break;
                }
                if (r->large ? s->nlarge == s->nlarge_cap : s->nfree == 0)
                        break;
                if (r->data != NULL && build(s, r)) {
                        /* not buildable: completes with an error */
                        if (r->ret != NULL) *r->ret = -ACD_SOLVE_NO_MEMORY;
                        if (s->ndropped < s->dropped_cap)
                                s->dropped[s->ndropped++] = r->tag;
                        s->qhead = (s->qhead + 1) % s->qcap;
                        s->qlen--;
                        continue;
                }
                acd_alm_upload(r->alm, r->x, r->lambda);
                struct acd_dev_lp * lp = r->alm->approx->dev;
                if (r->large) {
                        if (acd_dev_lp_async_begin(lp, &s->q)) return -1;
                        s->large[s->nlarge++] = *r;
                } else {
                        uint32_t slot = s->free_slots[--s->nfree];
                        if (acd_dev_pool_admit(s->pool, slot, lp, &s->q))
                                return -1;
                        s->slot[slot] = *r;
                        s->active[s->nactive++] = slot;
                }
                s->used += r->bytes;
                s->qhead = (s->qhead + 1) % s->qcap;
                s->qlen--;
        }
        return 0;
}

int acd_server_step(acd_server_t * s, uint64_t * OUT_done, size_t cap,
                    size_t * OUT_ndone)
{
        size_t nd = 0;
        if (admit(s)) return -1;
        /* requests that could not be built complete now, with their error */
        for (size_t i = 0; i < s->ndropped; i++) {
                if (nd < cap && OUT_done) OUT_done[nd] = s->dropped[i];
                nd++;
                s->completed++;
        }
        s->ndropped = 0;
        /* Only large LPs running: wait for the first one's chunk instead of
         * returning at once, so callers stepping in a loop do not spin. */
        if (s->nactive == 0 && s->nlarge > 0
            && acd_dev_lp_async_wait(s->large[0].alm->approx->dev))
                return -1;
        /* large LPs: launch their next chunks (non-blocking) */
        for (size_t i = 0; i < s->nlarge;) {
                struct acd_solve_result res;
                int st = acd_dev_lp_async_poll(s->large[i].alm->approx->dev,
                                               &res);
                if (st < 0) return -1;
                if (st == 1) {
                        finish(&s->large[i], &res);
                        if (nd < cap && OUT_done) OUT_done[nd] = s->large[i].tag;
                        nd++;
                        s->used -= s->large[i].bytes;
                        s->large[i] = s->large[--s->nlarge];
                        s->completed++;
                        continue;
                }
                i++;
        }
        /* one round of the batched kernel over the active slots */
        if (s->nactive) {
                uint32_t * flags = NULL;
                if (acd_dev_pool_round(s->pool, s->active, s->nactive,
                                       kRoundLaunches, kRoundBudget, &flags))
                        return -1;
                for (uint32_t i = 0; i < s->nactive;) {
                        uint32_t slot = s->active[i];
                        if (!flags[slot]) {
                                i++;
                                continue;
                        }
                        struct acd_solve_result res;
                        if (acd_dev_pool_result(s->pool, slot, &res))
                                return -1;
                        finish(&s->slot[slot], &res);
                        if (nd < cap && OUT_done) OUT_done[nd] = s->slot[slot].tag;
                        nd++;
                        s->used -= s->slot[slot].bytes;
                        s->free_slots[s->nfree++] = slot;
                        s->active[i] = s->active[--s->nactive];
                        s->completed++;
                }
        }
        if (OUT_ndone) *OUT_ndone = nd;
        return 0;
}
