/* host_server.c -- continuous-batching request loop (acd_gpu.h, acd_server_*).
 *
 * Requests are ALM objects (made with alm_make, so their matrices and
 * vectors already live on the device) with caller-owned x / lambda /
 * diagnosis buffers.  They wait in a FIFO queue until their estimated
 * device working set fits the budget next to the requests already running;
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

enum { kRoundLaunches = 2, kRoundBudget = 32 };

struct request {
        alm_t * alm;
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
};

/* Device bytes of an LP: the caller's and the permuted matrix (A and A'
 * in sliced ELL, FP64 + FP32 values, u32 columns, ~30% padding), and the
 * vectors of both copies. */
static size_t footprint(const alm_t * alm)
{
        size_t nnz = sparse_matrix_nnz(alm->matrix);
        size_t rows = alm->nrhs, cols = alm->nvars;
        return 4 * (size_t)(1.3 * nnz) * 16 + (rows + cols) * 8 * 40;
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
        if (s->pool == NULL || !s->slot || !s->free_slots || !s->active
            || !s->queue || !s->large) {
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
        free(s);
}

int acd_server_submit(acd_server_t * s, alm_t * alm, double * x,
                      double * lambda, double * OUT_diagnosis, int * OUT_ret,
                      uint64_t tag)
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
        struct request r;
        r.alm = alm;
        r.x = x;
        r.lambda = lambda;
        r.diag = OUT_diagnosis;
        r.ret = OUT_ret;
        r.tag = tag;
        r.bytes = footprint(alm);
        r.large = alm->nrhs >= s->large_rows;
        s->queue[(s->qhead + s->qlen) % s->qcap] = r;
        s->qlen++;
        return 0;
}

size_t acd_server_pending(const acd_server_t * s)
{
        return s->qlen + s->nactive + s->nlarge;
}

/* A request whose solve failed (non-finite iterate, positive model
 * estimate) completes with a negative return code; the loop goes on. */
static void finish(struct request * r, const struct acd_solve_result * res)
{
        acd_alm_download(r->alm, r->x, r->lambda, 1);
        if (r->diag != NULL)
                for (int k = 0; k < 3; k++) r->diag[k] = res->diag[k];
        if (r->ret != NULL) *r->ret = res->error ? -res->error : res->ret;
}

/* FIFO admission while the head fits: a free slot (small) or a large lane
 * (at most nlarge_cap), and the budget (always admitted when nothing runs,
 * so one request larger than the budget still makes progress). */
static int admit(acd_server_t * s)
{
        while (s->qlen) {
                struct request * r = &s->queue[s->qhead];
                int idle = s->nactive == 0 && s->nlarge == 0;
                if (!idle && s->used + r->bytes > s->budget) break;
                if (r->large ? s->nlarge == s->nlarge_cap : s->nfree == 0)
                        break;
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
