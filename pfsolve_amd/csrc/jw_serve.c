/*
 * jw_serve.c -- a serving loop for fused JW steps (pfsolve_amd.h, pfsolve_serve).
 *
 * Requests (bundles of real systems of one size N, with host input and output)
 * arrive over time.  The loop
 *   - admits arrived requests in arrival order while the device budget (one
 *     arena allocated at start, right-hand sides of requests in flight plus
 *     resident matrices) and the in-flight request limit allow;
 *   - keeps matrices resident per key: the first request of a key uploads its
 *     matrices and writes its coefficient table once, on its launch stream;
 *     later requests of the key (on any stream) wait for that event only;
 *   - merges every admitted, not yet launched request into one launch on the
 *     next free stream slot: host->device copies of the bundles, one persistent
 *     queue kernel per row class over all bundles of the launch (the kernels of
 *     the uneven-size batch), device->host copies, an event;
 *   - polls the slot events; a finished launch returns its memory and slots,
 *     so requests that arrived meanwhile are admitted and launched while other
 *     launches are still on the device (continuous batching).
 * Latency of a request = completion time - arrival time.
 * No counterpart in the reference (which has no request loop).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "jw_plan.h"

#define NCLS 5            /* row classes R = 2, 4, 8, 16, 32 */
#define ALIGN 256
enum { kUnitsPerReq = 64 };
/* no admission, launch or completion for this long while work is outstanding:
 * an error, not a wait (real-time runs sleep until the next arrival instead) */
static const double kStallLimitS = 30.0;

static int cmp_double(const void* a, const void* b)
{
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int row_class(int64_t N)
{
    int c = 0;
    while (c < NCLS - 1 && (int64_t)(128 << c) < N) c++;
    return c;
}

/* ---- arena: first-fit over a sorted free list, coalescing on free ---- */
typedef struct blk { int64_t off, len; } blk;
typedef struct arena {
    char* base;
    int64_t size, used, peak;
    blk* fr;              /* free blocks sorted by offset */
    int nfree, cap;
} arena;

static int arena_init(arena* a, int64_t size)
{
    memset(a, 0, sizeof(*a));
    if (hipMalloc((void**)&a->base, (size_t)size) != hipSuccess) return -1;
    a->size = size;
    a->cap = 64;
    a->fr = (blk*)malloc(sizeof(blk) * (size_t)a->cap);
    if (!a->fr) return -1;
    a->fr[0].off = 0;
    a->fr[0].len = size;
    a->nfree = 1;
    return 0;
}

static void arena_free_all(arena* a)
{
    if (a->base) hipFree(a->base);
    free(a->fr);
    memset(a, 0, sizeof(*a));
}

static int64_t arena_alloc(arena* a, int64_t len)
{
    len = (len + ALIGN - 1) / ALIGN * ALIGN;
    for (int i = 0; i < a->nfree; i++) {
        if (a->fr[i].len < len) continue;
        const int64_t off = a->fr[i].off;
        a->fr[i].off += len;
        a->fr[i].len -= len;
        if (a->fr[i].len == 0) {
            memmove(&a->fr[i], &a->fr[i + 1], sizeof(blk) * (size_t)(a->nfree - i - 1));
            a->nfree--;
        }
        a->used += len;
        if (a->used > a->peak) a->peak = a->used;
        return off;
    }
    return -1;
}

static int arena_release(arena* a, int64_t off, int64_t len)
{
    len = (len + ALIGN - 1) / ALIGN * ALIGN;
    int i = 0;
    while (i < a->nfree && a->fr[i].off < off) i++;
    if (a->nfree + 1 > a->cap) {
        blk* n = (blk*)realloc(a->fr, sizeof(blk) * (size_t)(a->cap * 2));
        if (!n) return -1;
        a->fr = n;
        a->cap *= 2;
    }
    memmove(&a->fr[i + 1], &a->fr[i], sizeof(blk) * (size_t)(a->nfree - i));
    a->fr[i].off = off;
    a->fr[i].len = len;
    a->nfree++;
    /* coalesce with the neighbours */
    if (i + 1 < a->nfree && a->fr[i].off + a->fr[i].len == a->fr[i + 1].off) {
        a->fr[i].len += a->fr[i + 1].len;
        memmove(&a->fr[i + 1], &a->fr[i + 2], sizeof(blk) * (size_t)(a->nfree - i - 2));
        a->nfree--;
    }
    if (i > 0 && a->fr[i - 1].off + a->fr[i - 1].len == a->fr[i].off) {
        a->fr[i - 1].len += a->fr[i].len;
        memmove(&a->fr[i], &a->fr[i + 1], sizeof(blk) * (size_t)(a->nfree - i - 1));
        a->nfree--;
    }
    a->used -= len;
    return 0;
}

/* ---- resident matrices per key (open addressing) ---- */
typedef struct resident {
    uint64_t key;
    int used;
    int64_t N;
    int R;
    int64_t off, bytes;   /* arena block: matrices (10 N doubles, padded) then the table */
    void* table;
    hipEvent_t ready;
} resident;

typedef struct keymap {
    resident* e;
    int64_t cap, n;
} keymap;

static resident* key_find(keymap* m, uint64_t key, int insert)
{
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    for (int64_t i = 0; i < m->cap; i++) {
        resident* r = &m->e[(h + (uint64_t)i) % (uint64_t)m->cap];
        if (r->used && r->key == key) return r;
        if (!r->used) {
            if (!insert) return NULL;
            r->used = 1;
            r->key = key;
            m->n++;
            return r;
        }
    }
    return NULL;
}

/* ---- launch slots ---- */
typedef struct slot {
    hipStream_t stream;
    hipEvent_t done;
    int busy;
    int64_t* reqs;            /* request indices of the launch */
    int nreqs;
    jw_batch_item* h_items;   /* pinned */
    jw_batch_unit* h_units;   /* pinned */
    jw_batch_item* d_items;
    jw_batch_unit* d_units;
    int* d_counters;
} slot;

typedef struct reqstate {
    int64_t off, bytes;       /* device block of the bundle, -1 when not admitted */
    int64_t stride;
    int state;                /* 0 waiting, 1 admitted, 2 launched, 3 done, 4 rejected */
} reqstate;

static PfSolveResult check_request(const pfsolve_serve_request* q)
{
    if (q->N < 2 || q->N > 2048) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    if (q->nsys < 0) return PFSOLVE_ERROR_EMPTY_size;
    if (!q->matrices) return PFSOLVE_ERROR_EMPTY_buffer;
    if (q->offsetV < 0 || q->offsetT < 0 || q->offsetM < 0) return PFSOLVE_ERROR_EMPTY_bufferSize;
    if (q->KL < 0 || q->KL > 2 || q->KU < 0 || q->KU > 2) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
    if (q->nsys > 0 && (!q->rhs_in || !q->rhs_out)) return PFSOLVE_ERROR_EMPTY_outputBuffer;
    return PFSOLVE_SUCCESS;
}

static int64_t matrix_extent(const pfsolve_serve_request* q)
{
    /* largest element index the prologue reads, + 1 (V band-major: KL + KU + 1 diagonals) */
    int64_t e = q->offsetV + (int64_t)(q->KL + q->KU + 1) * q->N;
    if (q->offsetT + 3 * q->N > e) e = q->offsetT + 3 * q->N;
    if (q->offsetM + 2 * q->N > e) e = q->offsetM + 2 * q->N;
    return e;
}

PfSolveResult pfsolve_serve(const pfsolve_serve_config* cfg, pfsolve_serve_request* reqs, int64_t nreq,
                            pfsolve_serve_stats* stats)
{
    if (!cfg || !stats || (nreq > 0 && !reqs)) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    memset(stats, 0, sizeof(*stats));
    if (cfg->hbm_budget < (1 << 20) || cfg->max_inflight < 1 || cfg->streams < 1 || cfg->streams > 16 ||
        cfg->max_batch < 1)
        return PFSOLVE_ERROR_EMPTY_size;
    for (int64_t i = 0; i < nreq; i++) {
        PfSolveResult r = check_request(&reqs[i]);
        if (r != PFSOLVE_SUCCESS) return r;
        if (i > 0 && reqs[i].arrival_s < reqs[i - 1].arrival_s) return PFSOLVE_ERROR_EMPTY_size;
    }
    if (nreq == 0) return PFSOLVE_SUCCESS;

    int cu = 0, lds = 0, maj = 0, min = 0;
    if (jwdev_device_info(&cu, &lds, &maj, &min) != 0) return PFSOLVE_ERROR_INVALID_DEVICE;

    PfSolveResult res = PFSOLVE_SUCCESS;
    arena A;
    memset(&A, 0, sizeof(A));
    keymap K;
    memset(&K, 0, sizeof(K));
    slot* S = (slot*)calloc((size_t)cfg->streams, sizeof(slot));
    reqstate* st = (reqstate*)calloc((size_t)nreq, sizeof(reqstate));
    int64_t* pending = (int64_t*)malloc(sizeof(int64_t) * (size_t)nreq);
    double* lat = (double*)malloc(sizeof(double) * (size_t)nreq);
    const int maxb = cfg->max_batch;
    /* units per launch: a request's passes are cut into at most kUnitsPerReq units */
    const int64_t max_units = (int64_t)maxb * kUnitsPerReq;
    if (!S || !st || !pending || !lat) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    if (arena_init(&A, cfg->hbm_budget) != 0) { res = PFSOLVE_ERROR_FAILED_TO_ALLOCATE; goto out; }
    K.cap = 4096;
    K.e = (resident*)calloc((size_t)K.cap, sizeof(resident));
    if (!K.e) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    for (int s = 0; s < cfg->streams; s++) {
        slot* x = &S[s];
        if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&x->done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&x->h_items, sizeof(jw_batch_item) * (size_t)maxb, 0) != hipSuccess ||
            hipHostMalloc((void**)&x->h_units, sizeof(jw_batch_unit) * (size_t)max_units, 0) != hipSuccess ||
            hipMalloc((void**)&x->d_items, sizeof(jw_batch_item) * (size_t)maxb) != hipSuccess ||
            hipMalloc((void**)&x->d_units, sizeof(jw_batch_unit) * (size_t)max_units) != hipSuccess ||
            hipMalloc((void**)&x->d_counters, sizeof(int) * NCLS) != hipSuccess) {
            res = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
            goto out;
        }
        x->reqs = (int64_t*)malloc(sizeof(int64_t) * (size_t)maxb);
        if (!x->reqs) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    }
    for (int64_t i = 0; i < nreq; i++) st[i].off = -1;

    /* launch shapes per class: the batch kernel's waves and workgroups per CU */
    int waves[NCLS], grid_max[NCLS];
    for (int c = 0; c < NCLS; c++) {
        const int R = 2 << c;
        waves[c] = jwdev_main64_waves(R);
        const int occ = waves[c] > 0 ? jwdev_batch64_occupancy(R, waves[c]) : 0;
        if (waves[c] < 1 || occ < 1) { res = PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY; goto out; }
        grid_max[c] = cu * occ;
    }

    const double t0 = now_s();
    int64_t next = 0, npend = 0, ndone = 0, inflight = 0, launches = 0, launched_reqs = 0;
    double last_progress = now_s();
    while (ndone < nreq) {
        int progress = 0;
        const double t = now_s() - t0;

        /* 1. admission in arrival order */
        while (next < nreq && (!cfg->realtime || reqs[next].arrival_s <= t) && inflight < cfg->max_inflight) {
            pfsolve_serve_request* q = &reqs[next];
            reqstate* r = &st[next];
            r->stride = q->N + (q->N & 1);
            r->bytes = q->nsys * r->stride * 8;
            /* the matrices of a new key are part of the budget too */
            resident* k = key_find(&K, q->matrix_key, 0);
            const int R = 2 << row_class(q->N);
            const int64_t mbytes = (matrix_extent(q) * 8 + ALIGN - 1) / ALIGN * ALIGN;
            const int64_t kbytes = k ? 0 : mbytes + jwdev_table64_elems(R) * 8;
            if (r->bytes + kbytes > A.size) {
                /* can never fit: rejected, not waited for */
                r->state = 4;
                q->status = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
                q->admit_s = q->done_s = t;
                ndone++;
                next++;
                progress = 1;
                continue;
            }
            if (!k) {
                const int64_t off = arena_alloc(&A, kbytes);
                if (off < 0) {
                    if (inflight == 0) goto reject_nomem;   /* resident keys leave no room: nothing will free */
                    break;
                }
                k = key_find(&K, q->matrix_key, 1);
                if (!k) { arena_release(&A, off, kbytes); res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
                k->N = q->N;
                k->R = R;
                k->off = off;
                k->bytes = kbytes;
                k->table = A.base + off + mbytes;
                k->ready = NULL;
                /* upload the matrices now; the table is written on the first launch stream */
                if (hipMemcpy(A.base + off, q->matrices, (size_t)(matrix_extent(q) * 8), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                    res = PFSOLVE_ERROR_FAILED_TO_COPY;
                    goto out;
                }
            } else if (k->N != q->N) {
                r->state = 4;   /* one key, one size */
                q->status = PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
                q->admit_s = q->done_s = t;
                ndone++;
                next++;
                progress = 1;
                continue;
            }
            if (r->bytes > 0) {
                const int64_t off = arena_alloc(&A, r->bytes);
                if (off < 0) {
                    if (inflight == 0) goto reject_nomem;
                    break;   /* wait for memory to free up */
                }
                r->off = off;
            }
            r->state = 1;
            q->admit_s = t;
            pending[npend++] = next;
            inflight++;
            next++;
            progress = 1;
            continue;
        reject_nomem:
            r->state = 4;
            q->status = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
            q->admit_s = q->done_s = t;
            ndone++;
            next++;
            progress = 1;
        }

        /* 2. one launch of the pending requests on a free stream slot */
        if (npend > 0) {
            int s = -1;
            for (int i = 0; i < cfg->streams; i++)
                if (!S[i].busy) { s = i; break; }
            if (s >= 0) {
                slot* x = &S[s];
                const int nb = npend < maxb ? (int)npend : maxb;
                x->nreqs = nb;
                int64_t nunits = 0;
                int64_t cls_first[NCLS], cls_n[NCLS];
                for (int c = 0; c < NCLS; c++) cls_n[c] = 0;
                /* resident tables: first use writes the table on this stream */
                for (int i = 0; i < nb; i++) {
                    const int64_t qi = pending[i];
                    pfsolve_serve_request* q = &reqs[qi];
                    resident* k = key_find(&K, q->matrix_key, 0);
                    if (!k->ready) {
                        jw_source_desc src;
                        memset(&src, 0, sizeof(src));
                        src.N = q->N;
                        src.band_kind = JW_BAND_GB;
                        src.offV = q->offsetV;
                        src.KL = q->KL;
                        src.KU = q->KU;
                        src.solve_kind = JW_SOLVE_TRI;
                        src.ldS = q->offsetT;
                        src.mdS = q->offsetT + q->N;
                        src.udS = q->offsetT + 2 * q->N;
                        src.has_m = 1;
                        src.mdM = q->offsetM;
                        src.udM = q->offsetM + q->N;
                        src.Msplit0 = 704;
                        src.permuted = !q->naturalOrder && q->N > 64;
                        if (jwdev_prologue(&src, A.base + k->off, k->table, 64, k->R, JW_PREC_F64, 2, x->stream) != 0 ||
                            hipEventCreateWithFlags(&k->ready, hipEventDisableTiming) != hipSuccess ||
                            hipEventRecord(k->ready, x->stream) != hipSuccess) {
                            res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                            goto out;
                        }
                    } else if (hipStreamWaitEvent(x->stream, k->ready, 0) != hipSuccess) {
                        res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                        goto out;
                    }
                }
                /* host -> device copies of the bundles (systems padded to an even stride) */
                for (int i = 0; i < nb; i++) {
                    const int64_t qi = pending[i];
                    pfsolve_serve_request* q = &reqs[qi];
                    reqstate* r = &st[qi];
                    if (q->nsys > 0 &&
                        hipMemcpy2DAsync(A.base + r->off, (size_t)(r->stride * 8), q->rhs_in, (size_t)(q->N * 8),
                                         (size_t)(q->N * 8), (size_t)q->nsys, hipMemcpyHostToDevice,
                                         x->stream) != hipSuccess) {
                        res = PFSOLVE_ERROR_FAILED_TO_COPY;
                        goto out;
                    }
                    resident* k = key_find(&K, q->matrix_key, 0);
                    jw_batch_item* it = &x->h_items[i];
                    memset(it, 0, sizeof(*it));
                    it->rhs = A.base + r->off;
                    it->table = k->table;
                    it->nsys = q->nsys;
                    it->stride = r->stride;
                    it->offSol = 0;
                    it->N = (int32_t)q->N;
                    it->kl = q->KL;
                    it->ku = q->KU;
                    it->apply_scale = q->scaleC != 1.0;
                    it->scale = q->scaleC;
                    x->reqs[i] = qi;
                }
                /* work units per row class, largest N first within a class */
                for (int c = NCLS - 1; c >= 0; c--) {
                    cls_first[c] = nunits;
                    for (int i = 0; i < nb; i++) {
                        const pfsolve_serve_request* q = &reqs[pending[i]];
                        if (row_class(q->N) != c) continue;
                        const int64_t passes = c == NCLS - 1 ? q->nsys : (q->nsys + 1) / 2;
                        /* units of >= 8 passes, at most kUnitsPerReq of them */
                        int64_t up = (passes + kUnitsPerReq - 1) / kUnitsPerReq;
                        if (up < 8) up = 8;
                        for (int64_t p = 0; p < passes; p += up) {
                            jw_batch_unit* u = &x->h_units[nunits++];
                            u->item = i;
                            u->pair0 = (int32_t)p;
                            u->pair1 = (int32_t)(p + up < passes ? p + up : passes);
                        }
                    }
                    cls_n[c] = nunits - cls_first[c];
                }
                if (hipMemcpyAsync(x->d_items, x->h_items, sizeof(jw_batch_item) * (size_t)nb, hipMemcpyHostToDevice,
                                   x->stream) != hipSuccess ||
                    (nunits > 0 && hipMemcpyAsync(x->d_units, x->h_units, sizeof(jw_batch_unit) * (size_t)nunits,
                                                  hipMemcpyHostToDevice, x->stream) != hipSuccess) ||
                    jwdev_zero_counters(x->d_counters, NCLS, x->stream) != 0) {
                    res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                    goto out;
                }
                for (int c = 0; c < NCLS; c++) {
                    if (!cls_n[c]) continue;
                    int64_t grid = cls_n[c] < grid_max[c] ? cls_n[c] : grid_max[c];
                    if (jwdev_batch64(x->d_items, x->d_units + cls_first[c], (int)cls_n[c], x->d_counters + c,
                                      2 << c, waves[c], (int)grid, x->stream) != 0) {
                        res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                        goto out;
                    }
                }
                for (int i = 0; i < nb; i++) {
                    const int64_t qi = pending[i];
                    pfsolve_serve_request* q = &reqs[qi];
                    reqstate* r = &st[qi];
                    if (q->nsys > 0 &&
                        hipMemcpy2DAsync(q->rhs_out, (size_t)(q->N * 8), A.base + r->off, (size_t)(r->stride * 8),
                                         (size_t)(q->N * 8), (size_t)q->nsys, hipMemcpyDeviceToHost,
                                         x->stream) != hipSuccess) {
                        res = PFSOLVE_ERROR_FAILED_TO_COPY;
                        goto out;
                    }
                    r->state = 2;
                }
                if (hipEventRecord(x->done, x->stream) != hipSuccess) { res = PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD; goto out; }
                x->busy = 1;
                launches++;
                launched_reqs += nb;
                memmove(pending, pending + nb, sizeof(int64_t) * (size_t)(npend - nb));
                npend -= nb;
                progress = 1;
            }
        }

        /* 3. completions */
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &S[s];
            if (!x->busy) continue;
            hipError_t e = hipEventQuery(x->done);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) { res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE; goto out; }
            const double td = now_s() - t0;
            for (int i = 0; i < x->nreqs; i++) {
                const int64_t qi = x->reqs[i];
                reqstate* r = &st[qi];
                if (r->bytes > 0) arena_release(&A, r->off, r->bytes);
                r->off = -1;
                r->state = 3;
                reqs[qi].done_s = td;
                reqs[qi].status = PFSOLVE_SUCCESS;
                inflight--;
                ndone++;
            }
            x->busy = 0;
            progress = 1;
        }
        if (progress) {
            last_progress = now_s();
        } else if (now_s() - last_progress > kStallLimitS) {
            /* a launch that never completes: give up instead of spinning for ever */
            res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
            goto out;
        }
        if (!progress) {
            /* nothing to do: the next arrival or a completion is ahead */
            if (cfg->realtime && next < nreq && npend == 0 && inflight == 0) {
                const double wait = reqs[next].arrival_s - (now_s() - t0);
                if (wait > 200e-6) usleep((useconds_t)((wait - 100e-6) * 1e6));
            } else {
                usleep(10);
            }
        }
    }

    {
        const double wall = now_s() - t0;
        int64_t nl = 0, nsys = 0, rejected = 0;
        for (int64_t i = 0; i < nreq; i++) {
            if (st[i].state != 3) { rejected++; continue; }
            lat[nl++] = reqs[i].done_s - (cfg->realtime ? reqs[i].arrival_s : 0.0);
            nsys += reqs[i].nsys;
        }
        qsort(lat, (size_t)nl, sizeof(double), cmp_double);
        stats->requests = nl;
        stats->rejected = rejected;
        stats->systems = nsys;
        stats->launches = launches;
        stats->wall_s = wall;
        stats->systems_per_s = wall > 0 ? (double)nsys / wall : 0.0;
        if (nl > 0) {
            stats->p50_ms = 1e3 * lat[(int64_t)(0.50 * (double)(nl - 1))];
            stats->p99_ms = 1e3 * lat[(int64_t)(0.99 * (double)(nl - 1))];
            stats->max_ms = 1e3 * lat[nl - 1];
        }
        stats->mean_batch = launches > 0 ? (double)launched_reqs / (double)launches : 0.0;
        stats->peak_bytes = A.peak;
        stats->resident_keys = K.n;
        int64_t rb = 0;
        for (int64_t i = 0; i < K.cap; i++)
            if (K.e[i].used) rb += K.e[i].bytes;
        stats->resident_bytes = rb;
    }

out:
    if (S) {
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &S[s];
            if (x->stream) hipStreamSynchronize(x->stream);
        }
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &S[s];
            if (x->done) hipEventDestroy(x->done);
            if (x->stream) hipStreamDestroy(x->stream);
            if (x->h_items) hipHostFree(x->h_items);
            if (x->h_units) hipHostFree(x->h_units);
            if (x->d_items) hipFree(x->d_items);
            if (x->d_units) hipFree(x->d_units);
            if (x->d_counters) hipFree(x->d_counters);
            free(x->reqs);
        }
    }
    if (K.e) {
        for (int64_t i = 0; i < K.cap; i++)
            if (K.e[i].used && K.e[i].ready) hipEventDestroy(K.e[i].ready);
        free(K.e);
    }
    arena_free_all(&A);
    free(S);
    free(st);
    free(pending);
    free(lat);
    return res;
}
