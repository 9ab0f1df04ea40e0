/*
 * jw_serve.c -- a serving loop for fused JW steps (pfsolve_amd.h, pfsolve_serve).
 *
 * Requests (bundles of real systems of one size N, with host input and output)
 * arrive over time.  The loop
 *   - admits arrived requests in arrival order while the device budget (one
 *     arena allocated at start: right-hand sides of requests in flight plus
 *     resident matrices) and the in-flight request limit allow;
 *   - keeps matrices resident per key in a growable table.  A key is pinned by
 *     every admitted, unfinished request that uses it; when the arena is full,
 *     unpinned keys are evicted least recently used first (a key with a request
 *     in flight is never evicted);
 *   - merges every admitted, not yet launched request into one launch on the
 *     next free stream slot: the matrices of keys new to the device (copied and
 *     turned into their coefficient table on that stream; later launches on other
 *     streams wait for that event), host->device copies of the bundles, one
 *     persistent queue kernel per row class over all bundles of the launch (the
 *     kernels of the uneven-size batch), device->host copies, an event;
 *   - finds finished launches by their events (blocking on the oldest one when
 *     nothing else can be done before the next arrival); a finished launch returns
 *     its memory, pins and slot, so requests that arrived meanwhile are admitted
 *     and launched while other launches are still on the device;
 *   - survives a failed launch: its stream is drained, its requests end with the
 *     failure's status, their memory and pins are returned, keys whose table that
 *     launch was to write are written again by a later launch, and the loop goes on.
 * The device work goes through a table of operations: the GPU, or a dry run on the
 * host (no kernels: results are the inputs, launches complete after a set number of
 * polls, and every completion checks that the matrices of its keys were not
 * overwritten while it was in flight), which exercises the admission, eviction and
 * recovery policy without a device.
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
/* an arrival closer than this is waited for by polling, not by blocking on a launch */
static const double kNearArrivalS = 200e-6;

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

/* ---- device operations: the GPU, or the host dry run ---- */
typedef struct serve_dev {
    int dry;                  /* 1: host dry run */
    int latency;              /* dry run: polls until a launch completes */
    int64_t polls;            /* dry run: event queries so far */
} serve_dev;

typedef struct dry_event { int64_t ready_at; int recorded; } dry_event;

static int dev_arena(const serve_dev* d, int64_t size, char** base)
{
    if (d->dry) {
        *base = (char*)malloc((size_t)size);
        return *base ? 0 : -1;
    }
    return hipMalloc((void**)base, (size_t)size) == hipSuccess ? 0 : -1;
}
static void dev_arena_free(const serve_dev* d, char* base)
{
    if (!base) return;
    if (d->dry) free(base);
    else (void)hipFree(base);
}
static int dev_stream(const serve_dev* d, hipStream_t* s)
{
    if (d->dry) { *s = (hipStream_t)(uintptr_t)1; return 0; }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking) == hipSuccess ? 0 : -1;
}
static void dev_stream_free(const serve_dev* d, hipStream_t s)
{
    if (!d->dry && s) (void)hipStreamDestroy(s);
}
static int dev_stream_sync(const serve_dev* d, hipStream_t s)
{
    return d->dry ? 0 : (hipStreamSynchronize(s) == hipSuccess ? 0 : -1);
}
static int dev_event(const serve_dev* d, hipEvent_t* e)
{
    if (d->dry) {
        dry_event* x = (dry_event*)calloc(1, sizeof(dry_event));
        *e = (hipEvent_t)x;
        return x ? 0 : -1;
    }
    /* blocking sync: a host thread waiting on it sleeps instead of spinning */
    return hipEventCreateWithFlags(e, hipEventDisableTiming | hipEventBlockingSync) == hipSuccess ? 0 : -1;
}
static void dev_event_free(const serve_dev* d, hipEvent_t e)
{
    if (!e) return;
    if (d->dry) free((dry_event*)e);
    else (void)hipEventDestroy(e);
}
static int dev_record(serve_dev* d, hipEvent_t e, hipStream_t s)
{
    if (d->dry) {
        dry_event* x = (dry_event*)e;
        x->ready_at = d->polls + d->latency;
        x->recorded = 1;
        return 0;
    }
    return hipEventRecord(e, s) == hipSuccess ? 0 : -1;
}
/* 0 done, 1 not ready, -1 error */
static int dev_query(serve_dev* d, hipEvent_t e)
{
    if (d->dry) {
        const dry_event* x = (const dry_event*)e;
        d->polls++;
        return (!x->recorded || d->polls >= x->ready_at) ? 0 : 1;
    }
    const hipError_t r = hipEventQuery(e);
    return r == hipSuccess ? 0 : (r == hipErrorNotReady ? 1 : -1);
}
static int dev_event_sync(serve_dev* d, hipEvent_t e)
{
    if (d->dry) {
        dry_event* x = (dry_event*)e;
        if (d->polls < x->ready_at) d->polls = x->ready_at;
        return 0;
    }
    return hipEventSynchronize(e) == hipSuccess ? 0 : -1;
}
static int dev_wait(const serve_dev* d, hipStream_t s, hipEvent_t e)
{
    return d->dry ? 0 : (hipStreamWaitEvent(s, e, 0) == hipSuccess ? 0 : -1);
}
static int dev_copy(const serve_dev* d, void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (d->dry) { memcpy(dst, src, bytes); return 0; }
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s) == hipSuccess ? 0 : -1;
}
static int dev_copy2d(const serve_dev* d, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                      size_t rows, hipStream_t s)
{
    if (d->dry) {
        for (size_t r = 0; r < rows; r++) memcpy((char*)dst + r * dpitch, (const char*)src + r * spitch, width);
        return 0;
    }
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDefault, s) == hipSuccess ? 0 : -1;
}
/* rows of `width` bytes; one contiguous copy when the pitches equal the width */
static int dev_copy_rows(const serve_dev* d, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                         size_t rows, hipStream_t s)
{
    if (dpitch == width && spitch == width) return dev_copy(d, dst, src, width * rows, s);
    return dev_copy2d(d, dst, dpitch, src, spitch, width, rows, s);
}
static int dev_host_alloc(const serve_dev* d, void** p, size_t bytes)
{
    if (d->dry) { *p = malloc(bytes); return *p ? 0 : -1; }
    return hipHostMalloc(p, bytes, 0) == hipSuccess ? 0 : -1;
}
static void dev_host_free(const serve_dev* d, void* p)
{
    if (!p) return;
    if (d->dry) free(p);
    else (void)hipHostFree(p);
}
static int dev_alloc(const serve_dev* d, void** p, size_t bytes)
{
    if (d->dry) { *p = malloc(bytes); return *p ? 0 : -1; }
    return hipMalloc(p, bytes) == hipSuccess ? 0 : -1;
}
static void dev_free(const serve_dev* d, void* p)
{
    if (!p) return;
    if (d->dry) free(p);
    else (void)hipFree(p);
}

/* ---- arena: first-fit over a sorted free list, coalescing on free ---- */
typedef struct blk { int64_t off, len; } blk;
typedef struct arena {
    char* base;
    int64_t size, used, peak;
    blk* fr;              /* free blocks sorted by offset */
    int nfree, cap;
} arena;

static int arena_init(arena* a, const serve_dev* d, int64_t size)
{
    memset(a, 0, sizeof(*a));
    if (dev_arena(d, size, &a->base) != 0) return -1;
    a->size = size;
    a->cap = 64;
    a->fr = (blk*)malloc(sizeof(blk) * (size_t)a->cap);
    if (!a->fr) return -1;
    a->fr[0].off = 0;
    a->fr[0].len = size;
    a->nfree = 1;
    return 0;
}

static void arena_free_all(arena* a, const serve_dev* d)
{
    dev_arena_free(d, a->base);
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

/* ---- resident matrices per key: entries in a pool (indices stay valid when it grows),
 * an open-addressing index with tombstones that doubles when half full, and an LRU list
 * (most recently used at the tail) ---- */
typedef struct resident {
    uint64_t key;
    int64_t N;
    int R;
    int64_t off, bytes;       /* arena block: matrices (padded) then the table */
    int64_t mat_bytes;        /* bytes of the matrices copied */
    const double* host_mat;   /* the caller's matrices (copied at the key's first launch) */
    jw_source_desc src;       /* where the table build finds them (offsets in the block) */
    void* table;
    hipEvent_t ready;         /* table written (recorded on the stream of the launch that wrote it) */
    int written;              /* the table build is enqueued (ready recorded) */
    int pins;                 /* admitted, unfinished requests using the key */
    int prev, next;           /* LRU links (pool indices, -1 none) */
    int live;
} resident;

typedef struct keycache {
    resident* pool;
    int npool, cap_pool, free_head;
    int* index;               /* -1 empty, -2 deleted, else pool index */
    int64_t nindex, n, deleted;
    int lru_head, lru_tail;
    int64_t evictions;
} keycache;

static uint64_t key_hash(uint64_t k) { return k * 0x9E3779B97F4A7C15ull; }

static int kc_init(keycache* c, int64_t nindex)
{
    memset(c, 0, sizeof(*c));
    c->nindex = nindex;
    c->index = (int*)malloc(sizeof(int) * (size_t)nindex);
    c->cap_pool = 64;
    c->pool = (resident*)calloc((size_t)c->cap_pool, sizeof(resident));
    if (!c->index || !c->pool) return -1;
    for (int64_t i = 0; i < nindex; i++) c->index[i] = -1;
    c->free_head = -1;
    c->lru_head = c->lru_tail = -1;
    return 0;
}

static void kc_free(keycache* c)
{
    free(c->index);
    free(c->pool);
    memset(c, 0, sizeof(*c));
}

static int kc_find(const keycache* c, uint64_t key)
{
    const uint64_t h = key_hash(key);
    for (int64_t i = 0; i < c->nindex; i++) {
        const int v = c->index[(h + (uint64_t)i) % (uint64_t)c->nindex];
        if (v == -1) return -1;
        if (v >= 0 && c->pool[v].key == key) return v;
    }
    return -1;
}

static void kc_place(keycache* c, uint64_t key, int v)
{
    const uint64_t h = key_hash(key);
    for (int64_t i = 0;; i++) {
        int* s = &c->index[(h + (uint64_t)i) % (uint64_t)c->nindex];
        if (*s < 0) {
            if (*s == -2) c->deleted--;
            *s = v;
            return;
        }
    }
}

static void lru_unlink(keycache* c, int v)
{
    resident* r = &c->pool[v];
    if (r->prev >= 0) c->pool[r->prev].next = r->next; else c->lru_head = r->next;
    if (r->next >= 0) c->pool[r->next].prev = r->prev; else c->lru_tail = r->prev;
    r->prev = r->next = -1;
}

static void lru_append(keycache* c, int v)
{
    resident* r = &c->pool[v];
    r->prev = c->lru_tail;
    r->next = -1;
    if (c->lru_tail >= 0) c->pool[c->lru_tail].next = v; else c->lru_head = v;
    c->lru_tail = v;
}

static void kc_touch(keycache* c, int v)
{
    lru_unlink(c, v);
    lru_append(c, v);
}

/* a new entry for `key` (not present); -1 when out of host memory */
static int kc_insert(keycache* c, uint64_t key)
{
    if ((c->n + c->deleted + 1) * 2 > c->nindex) {
        /* double the index (dropping tombstones) */
        const int64_t nn = c->nindex * 2;
        int* ni = (int*)malloc(sizeof(int) * (size_t)nn);
        if (!ni) return -1;
        free(c->index);
        c->index = ni;
        c->nindex = nn;
        c->deleted = 0;
        for (int64_t i = 0; i < nn; i++) ni[i] = -1;
        for (int v = 0; v < c->npool; v++)
            if (c->pool[v].live) kc_place(c, c->pool[v].key, v);
    }
    int v = c->free_head;
    if (v >= 0) {
        c->free_head = c->pool[v].next;
    } else {
        if (c->npool == c->cap_pool) {
            resident* np = (resident*)realloc(c->pool, sizeof(resident) * (size_t)c->cap_pool * 2);
            if (!np) return -1;
            memset(np + c->cap_pool, 0, sizeof(resident) * (size_t)c->cap_pool);
            c->pool = np;
            c->cap_pool *= 2;
        }
        v = c->npool++;
    }
    memset(&c->pool[v], 0, sizeof(resident));
    c->pool[v].key = key;
    c->pool[v].live = 1;
    c->pool[v].prev = c->pool[v].next = -1;
    kc_place(c, key, v);
    lru_append(c, v);
    c->n++;
    return v;
}

static void kc_remove(keycache* c, int v)
{
    const uint64_t h = key_hash(c->pool[v].key);
    for (int64_t i = 0; i < c->nindex; i++) {
        int* s = &c->index[(h + (uint64_t)i) % (uint64_t)c->nindex];
        if (*s == v) {
            *s = -2;
            c->deleted++;
            break;
        }
    }
    lru_unlink(c, v);
    c->pool[v].live = 0;
    c->pool[v].next = c->free_head;
    c->free_head = v;
    c->n--;
}

/* least recently used key with no request in flight, -1 if none */
static int kc_victim(const keycache* c)
{
    for (int v = c->lru_head; v >= 0; v = c->pool[v].next)
        if (c->pool[v].pins == 0) return v;
    return -1;
}

/* ---- launch slots ---- */
typedef struct slot {
    hipStream_t stream;
    hipEvent_t done;
    int busy;
    int64_t seq;              /* launch number (the oldest busy slot is waited for first) */
    int64_t* reqs;            /* request indices of the launch */
    int nreqs;
    int* wrote;               /* keys whose table this launch writes */
    int nwrote;
    jw_batch_item* h_items;   /* pinned */
    jw_batch_unit* h_units;   /* pinned */
    jw_batch_item* d_items;
    jw_batch_unit* d_units;
    int* d_counters;
} slot;

typedef struct reqstate {
    int64_t off, bytes;       /* device block of the bundle, -1 when not admitted */
    int64_t stride;
    int key;                  /* pool index of the request's key while pinned, else -1 */
    int state;                /* 0 waiting, 1 admitted, 2 launched, 3 done, 4 rejected, 5 failed */
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

/* where the table build of a request's key reads its matrices (the key's arena block) */
static void fused_source(const pfsolve_serve_request* q, jw_source_desc* src)
{
    memset(src, 0, sizeof(*src));
    src->N = q->N;
    src->band_kind = JW_BAND_GB;
    src->offV = q->offsetV;
    src->KL = q->KL;
    src->KU = q->KU;
    src->solve_kind = JW_SOLVE_TRI;
    src->ldS = q->offsetT;
    src->mdS = q->offsetT + q->N;
    src->udS = q->offsetT + 2 * q->N;
    src->has_m = 1;
    src->mdM = q->offsetM;
    src->udM = q->offsetM + q->N;
    src->Msplit0 = 704;
    src->permuted = !q->naturalOrder && q->N > 64;
}

typedef struct serve_run {
    serve_dev dev;
    arena A;
    keycache K;
    int nkeys_peak;
} serve_run;

/* frees the key's block (its launches are all finished: it has no pins) */
static void evict(serve_run* S, int v)
{
    resident* r = &S->K.pool[v];
    dev_event_free(&S->dev, r->ready);
    r->ready = NULL;
    arena_release(&S->A, r->off, r->bytes);
    kc_remove(&S->K, v);
    S->K.evictions++;
}

/* arena block of `len` bytes, evicting unpinned keys least recently used first; -1 if
 * every resident key is pinned and the arena still has no room */
static int64_t alloc_evicting(serve_run* S, int64_t len)
{
    for (;;) {
        const int64_t off = arena_alloc(&S->A, len);
        if (off >= 0) return off;
        const int v = kc_victim(&S->K);
        if (v < 0) return -1;
        evict(S, v);
    }
}

PfSolveResult pfsolve_serve(const pfsolve_serve_config* cfg, pfsolve_serve_request* reqs, int64_t nreq,
                            pfsolve_serve_stats* stats)
{
    if (!cfg || !stats || (nreq > 0 && !reqs)) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    memset(stats, 0, sizeof(*stats));
    if (cfg->hbm_budget < (1 << 20) || cfg->max_inflight < 1 || cfg->streams < 1 || cfg->streams > 16 ||
        cfg->max_batch < 1 || cfg->fail_launch < 0 || cfg->dry_run < 0)
        return PFSOLVE_ERROR_EMPTY_size;
    for (int64_t i = 0; i < nreq; i++) {
        PfSolveResult r = check_request(&reqs[i]);
        if (r != PFSOLVE_SUCCESS) return r;
        if (i > 0 && reqs[i].arrival_s < reqs[i - 1].arrival_s) return PFSOLVE_ERROR_EMPTY_size;
    }
    if (nreq == 0) return PFSOLVE_SUCCESS;

    serve_run S;
    memset(&S, 0, sizeof(S));
    S.dev.dry = cfg->dry_run > 0;
    S.dev.latency = cfg->dry_run;
    serve_dev* dev = &S.dev;
    int cu = 0;
    if (!dev->dry) {
        int lds = 0, maj = 0, min = 0;
        if (jwdev_device_info(&cu, &lds, &maj, &min) != 0) return PFSOLVE_ERROR_INVALID_DEVICE;
    }

    PfSolveResult res = PFSOLVE_SUCCESS;
    slot* SL = (slot*)calloc((size_t)cfg->streams, sizeof(slot));
    reqstate* st = (reqstate*)calloc((size_t)nreq, sizeof(reqstate));
    int64_t* pending = (int64_t*)malloc(sizeof(int64_t) * (size_t)nreq);
    double* lat = (double*)malloc(sizeof(double) * (size_t)nreq);
    const int maxb = cfg->max_batch;
    /* units per launch: a request's passes are cut into at most kUnitsPerReq units */
    const int64_t max_units = (int64_t)maxb * kUnitsPerReq;
    if (!SL || !st || !pending || !lat) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    if (arena_init(&S.A, dev, cfg->hbm_budget) != 0) { res = PFSOLVE_ERROR_FAILED_TO_ALLOCATE; goto out; }
    if (kc_init(&S.K, 1024) != 0) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    for (int s = 0; s < cfg->streams; s++) {
        slot* x = &SL[s];
        if (dev_stream(dev, &x->stream) != 0 || dev_event(dev, &x->done) != 0 ||
            dev_host_alloc(dev, (void**)&x->h_items, sizeof(jw_batch_item) * (size_t)maxb) != 0 ||
            dev_host_alloc(dev, (void**)&x->h_units, sizeof(jw_batch_unit) * (size_t)max_units) != 0 ||
            dev_alloc(dev, (void**)&x->d_items, sizeof(jw_batch_item) * (size_t)maxb) != 0 ||
            dev_alloc(dev, (void**)&x->d_units, sizeof(jw_batch_unit) * (size_t)max_units) != 0 ||
            dev_alloc(dev, (void**)&x->d_counters, sizeof(int) * 2 * NCLS) != 0) {
            res = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
            goto out;
        }
        /* the slot's queues start empty (each queue kernel leaves them so) */
        if (!dev->dry && jwdev_zero_counters(x->d_counters, 2 * NCLS, x->stream) != 0) {
            res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            goto out;
        }
        x->reqs = (int64_t*)malloc(sizeof(int64_t) * (size_t)maxb);
        x->wrote = (int*)malloc(sizeof(int) * (size_t)maxb);
        if (!x->reqs || !x->wrote) { res = PFSOLVE_ERROR_MALLOC_FAILED; goto out; }
    }
    for (int64_t i = 0; i < nreq; i++) {
        st[i].off = -1;
        st[i].key = -1;
    }

    /* launch shapes per class: the batch kernel's waves and workgroups per CU */
    int waves[NCLS], grid_max[NCLS];
    for (int c = 0; c < NCLS; c++) {
        waves[c] = 1;
        grid_max[c] = 1;
        if (dev->dry) continue;
        const int R = 2 << c;
        waves[c] = jwdev_main64_waves(R);
        const int occ = waves[c] > 0 ? jwdev_batch64_occupancy(R, waves[c]) : 0;
        if (waves[c] < 1 || occ < 1) { res = PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY; goto out; }
        grid_max[c] = cu * occ;
    }

    const double t0 = now_s();
    int64_t next = 0, npend = 0, ndone = 0, inflight = 0, launches = 0, launched_reqs = 0, failed = 0;
    int64_t integrity = 0;
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
            int v = kc_find(&S.K, q->matrix_key);
            const int R = 2 << row_class(q->N);
            const int64_t mbytes = (matrix_extent(q) * 8 + ALIGN - 1) / ALIGN * ALIGN;
            const int64_t kbytes = mbytes + jwdev_table64_elems(R) * 8;
            int status = PFSOLVE_SUCCESS;
            jw_source_desc qs;
            fused_source(q, &qs);
            if (v >= 0 && S.K.pool[v].N != q->N) {
                status = PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;   /* one key, one size */
            } else if (v >= 0 && (S.K.pool[v].src.KL != qs.KL || S.K.pool[v].src.KU != qs.KU ||
                                  S.K.pool[v].src.offV != qs.offV || S.K.pool[v].src.ldS != qs.ldS ||
                                  S.K.pool[v].src.mdM != qs.mdM || S.K.pool[v].src.permuted != qs.permuted)) {
                /* the key's table was built for another band, layout or storage order */
                status = PFSOLVE_ERROR_UNSUPPORTED_RADIX;
            } else if (r->bytes + (v >= 0 ? 0 : kbytes) > S.A.size) {
                status = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;      /* can never fit: rejected, not waited for */
            }
            if (status != PFSOLVE_SUCCESS) {
                r->state = 4;
                q->status = status;
                q->admit_s = q->done_s = t;
                ndone++;
                next++;
                progress = 1;
                continue;
            }
            int64_t koff = -1, roff = -1;
            if (v < 0) {
                koff = alloc_evicting(&S, kbytes);
                if (koff < 0) {
                    if (inflight == 0) goto reject_nomem;   /* only possible if it can never fit */
                    break;                                  /* wait for memory to free up */
                }
            } else {
                /* pinned below, so that the request's own allocation cannot evict it */
                S.K.pool[v].pins++;
            }
            if (r->bytes > 0) {
                roff = alloc_evicting(&S, r->bytes);
                if (roff < 0) {
                    if (koff >= 0) arena_release(&S.A, koff, kbytes);
                    if (v >= 0) S.K.pool[v].pins--;
                    if (inflight == 0) goto reject_nomem;
                    break;
                }
            }
            if (v < 0) {
                v = kc_insert(&S.K, q->matrix_key);
                if (v < 0) {
                    arena_release(&S.A, koff, kbytes);
                    if (roff >= 0) arena_release(&S.A, roff, r->bytes);
                    res = PFSOLVE_ERROR_MALLOC_FAILED;
                    goto out;
                }
                resident* k = &S.K.pool[v];
                k->N = q->N;
                k->R = R;
                k->off = koff;
                k->bytes = kbytes;
                k->mat_bytes = matrix_extent(q) * 8;
                k->host_mat = q->matrices;
                fused_source(q, &k->src);
                k->table = S.A.base + koff + mbytes;
                k->pins = 1;
                if ((int64_t)S.K.n > S.nkeys_peak) S.nkeys_peak = (int)S.K.n;
            }
            kc_touch(&S.K, v);
            r->off = roff;
            r->key = v;
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
                if (!SL[i].busy) { s = i; break; }
            if (s >= 0) {
                slot* x = &SL[s];
                const int nb = npend < maxb ? (int)npend : maxb;
                x->nreqs = nb;
                x->nwrote = 0;
                for (int i = 0; i < nb; i++) x->reqs[i] = pending[i];
                memmove(pending, pending + nb, sizeof(int64_t) * (size_t)(npend - nb));
                npend -= nb;
                launches++;
                launched_reqs += nb;
                for (int i = 0; i < nb; i++) st[x->reqs[i]].state = 2;
                int err = 0;
                int64_t nunits = 0;
                int64_t cls_first[NCLS], cls_n[NCLS];
                for (int c = 0; c < NCLS; c++) cls_n[c] = 0;
                /* keys new to the device: matrices copied and the table written on this
                 * stream; keys written by an earlier launch: wait for that event */
                for (int i = 0; i < nb && !err; i++) {
                    resident* k = &S.K.pool[st[x->reqs[i]].key];
                    if (!k->written) {
                        if (dev_copy(dev, S.A.base + k->off, k->host_mat, (size_t)k->mat_bytes, x->stream) != 0 ||
                            (!k->ready && dev_event(dev, &k->ready) != 0)) {
                            err = PFSOLVE_ERROR_FAILED_TO_COPY;
                            break;
                        }
                        if (!dev->dry &&
                            jwdev_prologue(&k->src, S.A.base + k->off, k->table, 64, k->R, JW_PREC_F64, 2, x->stream) != 0) {
                            err = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                            break;
                        }
                        if (dev_record(dev, k->ready, x->stream) != 0) { err = PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD; break; }
                        k->written = 1;
                        x->wrote[x->nwrote++] = st[x->reqs[i]].key;
                    } else if (dev_wait(dev, x->stream, k->ready) != 0) {
                        err = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                    }
                }
                /* host -> device copies of the bundles (systems padded to an even stride) */
                for (int i = 0; i < nb && !err; i++) {
                    const int64_t qi = x->reqs[i];
                    pfsolve_serve_request* q = &reqs[qi];
                    reqstate* r = &st[qi];
                    if (q->nsys > 0 && dev_copy_rows(dev, S.A.base + r->off, (size_t)(r->stride * 8), q->rhs_in,
                                                     (size_t)(q->N * 8), (size_t)(q->N * 8), (size_t)q->nsys,
                                                     x->stream) != 0) {
                        err = PFSOLVE_ERROR_FAILED_TO_COPY;
                        break;
                    }
                    const resident* k = &S.K.pool[r->key];
                    jw_batch_item* it = &x->h_items[i];
                    memset(it, 0, sizeof(*it));
                    it->rhs = S.A.base + r->off;
                    it->table = k->table;
                    it->nsys = q->nsys;
                    it->stride = r->stride;
                    it->offSol = 0;
                    it->N = (int32_t)q->N;
                    it->kl = q->KL;
                    it->ku = q->KU;
                    it->apply_scale = q->scaleC != 1.0;
                    it->scale = q->scaleC;
                }
                /* work units per row class, largest N first within a class */
                for (int c = NCLS - 1; c >= 0 && !err; c--) {
                    cls_first[c] = nunits;
                    for (int i = 0; i < nb; i++) {
                        const pfsolve_serve_request* q = &reqs[x->reqs[i]];
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
                /* a launch failure injected from the host: the kernels of this launch */
                if (!err && cfg->fail_launch == launches) err = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                if (!err && !dev->dry) {
                    /* the queues are empty: zeroed at set-up, left empty by every queue kernel */
                    if (dev_copy(dev, x->d_items, x->h_items, sizeof(jw_batch_item) * (size_t)nb, x->stream) != 0 ||
                        (nunits > 0 &&
                         dev_copy(dev, x->d_units, x->h_units, sizeof(jw_batch_unit) * (size_t)nunits, x->stream) != 0))
                        err = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                    for (int c = 0; c < NCLS && !err; c++) {
                        if (!cls_n[c]) continue;
                        int64_t grid = cls_n[c] < grid_max[c] ? cls_n[c] : grid_max[c];
                        /* R = 2: sub-wavefront groups for requests with N <= 64 (8 systems
                         * per wavefront pass for N <= 32, 4 for N <= 64) */
                        if (jwdev_batch64(x->d_items, x->d_units + cls_first[c], (int)cls_n[c], x->d_counters + 2 * c,
                                          2 << c, waves[c], (int)grid, 8, x->stream) != 0)
                            err = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                    }
                }
                for (int i = 0; i < nb && !err; i++) {
                    const int64_t qi = x->reqs[i];
                    pfsolve_serve_request* q = &reqs[qi];
                    reqstate* r = &st[qi];
                    if (q->nsys > 0 && dev_copy_rows(dev, q->rhs_out, (size_t)(q->N * 8), S.A.base + r->off,
                                                     (size_t)(r->stride * 8), (size_t)(q->N * 8), (size_t)q->nsys,
                                                     x->stream) != 0)
                        err = PFSOLVE_ERROR_FAILED_TO_COPY;
                }
                if (!err && dev_record(dev, x->done, x->stream) != 0) err = PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
                if (err) {
                    /* drain what was enqueued, then end the launch's requests with the
                     * failure; its keys' tables are written again by a later launch */
                    if (dev_stream_sync(dev, x->stream) != 0) { res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE; goto out; }
                    for (int w = 0; w < x->nwrote; w++) S.K.pool[x->wrote[w]].written = 0;
                    const double td = now_s() - t0;
                    for (int i = 0; i < nb; i++) {
                        const int64_t qi = x->reqs[i];
                        reqstate* r = &st[qi];
                        if (r->bytes > 0) arena_release(&S.A, r->off, r->bytes);
                        r->off = -1;
                        S.K.pool[r->key].pins--;
                        r->key = -1;
                        r->state = 5;
                        reqs[qi].status = err;
                        reqs[qi].done_s = td;
                        inflight--;
                        ndone++;
                        failed++;
                    }
                } else {
                    x->busy = 1;
                    x->seq = launches;
                }
                progress = 1;
            }
        }

        /* 3. completions */
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &SL[s];
            if (!x->busy) continue;
            const int e = dev_query(dev, x->done);
            if (e == 1) continue;
            if (e < 0) { res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE; goto out; }
            const double td = now_s() - t0;
            for (int i = 0; i < x->nreqs; i++) {
                const int64_t qi = x->reqs[i];
                reqstate* r = &st[qi];
                resident* k = &S.K.pool[r->key];
                int status = PFSOLVE_SUCCESS;
                /* dry run: the key's matrices must be what was copied for it */
                if (dev->dry && memcmp(S.A.base + k->off, k->host_mat, (size_t)k->mat_bytes) != 0) {
                    status = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
                    integrity++;
                }
                if (r->bytes > 0) arena_release(&S.A, r->off, r->bytes);
                r->off = -1;
                k->pins--;
                r->key = -1;
                r->state = 3;
                reqs[qi].done_s = td;
                reqs[qi].status = status;
                inflight--;
                ndone++;
            }
            x->busy = 0;
            progress = 1;
        }
        if (progress) {
            last_progress = now_s();
            continue;
        }
        if (now_s() - last_progress > kStallLimitS) {
            /* a launch that never completes: give up instead of spinning for ever */
            res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
            goto out;
        }
        /* nothing to do until a launch finishes or the next request arrives */
        const double until = (cfg->realtime && next < nreq) ? reqs[next].arrival_s - (now_s() - t0) : 1e30;
        int oldest = -1;
        for (int s = 0; s < cfg->streams; s++)
            if (SL[s].busy && (oldest < 0 || SL[s].seq < SL[oldest].seq)) oldest = s;
        if (oldest >= 0 && until > kNearArrivalS) {
            /* block on the oldest launch (the host thread sleeps in the driver) */
            if (dev_event_sync(dev, SL[oldest].done) != 0) { res = PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE; goto out; }
        } else if (oldest < 0 && until < 1e29) {
            if (until > kNearArrivalS) usleep((useconds_t)((until - 100e-6) * 1e6));
        } else if (!dev->dry) {
            usleep(10);
        }
    }

    {
        const double wall = now_s() - t0;
        int64_t nl = 0, nsys = 0, rejected = 0;
        for (int64_t i = 0; i < nreq; i++) {
            if (st[i].state == 4) rejected++;
            if (st[i].state != 3 || reqs[i].status != PFSOLVE_SUCCESS) continue;
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
        stats->peak_bytes = S.A.peak;
        stats->resident_keys = S.K.n;
        int64_t rb = 0;
        for (int v = 0; v < S.K.npool; v++)
            if (S.K.pool[v].live) rb += S.K.pool[v].bytes;
        stats->resident_bytes = rb;
        stats->evictions = S.K.evictions;
        stats->failed = failed;
        stats->integrity_failures = integrity;
        stats->peak_keys = S.nkeys_peak;
    }

out:
    if (SL) {
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &SL[s];
            if (x->stream) dev_stream_sync(dev, x->stream);
        }
        for (int s = 0; s < cfg->streams; s++) {
            slot* x = &SL[s];
            dev_event_free(dev, x->done);
            dev_stream_free(dev, x->stream);
            dev_host_free(dev, x->h_items);
            dev_host_free(dev, x->h_units);
            dev_free(dev, x->d_items);
            dev_free(dev, x->d_units);
            dev_free(dev, x->d_counters);
            free(x->reqs);
            free(x->wrote);
        }
    }
    for (int v = 0; v < S.K.npool; v++)
        if (S.K.pool && S.K.pool[v].live) dev_event_free(dev, S.K.pool[v].ready);
    kc_free(&S.K);
    arena_free_all(&S.A, dev);
    free(SL);
    free(st);
    free(pending);
    free(lat);
    return res;
}
