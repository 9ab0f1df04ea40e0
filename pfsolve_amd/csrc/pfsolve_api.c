/*
 * pfsolve_api.c -- the drop-in C entry points (pfSolve.h) and the extension
 * introspection functions (pfsolve_amd.h).  Host code in C; device work is
 * enqueued through the launchers of jw_device.hip.
 *
 *   initializePfSolve  PfSolve/pfSolve_AppManagement/pfSolve_InitializeApp.h:1480-1644
 *   PfSolveAppend      PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:80-159
 *   deletePfSolve      PfSolve/pfSolve_AppManagement/pfSolve_DeleteApp.h:28-312
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jw_plan.h"

int jwdev_main_occupancy(int L, int R, int precision, int waves);

/* Stream slot of the next launch: with num_streams > 1 launches go round robin over
 * stream[0..num_streams) (pfSolve_DispatchPlan.h:337-362), each slot with its own event
 * and its own coefficient table, so that launches in flight on different streams never
 * share device scratch. */
static uint64_t plan_slot(const PfSolveApplication* app)
{
    const PfSolveConfiguration* c = &app->configuration;
    return (c->num_streams > 1 && c->stream) ? c->streamID % c->num_streams : 0;
}
static hipStream_t slot_stream(const PfSolveApplication* app, uint64_t slot)
{
    return app->configuration.stream ? app->configuration.stream[slot] : (hipStream_t)0;
}
static void* slot_table(const struct pfsolve_plan* p, uint64_t slot)
{
    return p->table ? (char*)p->table + slot * p->table_stride : NULL;
}

static PfSolveResult fill_geometry(const PfSolveConfiguration* cfg, pfsolve_geometry* g)
{
    memset(g, 0, sizeof(*g));
    g->precision = jw_precision(cfg);
    g->waves = 4;
    if (jw_plan_kind(cfg) != PF_KIND_JW) return PFSOLVE_SUCCESS;
    int L = 0, R = 0;
    /* the ALT variant solves two subsystems of ceil(N/2) and floor(N/2) rows */
    const int alt = (cfg->jw_type % 1000) / 100 == 1;
    if (!alt && cfg->numConsecutiveJWIterations <= 1 && jw_is_chunked(cfg->M_size, g->precision)) {
        /* long systems: chunks of 64 R rows on the chunk kernel, one table per chunk */
        const int kp = g->precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD;
        const int64_t rows = 64 * (int64_t)jwdev_long_rows(kp);
        const int64_t build = rows * jwdev_long_nsub(kp);
        const int64_t ntab = (cfg->M_size + build - 1) / build * jwdev_long_nsub(kp);
        g->lanes = 64;
        g->rows_per_lane = jwdev_long_rows(kp);
        g->waves = 4;
        g->table_bytes = ntab * jwdev_table_elems(64, g->rows_per_lane) * (kp == JW_PREC_F64 ? 8 : 16);
        g->lds_bytes = jwdev_long_lds_bytes(kp);
        return PFSOLVE_SUCCESS;
    }
    PfSolveResult r = jw_select_geometry(alt ? (cfg->M_size + 1) / 2 : cfg->M_size, g->precision, &L, &R);
    if (r != PFSOLVE_SUCCESS) return r;
    if (R > 16) {
        /* long systems: only the wave-per-system kernel (no general-kernel geometry) */
        g->lanes = 0;
        g->rows_per_lane = 0;
        g->table_bytes = jwdev_table64_elems(R) * 8;
        g->lds_bytes = 0;
        return PFSOLVE_SUCCESS;
    }
    g->lanes = L;
    g->rows_per_lane = R;
    g->table_bytes = jwdev_table_elems(L, R) * (g->precision == JW_PREC_F64 ? 8 : 16);
    g->lds_bytes = jwdev_lds_bytes(L, R, g->precision);
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_validate_configuration(const PfSolveConfiguration* cfg, pfsolve_geometry* g)
{
    if (!cfg) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    PfSolveConfiguration c;
    PfSolveResult r = jw_configure(cfg, &c);
    if (r != PFSOLVE_SUCCESS) return r;
    pfsolve_geometry tmp;
    r = fill_geometry(&c, g ? g : &tmp);
    if (r != PFSOLVE_SUCCESS) return r;
    if (jw_plan_kind(&c) == PF_KIND_JW) {
        const pfsolve_geometry* gg = g ? g : &tmp;
        /* lanes == 0: 1024 < N <= 2048 in FP64, the wave-per-system kernel or chunks */
        if (gg->lanes > 0 &&
            !jwdev_supported(gg->lanes, gg->rows_per_lane, gg->precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD))
            return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    }
    return PFSOLVE_SUCCESS;
}

void deletePfSolve(PfSolveApplication* app)
{
    if (!app) return;
    struct pfsolve_plan* p = app->localFFTPlan;
    if (p) {
        if (p->table) hipFree(p->table);
        if (p->d_err) hipFree(p->d_err);
        free(p);
    }
    app->localFFTPlan = NULL;
    PfSolveConfiguration* c = &app->configuration;
    if (c->stream_event) {
        for (uint64_t i = 0; i < c->num_streams; i++)
            if (c->stream_event[i]) (void)hipEventDestroy(c->stream_event[i]);
        free(c->stream_event);
        c->stream_event = NULL;
    }
}

/* The chunked path of long systems: chunk geometry, the tables of every chunk and the
 * reduced data after whatever the slot's area already holds (p->table_bytes). */
static PfSolveResult long_setup(PfSolveApplication* app, struct pfsolve_plan* p)
{
    const PfSolveConfiguration* cfg = &app->configuration;
    const int kp = p->precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD;
    const size_t elt = kp == JW_PREC_F64 ? 8 : 16;
    const int R = jwdev_long_rows(kp);
    const int64_t rows = 64 * (int64_t)R;
    const int64_t build = rows * jwdev_long_nsub(kp);
    p->long_R = R;
    p->long_nc = (p->N + rows - 1) / rows;
    p->long_ntab = (p->N + build - 1) / build * jwdev_long_nsub(kp);
    p->long_tab_elems = jwdev_table_elems(64, R);
    p->long_occ = jwdev_long_occupancy(kp);
    if (p->long_occ < 1) return PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY;
    /* FP64: 16-byte aligned systems may take the pair path (its tables share the space below) */
    p->long64_occ = (kp == JW_PREC_F64 && R == 8) ? jwdev_long64_occupancy() : 0;
    /* systems per slice: every system of the plan, up to 64 MiB of reduced data */
    int64_t nsys = (int64_t)(cfg->size[1] * cfg->size[2]);
    int64_t cap = ((int64_t)64 << 20) / (JW_LONG_ARRAYS * p->long_nc * (int64_t)elt);
    if (nsys < cap) cap = nsys;
    if (cap < 2) cap = 2;
    cap = (cap + 1) / 2 * 2;
    p->long_red_sys = cap;
    p->long_tab_off = (p->table_bytes + 255) / 256 * 256;
    /* the pair path's format-2 tables (builds of jwdev_long64_nsub chunks) in the same space */
    size_t tab_bytes = (size_t)p->long_ntab * p->long_tab_elems * elt;
    if (kp == JW_PREC_F64) {
        const int64_t b64 = rows * jwdev_long64_nsub();
        const size_t t64 = (size_t)((p->N + b64 - 1) / b64 * jwdev_long64_nsub()) * jwdev_table64_elems(8) * elt;
        if (t64 > tab_bytes) tab_bytes = t64;
    }
    p->long_red_off = p->long_tab_off + (tab_bytes + 255) / 256 * 256;
    const int64_t nbuild = p->long_ntab / jwdev_long_nsub(kp);
    p->long_sync_off = p->long_red_off +
                       ((size_t)(JW_LONG_ARRAYS * p->long_nc * cap + 9 * p->long_nc) * elt + 255) / 256 * 256;
    /* ticket and pivot flags (chunk kernel); counters and table flags (pair path: 32 + nbuild) */
    p->long_gval_off = p->long_sync_off + ((size_t)(nbuild + 32) * sizeof(int) + 255) / 256 * 256;
    p->table_bytes = p->long_gval_off + (size_t)nbuild * elt;
    p->table_stride = (p->table_bytes + 255) / 256 * 256;
    return PFSOLVE_SUCCESS;
}

/* Everything a plan needs that allocates nothing: kind, precision, device limits,
 * geometry and kernel choice (pfSolve_InitializeApp.h:1480-1644 decides these while it
 * generates code).  The caller provides the device table and error flag. */
static PfSolveResult plan_setup(PfSolveApplication* app, struct pfsolve_plan* p)
{
    PfSolveConfiguration* cfg = &app->configuration;
    PfSolveResult res = PFSOLVE_SUCCESS;
    p->kind = jw_plan_kind(cfg);
    p->precision = jw_precision(cfg);
    p->N = cfg->M_size;
    p->chain_K = 1;

    int cu = 0, lds_blk = 0, maj = 0, min = 0;
    if (jwdev_device_info(&cu, &lds_blk, &maj, &min) != 0) return PFSOLVE_ERROR_INVALID_DEVICE;
    p->cu_count = cu;
    cfg->computeCapabilityMajor = (uint64_t)maj;
    cfg->computeCapabilityMinor = (uint64_t)min;
    cfg->sharedMemorySize = (uint64_t)lds_blk;
    cfg->maxThreadsNum = 1024;
    if (p->kind == PF_KIND_JW) {
        snprintf(app->kernelName, sizeof(app->kernelName), "PfSolve_jw_%llu_%llu_%d_%d_%lld_%llu_%d_%d",
                 (unsigned long long)cfg->size[0], (unsigned long long)cfg->size[1], cfg->upperBanded,
                 cfg->jw_type, (long long)cfg->jw_control_bitmask, (unsigned long long)cfg->Msplit[0],
                 cfg->numConsecutiveJWIterations, cfg->useMultipleInputBuffers);
        if (cfg->numConsecutiveJWIterations > 1) {
            /* chained steps: K format-2 tables (one per iteration) per stream slot and the
             * chained kernel, R rows per lane for the largest iteration */
            const int K = cfg->numConsecutiveJWIterations;
            const int64_t nmax = cfg->upperBanded ? p->N : p->N + K - 1;
            int r64 = 2;
            while (r64 * 64 < nmax) r64 *= 2;
            p->chain_K = K;
            p->chain_R = r64;
            p->chain_occ = jwdev_chain64_occupancy(r64);
            if (p->chain_occ < 1) return PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY;
            p->chain_tab = (size_t)jwdev_table64_elems(r64) * 8;
            p->L = 64;
            p->R64 = r64;
            p->table_bytes = (size_t)K * p->chain_tab;
            p->table_stride = (p->table_bytes + 255) / 256 * 256;
            return PFSOLVE_SUCCESS;
        }
        pfsolve_geometry g;
        res = fill_geometry(cfg, &g);
        if (res != PFSOLVE_SUCCESS) return res;
        if ((cfg->jw_type % 1000) / 100 != 1 && jw_is_chunked(p->N, p->precision)) {
            /* longer than one wavefront holds: chunks only */
            p->long_only = 1;
            p->waves = g.waves;
            p->lds_bytes = g.lds_bytes;
            p->table_bytes = 0;
            return long_setup(app, p);
        }
        p->L = g.lanes;
        p->R = g.rows_per_lane;
        p->waves = g.waves;
        p->lds_bytes = g.lds_bytes;
        p->table_bytes = (size_t)g.table_bytes;
        const int kp = p->precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD;
        const int long_n = p->L == 0;   /* 1024 < N <= 2048: the wave-per-system kernel or chunks */
        if (!long_n && !jwdev_supported(p->L, p->R, kp)) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        if ((cfg->jw_type % 1000) / 100 == 1) {
            /* ALT: each parity subsystem on the general kernel (rows two elements apart), one
             * format-1 table per parity */
            if (long_n) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
            p->alt = 1;
            p->alt_tab = (p->table_bytes + 255) / 256 * 256;
            p->table_bytes = 2 * p->alt_tab;
            p->table_stride = (p->table_bytes + 255) / 256 * 256;
            p->occupancy = jwdev_main_occupancy(p->L, p->R, p->precision, p->waves);
            if (p->occupancy < 1) return PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY;
            return PFSOLVE_SUCCESS;
        }
        /* wave-per-system path: FP64, 64 lanes, R a power of two in [2, 32] */
        if (p->precision == JW_PREC_F64 && cfg->kernelVariant != 1) {
            int r64 = 2;
            while (r64 * 64 < p->N) r64 *= 2;
            if (r64 <= 32) {
                p->R64 = r64;
                p->waves64 = jwdev_main64_waves(r64);
                p->occupancy64 = p->waves64 ? jwdev_main64_occupancy(r64, p->waves64) : 0;
                if (p->occupancy64 < 1) p->R64 = 0;
            }
            if (p->R64) {
                size_t b2 = (size_t)jwdev_table64_elems(p->R64) * 8;
                if (b2 > p->table_bytes) p->table_bytes = b2;
            }
        }
        /* wave-per-system double-double path: 64 lanes, R a power of two in [2, 8] */
        if (p->precision != JW_PREC_F64 && cfg->kernelVariant != 1 && !long_n) {
            int r = 2;
            while (r * 64 < p->N) r *= 2;
            if (r <= 8) {
                p->dd64_waves = 8;
                p->dd64_occ = jwdev_dd64_occupancy(r, p->dd64_waves);
                if (p->dd64_occ >= 1) {
                    p->dd64_R = r;
                    size_t b2 = (size_t)jwdev_dd64_table_elems(r) * 16;
                    if (b2 > p->table_bytes) p->table_bytes = b2;
                }
            }
        }
        if (cfg->kernelVariant == 2 && !p->R64 && !p->dd64_R) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        /* compute / store split form of the wave-per-system kernel (R <= 8) */
        if (p->R64 && p->R64 <= 8) {
            int c = 0, st = 0;
            if (jwdev_main64s_shape(p->R64, &c, &st)) {
                p->occupancy_split = jwdev_main64s_occupancy(p->R64);
                if (p->occupancy_split >= 1) p->split64 = c;
            }
        }
        p->table_stride = (p->table_bytes + 255) / 256 * 256;
        if (!long_n) {
            p->occupancy = jwdev_main_occupancy(p->L, p->R, p->precision, p->waves);
            if (p->occupancy < 1) return PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY;
        } else if (cfg->kernelVariant != 2) {
            /* systems the wave-per-system kernel does not take (unaligned, or the general
             * kernel asked for) run in chunks */
            p->main_bytes = p->table_stride;
            res = long_setup(app, p);
            if (res != PFSOLVE_SUCCESS) return res;
        }
    } else if (p->kind == PF_KIND_BLOCK) {
        snprintf(app->kernelName, sizeof(app->kernelName), "PfSolve_block_%llu_%llu_%d_%d_%lld",
                 (unsigned long long)cfg->size[0], (unsigned long long)cfg->size[1], cfg->block, cfg->lshift,
                 (long long)cfg->jw_control_bitmask);
    } else {
        snprintf(app->kernelName, sizeof(app->kernelName), "PfSolve_JW_dgbmv_%llu_%llu_%d_%d_%d_%lld",
                 (unsigned long long)cfg->size[0], (unsigned long long)cfg->size[1], cfg->LDA, cfg->KU, cfg->KL,
                 (long long)cfg->jw_control_bitmask);
    }
    return PFSOLVE_SUCCESS;
}


PfSolveResult initializePfSolve(PfSolveApplication* app, PfSolveConfiguration in)
{
    if (!app) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    /* members after the configuration start zero (the FFT machinery's stay so) */
    memset((char*)app + offsetof(PfSolveApplication, localFFTPlan), 0,
           sizeof(*app) - offsetof(PfSolveApplication, localFFTPlan));
    /* HIP branch of setConfigurationPfSolve: a device pointer is required (pfSolve_InitializeApp.h:656-659) */
    if (in.device == NULL) return PFSOLVE_ERROR_INVALID_DEVICE;
    PfSolveResult res = jw_configure(&in, &app->configuration);
    PfSolveConfiguration* cfg = &app->configuration;
    /* stream bookkeeping belongs to the plan (pfSolve_InitializeApp.h:661-663, 738-748) */
    cfg->stream_event = NULL;
    cfg->streamID = 0;
    cfg->streamCounter = 0;
    if (res != PFSOLVE_SUCCESS) { deletePfSolve(app); return res; }
    if (cfg->num_streams == 0) cfg->num_streams = 1;
    if (cfg->num_streams > 1) {
        if (!cfg->stream) { deletePfSolve(app); return PFSOLVE_ERROR_INVALID_QUEUE; }
        cfg->stream_event = (hipEvent_t*)calloc(cfg->num_streams, sizeof(hipEvent_t));
        if (!cfg->stream_event) { deletePfSolve(app); return PFSOLVE_ERROR_MALLOC_FAILED; }
        for (uint64_t i = 0; i < cfg->num_streams; i++)
            if (hipEventCreateWithFlags(&cfg->stream_event[i], hipEventDisableTiming) != hipSuccess) {
                deletePfSolve(app);
                return PFSOLVE_ERROR_FAILED_TO_CREATE_EVENT;
            }
    }

    struct pfsolve_plan* p = (struct pfsolve_plan*)calloc(1, sizeof(*p));
    if (!p) { deletePfSolve(app); return PFSOLVE_ERROR_MALLOC_FAILED; }
    app->localFFTPlan = p;
    res = plan_setup(app, p);
    if (res != PFSOLVE_SUCCESS) { deletePfSolve(app); return res; }
    if (p->kind == PF_KIND_JW) {
        if (hipMalloc((void**)&p->d_err, 256) != hipSuccess) { deletePfSolve(app); return PFSOLVE_ERROR_FAILED_TO_ALLOCATE; }
        if (hipMemset(p->d_err, 0, 256) != hipSuccess) { deletePfSolve(app); return PFSOLVE_ERROR_FAILED_TO_COPY; }
        /* one table per stream slot */
        if (hipMalloc(&p->table, p->table_stride * cfg->num_streams) != hipSuccess) {
            deletePfSolve(app);
            return PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
        }
    }
    return PFSOLVE_SUCCESS;
}

/* JW step with fused block operators: complex systems in (block layout 4,
 * PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:2275-2371) and either complex
 * out (layout 7, optional accumulate) or the step's real systems out, each row scaled by the
 * block's operator (pfSolve_RegistersScaleC.h:27-201); the re/im halves of complex system q are
 * the step's pair of real systems 2q, 2q+1.  The input keep-range selects the rows read, the
 * output keep-range the rows written.  Runs on the compute / store split kernel. */
static PfSolveResult append_fused_blocks(PfSolveApplication* app, const PfSolveLaunchParams* lp,
                                         const jw_source_desc* src, jw_run_desc* run, int64_t nsys, hipStream_t s)
{
    struct pfsolve_plan* p = app->localFFTPlan;
    PfSolveConfiguration* cfg = &app->configuration;
    /* block codes, precision and N were checked when the plan was configured */
    const int64_t ib = cfg->inputBlock, ob = cfg->outputBlock;
    if (p->precision != JW_PREC_F64 || !p->split64) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    if (nsys % 2) return PFSOLVE_ERROR_EMPTY_size;
    if (!cfg->inputBuffer || !cfg->inputBuffer[0] || cfg->inputBuffer == cfg->buffer)
        return PFSOLVE_ERROR_EMPTY_inputBuffer;
    if (!lp) lp = &(const PfSolveLaunchParams){0};
    const int64_t m = cfg->jw_control_bitmask;
    const int64_t ncx = nsys / 2;
    run->fused_io = ob ? 1 : 2;
    run->in_c = cfg->inputBuffer[0];
    run->out_c = cfg->outputBuffer[0];
    run->rhs = cfg->outputBuffer[0];
    run->in_stride_c = (m & RUNTIME_INPUTBUFFERSTRIDE) ? (int64_t)lp->inputBufferStride : (int64_t)cfg->inputBufferStride[0];
    run->in_off_c = cfg->inputBlockOffset;
    run->out_stride_c = run->stride;
    run->out_off_c = run->offSol;
    run->in_op = (int32_t)(ib / 100);
    run->out_op = (int32_t)(ob / 100);
    run->out_acc = (ob / 10) % 10 == 2;
    run->in_lshift = (int32_t)cfg->inputBlockLshift;
    run->out_lshift = (int32_t)cfg->outputBlockLshift;
    run->in_scale = cfg->inputBlockScale;
    run->out_scale = cfg->outputBlockScale;
    run->vec2 = 1;
    /* complex input: 16-byte elements; real output (mode 2): 16-byte aligned pairs of rows */
    if ((uintptr_t)run->in_c % 16) return PFSOLVE_ERROR_EMPTY_inputBufferSize;
    if (run->fused_io == 1 ? ((uintptr_t)run->out_c % 16) != 0
                           : (((uintptr_t)run->rhs % 16) || (run->stride % 2) || (run->offSol % 2)))
        return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    if (run->in_stride_c < 0 || run->in_off_c < 0 || (ncx > 1 && run->in_stride_c < p->N))
        return PFSOLVE_ERROR_EMPTY_inputBufferSize;
    if (run->stride < 0 || run->offSol < 0 || (nsys > 1 && run->stride < p->N)) return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    if (ncx > 0) {
        const int64_t imax = (ncx - 1) * run->in_stride_c + run->in_off_c + p->N;   /* complex elements */
        if (cfg->inputBufferSize && (uint64_t)(imax * 16) > cfg->inputBufferSize[0]) return PFSOLVE_ERROR_EMPTY_inputBufferSize;
        /* bytes up to the last output element: complex elements (mode 1) or doubles (mode 2) */
        const int64_t oend = run->fused_io == 1 ? ((ncx - 1) * run->stride + run->offSol + p->N) * 16
                                                : ((nsys - 1) * run->stride + run->offSol + p->N) * 8;
        if (cfg->outputBufferSize && (uint64_t)oend > cfg->outputBufferSize[0]) return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    }
    {
        int64_t mmax, rmax;
        jw_index_extent(src, run, &mmax, &rmax);
        if (cfg->bufferSize && mmax >= 0 && (uint64_t)((mmax + 1) * 8) > cfg->bufferSize[0])
            return PFSOLVE_ERROR_EMPTY_bufferSize;
    }
    if (ncx <= 0) return PFSOLVE_SUCCESS;
    int64_t need = (ncx + p->split64 - 1) / p->split64;
    int64_t grid = (int64_t)p->cu_count * p->occupancy_split;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    run->wait_polls = p->wait_polls;
    if (jwdev_main64s(run, src, cfg->buffer[0], p->R64, (int)grid, p->d_err, s) != 0)
        return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    return PFSOLVE_SUCCESS;
}

/* Chained steps (numConsecutiveJWIterations = K > 1): the K iterations' tables in one
 * launch, then one launch of the chained kernel that keeps each pair of systems in
 * registers through the K iterations (pfSolve_JonesWorland.h:95-374).  Iteration i reads its
 * matrices from buffer[i] with useMultipleInputBuffers, from buffer[0] otherwise
 * (pfSolve_DispatchPlan.h:231-235); with useMultipleInputBuffers a declared bufferSize holds
 * one size per buffer. */
static PfSolveResult append_chain(PfSolveApplication* app, const PfSolveLaunchParams* lp, jw_run_desc* run,
                                  hipStream_t s, void* table)
{
    struct pfsolve_plan* p = app->localFFTPlan;
    PfSolveConfiguration* cfg = &app->configuration;
    const int K = p->chain_K;
    const int multi = cfg->useMultipleInputBuffers != 0;
    if (!lp) lp = &(const PfSolveLaunchParams){0};
    /* the output keep-range is that of the last iteration's rows */
    const int64_t m = cfg->jw_control_bitmask;
    const int64_t nlast = cfg->upperBanded ? p->N - K + 1 : p->N + K - 1;
    const int ozrt = (m & RUNTIME_OUTPUTZEROPAD) != 0;
    jw_keep_range(ozrt ? (int64_t)lp->outputZeropad[0] : cfg->outputZeropad[0],
                  ozrt ? (int64_t)lp->outputZeropad[1] : cfg->outputZeropad[1], ozrt, nlast, &run->out_lo,
                  &run->out_hi);
    jw_chain_jobs jobs;
    memset(&jobs, 0, sizeof(jobs));
    for (int i = 0; i < K; i++) {
        jw_source_desc* src = &jobs.src[i];
        PfSolveResult r = jw_resolve_chain_iteration(cfg, lp, i, src);
        if (r != PFSOLVE_SUCCESS) return r;
        r = jw_check_offsets(src, run);
        if (r != PFSOLVE_SUCCESS) return r;
        const void* mat = cfg->buffer[multi ? i : 0];
        if (!mat) return PFSOLVE_ERROR_EMPTY_buffer;
        jobs.matrices[i] = mat;
        int64_t mmax, rmax;
        jw_index_extent(src, run, &mmax, &rmax);
        if (cfg->bufferSize && mmax >= 0 && (uint64_t)((mmax + 1) * 8) > cfg->bufferSize[multi ? i : 0])
            return PFSOLVE_ERROR_EMPTY_bufferSize;
        if (i == 0 && cfg->outputBufferSize && rmax >= 0 && (uint64_t)((rmax + 1) * 8) > cfg->outputBufferSize[0])
            return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    }
    run->rhs = cfg->outputBuffer[0];
    run->vec2 = ((uintptr_t)run->rhs % 16) == 0 && (run->stride % 2) == 0 && (run->offSol % 2) == 0;
    if (!run->vec2) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;   /* 16-byte aligned pairs of systems */
    const int64_t nsys = run->nsys;
    if (nsys <= 0) return PFSOLVE_SUCCESS;
    jobs.table = table;
    jobs.tab_stride = (int64_t)p->chain_tab;
    jobs.K = K;
    jobs.R = p->chain_R;
    if (jwdev_prologue_chain(&jobs, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    const int64_t npass = (nsys + 1) / 2;
    int64_t need = (npass + 3) / 4;
    int64_t grid = (int64_t)p->cu_count * p->chain_occ;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    if (jwdev_chain64(run, table, (int64_t)(p->chain_tab / 8), K, p->chain_R, (int)grid, s) != 0)
        return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    return PFSOLVE_SUCCESS;
}

/* ALT variant (jw_type hundreds digit 1): rows couple with rows two apart, so the even and
 * the odd rows of every system are two independent systems of ceil(N/2) and floor(N/2)
 * rows (pfSolve_TridiagonalSolve.h:1622-1646, pfSolve_MatVecMul.h:295-325); their matrix
 * entries are interleaved in reorder_i_alt order (pfSolve_Plan_JonesWorland.h:144-219).
 * Each parity: a table from its entries, then the general kernel on its rows (two elements
 * apart); row 0 of the system (the even subsystem's first row) takes the row-0 scaleC. */
static PfSolveResult append_alt(PfSolveApplication* app, const jw_source_desc* full, const jw_run_desc* frun,
                                hipStream_t s, void* table)
{
    struct pfsolve_plan* p = app->localFFTPlan;
    PfSolveConfiguration* cfg = &app->configuration;
    PfSolveResult r = jw_check_offsets(full, frun);
    if (r != PFSOLVE_SUCCESS) return r;
    const int64_t N = p->N;
    const int64_t NB = jw_alt_stride(N, cfg->Msplit[0]);
    {
        /* matrix extent: every diagonal is a block of NB entries */
        int64_t mmax = -1;
        const int64_t bases[5] = {full->mdA, full->udA, full->ldA, full->ldS, full->udS};
        const int used[5] = {full->band_kind != JW_BAND_NONE,
                             full->band_kind == JW_BAND_UPPER, full->band_kind == JW_BAND_LOWER,
                             full->solve_kind == JW_SOLVE_LOWER || full->solve_kind == JW_SOLVE_TRI,
                             full->solve_kind == JW_SOLVE_UPPER || full->solve_kind == JW_SOLVE_TRI};
        for (int k = 0; k < 5; k++)
            if (used[k] && bases[k] + NB - 1 > mmax) mmax = bases[k] + NB - 1;
        if (cfg->bufferSize && mmax >= 0 && (uint64_t)((mmax + 1) * (p->precision == JW_PREC_DD ? 16 : 8)) >
                                                 cfg->bufferSize[0])
            return PFSOLVE_ERROR_EMPTY_bufferSize;
        int64_t dummy, rmax;
        jw_index_extent(full, frun, &dummy, &rmax);
        const int64_t elt = p->precision == JW_PREC_DD ? 16 : 8;
        if (cfg->outputBufferSize && rmax >= 0 && (uint64_t)((rmax + 1) * elt) > cfg->outputBufferSize[0])
            return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    }
    if (frun->nsys <= 0) return PFSOLVE_SUCCESS;
    for (int par = 0; par < 2; par++) {
        const int64_t n = (N + 1 - par) / 2;
        if (n <= 0) continue;
        jw_source_desc src = *full;
        src.N = n;
        src.alt = 1;
        src.alt_parity = par;
        src.alt_N = N;
        jw_run_desc run = *frun;
        run.rhs = cfg->outputBuffer[0];
        run.N = (int32_t)n;
        run.row_step = 2;
        run.vec2 = 0;
        run.offSol = frun->offSol + par;
        /* rows r = 2k + par inside [lo, hi) <=> k in [ceil((lo - par) / 2), ceil((hi - par) / 2)) */
        run.in_lo = frun->in_lo > par ? (frun->in_lo - par + 1) / 2 : 0;
        run.in_hi = frun->in_hi > par ? (frun->in_hi - par + 1) / 2 : 0;
        run.out_lo = frun->out_lo > par ? (frun->out_lo - par + 1) / 2 : 0;
        run.out_hi = frun->out_hi > par ? (frun->out_hi - par + 1) / 2 : 0;
        if (run.in_hi < run.in_lo) run.in_hi = run.in_lo;
        if (run.out_hi < run.out_lo) run.out_hi = run.out_lo;
        if (par == 1) run.apply_scale = 0;
        void* tab = (char*)table + (size_t)par * p->alt_tab;
        run.table = tab;
        if (jwdev_prologue(&src, cfg->buffer[0], tab, p->L, p->R, p->precision, 1, s) != 0)
            return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        const int G = 64 / p->L;
        const int64_t npass = (run.nsys + 2 * G - 1) / (2 * G);
        int64_t need = (npass + p->waves - 1) / p->waves;
        int64_t grid = (int64_t)p->cu_count * p->occupancy;
        if (grid > need) grid = need;
        if (grid < 1) grid = 1;
        if (jwdev_main(&run, p->L, p->R, p->precision, p->waves, (int)grid, s) != 0)
            return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    return PFSOLVE_SUCCESS;
}

/* Long systems in chunks: the tables of every chunk, the chunks' responses to unit
 * carries, then per slice of systems the reduction, the reduced solve and the finishing
 * pass (jw_long.hip). */
static PfSolveResult append_long(PfSolveApplication* app, const jw_source_desc* src, const jw_run_desc* run,
                                 hipStream_t s, void* slot_area)
{
    struct pfsolve_plan* p = app->localFFTPlan;
    PfSolveConfiguration* cfg = &app->configuration;
    if (!p->long_R) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    if (p->area_cap && p->table_bytes > p->area_cap) return PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER;
    if (run->nsys <= 0) return PFSOLVE_SUCCESS;
    jw_long_desc d;
    memset(&d, 0, sizeof(d));
    d.run = *run;
    d.table = (char*)slot_area + p->long_tab_off;
    d.tab_elems = p->long_tab_elems;
    d.red = (char*)slot_area + p->long_red_off;
    d.red_sys = p->long_red_sys;
    d.nchunks = (int32_t)p->long_nc;
    /* FP64 systems 16 bytes aligned (every slice then too): row pairs move as 16-byte accesses */
    d.run.vec2 = p->precision == JW_PREC_F64 && ((uintptr_t)run->rhs % 16) == 0 && (run->stride % 2) == 0 &&
                 (run->offSol % 2) == 0;
    if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    int* sync = (int*)((char*)slot_area + p->long_sync_off);
    void* gval = (char*)slot_area + p->long_gval_off;
    if (d.run.vec2 && p->long64_occ > 0 && cfg->kernelVariant != 1) {
        /* aligned FP64 systems: the chunks on the pair machinery (format-2 tables); the
         * table builds, COEF and the first slice's REDUCE in one launch (kernelVariant 1:
         * the chunk kernel, as for every other long system) */
        d.tab_elems = jwdev_table64_elems(8);
        for (int64_t s0 = 0; s0 < run->nsys; s0 += p->long_red_sys) {
            const int64_t n = run->nsys - s0 < p->long_red_sys ? run->nsys - s0 : p->long_red_sys;
            d.run.nsys = n;
            d.run.offSol = run->offSol + s0 * run->stride;
            d.mode = JW_LONG_REDUCE;
            if (s0 == 0) {
                if (jwdev_long64_front(src, cfg->buffer[0], &d, sync, gval, p->d_err, p->cu_count, p->long64_occ, s) != 0)
                    return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            } else if (jwdev_long64(&d, p->cu_count, p->long64_occ, s) != 0) {
                return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            }
            if (jwdev_long_reduced(&d, p->precision, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            d.mode = JW_LONG_FINISH;
            if (jwdev_long64(&d, p->cu_count, p->long64_occ, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        }
        if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
        return PFSOLVE_SUCCESS;
    }
    if (jwdev_long_prologue(src, cfg->buffer[0], (void*)d.table, d.tab_elems, p->precision, sync, gval, p->d_err,
                            s) != 0)
        return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    d.mode = JW_LONG_COEF;
    d.run.nsys = 3;
    if (jwdev_long_chunks(&d, p->precision, 1, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    /* workgroups per chunk: enough for two rounds of the whole GPU, at most one per 4 pairs
     * (a workgroup of 4 wavefronts, one pair of systems per wavefront pass) */
    const int64_t target = 2 * (int64_t)p->cu_count * p->long_occ;
    for (int64_t s0 = 0; s0 < run->nsys; s0 += p->long_red_sys) {
        const int64_t n = run->nsys - s0 < p->long_red_sys ? run->nsys - s0 : p->long_red_sys;
        d.run.nsys = n;
        d.run.offSol = run->offSol + s0 * run->stride;
        const int64_t need = ((n + 1) / 2 + 3) / 4;
        int64_t gx = (target + p->long_nc - 1) / p->long_nc;
        if (gx > need) gx = need;
        if (gx < 1) gx = 1;
        d.mode = JW_LONG_REDUCE;
        if (jwdev_long_chunks(&d, p->precision, (int)gx, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        if (jwdev_long_reduced(&d, p->precision, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        d.mode = JW_LONG_FINISH;
        if (jwdev_long_chunks(&d, p->precision, (int)gx, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    return PFSOLVE_SUCCESS;
}

static PfSolveResult append_impl(PfSolveApplication* app, const PfSolveLaunchParams* lp, uint64_t slot)
{
    struct pfsolve_plan* p = app->localFFTPlan;
    PfSolveConfiguration* cfg = &app->configuration;
    void* const table = slot_table(p, slot);
    /* buffer swap and checks: PfSolve/pfSolve_PlanManagement/pfSolve_API_handles/pfSolve_UpdateBuffers.h:616-699 */
    if (lp) {
        if (lp->buffer) cfg->buffer = lp->buffer;
        if (lp->outputBuffer) cfg->outputBuffer = lp->outputBuffer;
        if (lp->inputBuffer) cfg->inputBuffer = lp->inputBuffer;
        if (lp->kernel) cfg->kernel = lp->kernel;
    }
    if (cfg->inputBuffer == NULL) cfg->inputBuffer = cfg->buffer;
    if (cfg->outputBuffer == NULL) cfg->outputBuffer = cfg->buffer;
    if (cfg->buffer == NULL || cfg->buffer[0] == NULL) return PFSOLVE_ERROR_EMPTY_buffer;
    if (cfg->outputBuffer == NULL || cfg->outputBuffer[0] == NULL) return PFSOLVE_ERROR_EMPTY_outputBuffer;
    /* batchSize > 0 replaces size[1] (pfSolve_UpdateBuffers.h:1558) */
    if (lp && lp->batchSize > 0) cfg->size[1] = lp->batchSize;
    hipStream_t s = slot_stream(app, slot);

    if (p->kind == PF_KIND_JW) {
        const int64_t nsys = (int64_t)(cfg->size[1] * cfg->size[2]);
        jw_source_desc src;
        jw_run_desc run;
        PfSolveResult r = jw_resolve_launch(cfg, lp, nsys, &src, &run);
        if (r != PFSOLVE_SUCCESS) return r;
        if (p->chain_K > 1) return append_chain(app, lp, &run, s, table);
        if (p->alt) return append_alt(app, &src, &run, s, table);
        /* offsets and strides: no matrix diagonal or system may start before its buffer */
        r = jw_check_offsets(&src, &run);
        if (r != PFSOLVE_SUCCESS) return r;
        run.table = table;
        if (cfg->inputBlock || cfg->outputBlock) return append_fused_blocks(app, lp, &src, &run, nsys, s);
        /* index guard: every element the kernels touch must lie inside the
         * buffers whose sizes the caller declared (bufferSize/outputBufferSize) */
        {
            int64_t mmax, rmax;
            jw_index_extent(&src, &run, &mmax, &rmax);
            const int64_t elt = p->precision == JW_PREC_DD ? 16 : 8;
            if (cfg->bufferSize && mmax >= 0 && (uint64_t)((mmax + 1) * elt) > cfg->bufferSize[0])
                return PFSOLVE_ERROR_EMPTY_bufferSize;
            if (cfg->outputBufferSize && rmax >= 0 && (uint64_t)((rmax + 1) * elt) > cfg->outputBufferSize[0])
                return PFSOLVE_ERROR_EMPTY_outputBufferSize;
        }
        run.rhs = cfg->outputBuffer[0];
        if (p->long_only) return append_long(app, &src, &run, s, table);
        const uintptr_t base = (uintptr_t)run.rhs;
        run.vec2 = p->precision == JW_PREC_F64 && (base % 16) == 0 && (run.stride % 2) == 0 && (run.offSol % 2) == 0;
        if (p->R64 && run.vec2 && cfg->kernelVariant != 1) {
            /* wave-per-system kernel: table format 2, built by a prologue launch or,
             * in the compute / store split kernel, by every workgroup in LDS */
            if (nsys <= 0) return PFSOLVE_SUCCESS;
            if (!p->split64 && jwdev_prologue(&src, cfg->buffer[0], table, 64, p->R64, p->precision, 2, s) != 0)
                return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            /* wavefront passes: pairs of systems, single systems for R = 32 */
            const int64_t npass = p->R64 == 32 ? nsys : (nsys + 1) / 2;
            const int per_wg = p->split64 ? p->split64 : p->waves64;
            int64_t need = (npass + per_wg - 1) / per_wg;
            int64_t grid = (int64_t)p->cu_count * (p->split64 ? p->occupancy_split : p->occupancy64);
            if (grid > need) grid = need;
            if (grid < 1) grid = 1;
            if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
            if (p->split64) {
                run.wait_polls = p->wait_polls;
                if (jwdev_main64s(&run, &src, cfg->buffer[0], p->R64, (int)grid, p->d_err, s) != 0)
                    return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            } else if (jwdev_main64(&run, p->R64, p->waves64, (int)grid, s) != 0) {
                return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            }
            if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
            return PFSOLVE_SUCCESS;
        }
        if (p->dd64_R) {
            /* double-double wave-per-system kernel: a format-2 table in double-double */
            if (nsys <= 0) return PFSOLVE_SUCCESS;
            if (jwdev_prologue(&src, cfg->buffer[0], table, 64, p->dd64_R, p->precision, 2, s) != 0)
                return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            const int64_t npass = (nsys + 1) / 2;
            int64_t need = (npass + p->dd64_waves - 1) / p->dd64_waves;
            int64_t grid = (int64_t)p->cu_count * p->dd64_occ;
            if (grid > need) grid = need;
            if (grid < 1) grid = 1;
            if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
            if (jwdev_dd64(&run, p->dd64_R, p->dd64_waves, (int)grid, s) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
            return PFSOLVE_SUCCESS;
        }
        /* 1024 < N <= 2048 without 16-byte aligned systems (or with the general kernel
         * asked for): in chunks */
        if (p->L == 0) return append_long(app, &src, &run, s, table);
        if (jwdev_prologue(&src, cfg->buffer[0], table, p->L, p->R, p->precision, 1, s) != 0)
            return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        if (nsys <= 0) return PFSOLVE_SUCCESS;
        const int G = 64 / p->L;
        const int64_t npass = (nsys + 2 * G - 1) / (2 * G);
        int64_t need = (npass + p->waves - 1) / p->waves;
        int64_t grid = (int64_t)p->cu_count * p->occupancy;
        if (grid > need) grid = need;
        if (grid < 1) grid = 1;
        if (p->ev_start && hipEventRecord(p->ev_start, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
        if (jwdev_main(&run, p->L, p->R, p->precision, p->waves, (int)grid, s) != 0)
            return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        if (p->ev_stop && hipEventRecord(p->ev_stop, s) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
        return PFSOLVE_SUCCESS;
    }
    if (p->kind == PF_KIND_BLOCK) {
        jw_block_desc d;
        PfSolveResult r = jw_resolve_block(cfg, lp, (int64_t)cfg->size[1], &d);
        if (r != PFSOLVE_SUCCESS) return r;
        d.in = cfg->buffer[0];
        d.out = cfg->outputBuffer[0];
        if (d.N * d.B <= 0) return PFSOLVE_SUCCESS;
        if (d.inStride < 0 || d.outStride < 0 || d.offIn < 0 || d.offOut < 0) return PFSOLVE_ERROR_EMPTY_bufferSize;
        {
            /* doubles touched: complex buffers count 2 per element */
            const int L = d.layout;
            int64_t in_max = -1, out_max = -1;
            if (d.in_hi > 0 && !d.skip_read) {
                int64_t e = (d.B - 1) * d.inStride + d.offIn;
                if (L == 2 || L == 3 || L == 4) in_max = 2 * (e + d.in_hi - 1) + 1;
                else if (L == 7) in_max = 2 * e + d.in_hi - 1 + d.inStride;
                else in_max = e + d.in_hi - 1;
            }
            if (d.out_hi > 0) {
                int64_t e = (d.B - 1) * d.outStride + d.offOut;
                if (L == 4) out_max = 2 * e + d.out_hi - 1 + d.outStride;
                else if (L == 5 || L == 6 || L == 7) out_max = 2 * (e + d.out_hi - 1) + 1;
                else out_max = e + d.out_hi - 1;
            }
            if (cfg->bufferSize && in_max >= 0 && (uint64_t)((in_max + 1) * 8) > cfg->bufferSize[0])
                return PFSOLVE_ERROR_EMPTY_bufferSize;
            if (cfg->outputBufferSize && out_max >= 0 && (uint64_t)((out_max + 1) * 8) > cfg->outputBufferSize[0])
                return PFSOLVE_ERROR_EMPTY_outputBufferSize;
        }
        return jwdev_block(&d, s) == 0 ? PFSOLVE_SUCCESS : PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    if (p->kind == PF_KIND_GBMV) {
        jw_gbmv_desc d;
        PfSolveResult r = jw_resolve_gbmv(cfg, lp, (int64_t)cfg->size[1], &d);
        if (r != PFSOLVE_SUCCESS) return r;
        d.mat = cfg->buffer[0];
        d.rhs = cfg->outputBuffer[0];
        if (d.nsys <= 0) return PFSOLVE_SUCCESS;
        if (d.offM < 0 || d.offSol < 0 || d.stride < 0 || (d.nsys > 1 && d.stride < d.N))
            return PFSOLVE_ERROR_EMPTY_outputBufferSize;
        if (cfg->bufferSize && (uint64_t)((d.offM + (int64_t)d.LDA * d.N) * 8) > cfg->bufferSize[0])
            return PFSOLVE_ERROR_EMPTY_bufferSize;
        if (cfg->outputBufferSize && (uint64_t)(((d.nsys - 1) * d.stride + d.offSol + d.N) * 8) > cfg->outputBufferSize[0])
            return PFSOLVE_ERROR_EMPTY_outputBufferSize;
        return jwdev_gbmv(&d, s) == 0 ? PFSOLVE_SUCCESS : PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
}

PfSolveResult PfSolveAppend(PfSolveApplication* app, int inverse, PfSolveLaunchParams* lp)
{
    (void)inverse;
    if (!app || !app->localFFTPlan) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    PfSolveConfiguration* cfg = &app->configuration;
    const uint64_t slot = plan_slot(app);
    PfSolveResult r = append_impl(app, lp, slot);
    if (r != PFSOLVE_SUCCESS || !cfg->stream_event) return r;
    /* round robin: the slot's event marks its last launch, the next launch takes the
     * next stream (pfSolve_DispatchPlan.h:355-362) */
    if (hipEventRecord(cfg->stream_event[slot], slot_stream(app, slot)) != hipSuccess)
        return PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD;
    cfg->streamCounter++;
    cfg->streamID = cfg->streamCounter % cfg->num_streams;
    return PFSOLVE_SUCCESS;
}

/* initializePfSolve + PfSolveAppend + deletePfSolve with the plan on the stack and the
 * device table and error flag in the library's static device scratch: no allocation. */
PfSolveResult pfsolve_append_noalloc(const PfSolveConfiguration* configuration,
                                     const PfSolveLaunchParams* launchParams)
{
    if (!configuration) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    if (configuration->device == NULL) return PFSOLVE_ERROR_INVALID_DEVICE;
    if (configuration->num_streams > 1) return PFSOLVE_ERROR_INVALID_QUEUE;
    PfSolveApplication app;
    memset(&app, 0, sizeof(app));
    PfSolveResult res = jw_configure(configuration, &app.configuration);
    if (res != PFSOLVE_SUCCESS) return res;
    app.configuration.stream_event = NULL;
    app.configuration.streamID = 0;
    app.configuration.streamCounter = 0;
    app.configuration.num_streams = 1;
    struct pfsolve_plan plan;
    memset(&plan, 0, sizeof(plan));
    res = plan_setup(&app, &plan);
    if (res != PFSOLVE_SUCCESS) return res;
    if (plan.kind == PF_KIND_JW) {
        void* table = NULL;
        int* err = NULL;
        size_t cap = 0;
        if (jwdev_noalloc_scratch(&table, &err, &cap) != 0) return PFSOLVE_ERROR_INVALID_DEVICE;
        /* a chunk fallback that does not fit is refused when a launch needs it */
        const size_t need = plan.main_bytes ? plan.main_bytes : plan.table_stride;
        if (need > cap) return PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER;
        plan.area_cap = cap;
        plan.table = table;
        plan.d_err = err;
        /* the flag starts at zero for this call (a fill kernel: callers may capture it) */
        if (jwdev_zero_counters(err, 1, slot_stream(&app, 0)) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    app.localFFTPlan = &plan;
    return append_impl(&app, launchParams, 0);
}

size_t pfsolve_noalloc_scratch_bytes(void) { return 144 * 1024 + 256; }

/* Joins the plan's work and reports a raised device error flag (a bounded wait inside a
 * kernel ran out).  With num_streams > 1 it waits for every stream's event and restarts
 * the round robin, as pfSolve_RunApp.h:28-57 does; with one stream it synchronises that
 * stream (the reference returns at once there; the wait is what lets the flag be read). */
PfSolveResult PfSolveSync(PfSolveApplication* app)
{
    if (!app) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    struct pfsolve_plan* p = app->localFFTPlan;
    if (!p) return PFSOLVE_SUCCESS;   /* as the reference: nothing to join */
    PfSolveConfiguration* cfg = &app->configuration;
    if (cfg->stream_event) {
        for (uint64_t i = 0; i < cfg->num_streams; i++)
            if (hipEventSynchronize(cfg->stream_event[i]) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
        cfg->streamCounter = 0;
        cfg->streamID = 0;
    } else if (hipStreamSynchronize(slot_stream(app, 0)) != hipSuccess) {
        return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
    }
    if (p->d_err) {
        int e = 0;
        if (hipMemcpy(&e, p->d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
        if (e) {
            (void)hipMemset(p->d_err, 0, sizeof(int));
            return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
        }
    }
    return PFSOLVE_SUCCESS;
}

/* The reference returns the literal 00102, an octal constant (pfSolve.h:101-103). */
int PfSolveGetVersion(void) { return 00102; }

size_t pfsolve_sizeof_ext(const char* type_name)
{
    static const struct { const char* n; size_t s; } k[] = {
        {"pfsolve_geometry", sizeof(pfsolve_geometry)},
        {"pfsolve_batch_set", sizeof(pfsolve_batch_set)},
        {"pfsolve_batch_tuning", sizeof(pfsolve_batch_tuning)},
        {"pfsolve_batch_info", sizeof(pfsolve_batch_info)},
        {"pfsolve_serve_config", sizeof(pfsolve_serve_config)},
        {"pfsolve_serve_request", sizeof(pfsolve_serve_request)},
        {"pfsolve_serve_stats", sizeof(pfsolve_serve_stats)},
    };
    if (!type_name) return 0;
    for (size_t i = 0; i < sizeof(k) / sizeof(k[0]); i++)
        if (strcmp(k[i].n, type_name) == 0) return k[i].s;
    return 0;
}

typedef struct { int code; const char* name; } err_name;
static const err_name k_err_names[] = {
    {0, "PFSOLVE_SUCCESS"},
    {1, "PFSOLVE_ERROR_MALLOC_FAILED"},
    {2, "PFSOLVE_ERROR_INSUFFICIENT_CODE_BUFFER"},
    {3, "PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER"},
    {4, "PFSOLVE_ERROR_PLAN_NOT_INITIALIZED"},
    {5, "PFSOLVE_ERROR_NULL_TEMP_PASSED"},
    {1001, "PFSOLVE_ERROR_INVALID_PHYSICAL_DEVICE"},
    {1002, "PFSOLVE_ERROR_INVALID_DEVICE"},
    {1003, "PFSOLVE_ERROR_INVALID_QUEUE"},
    {1004, "PFSOLVE_ERROR_INVALID_COMMAND_POOL"},
    {1005, "PFSOLVE_ERROR_INVALID_FENCE"},
    {1006, "PFSOLVE_ERROR_ONLY_FORWARD_FFT_INITIALIZED"},
    {1007, "PFSOLVE_ERROR_ONLY_INVERSE_FFT_INITIALIZED"},
    {1008, "PFSOLVE_ERROR_INVALID_CONTEXT"},
    {1009, "PFSOLVE_ERROR_INVALID_PLATFORM"},
    {1010, "PFSOLVE_ERROR_ENABLED_saveApplicationToString"},
    {1011, "PFSOLVE_ERROR_EMPTY_FILE"},
    {2001, "PFSOLVE_ERROR_EMPTY_FFTdim"},
    {2002, "PFSOLVE_ERROR_EMPTY_size"},
    {2003, "PFSOLVE_ERROR_EMPTY_bufferSize"},
    {2004, "PFSOLVE_ERROR_EMPTY_buffer"},
    {2005, "PFSOLVE_ERROR_EMPTY_tempBufferSize"},
    {2006, "PFSOLVE_ERROR_EMPTY_tempBuffer"},
    {2007, "PFSOLVE_ERROR_EMPTY_inputBufferSize"},
    {2008, "PFSOLVE_ERROR_EMPTY_inputBuffer"},
    {2009, "PFSOLVE_ERROR_EMPTY_outputBufferSize"},
    {2010, "PFSOLVE_ERROR_EMPTY_outputBuffer"},
    {2011, "PFSOLVE_ERROR_EMPTY_kernelSize"},
    {2012, "PFSOLVE_ERROR_EMPTY_kernel"},
    {2013, "PFSOLVE_ERROR_EMPTY_applicationString"},
    {2014, "PFSOLVE_ERROR_EMPTY_useCustomBluesteinPaddingPattern_arrays"},
    {3001, "PFSOLVE_ERROR_UNSUPPORTED_RADIX"},
    {3002, "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH"},
    {3003, "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH_R2C"},
    {3004, "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH_DCT"},
    {3005, "PFSOLVE_ERROR_UNSUPPORTED_FFT_OMIT"},
    {4001, "PFSOLVE_ERROR_FAILED_TO_ALLOCATE"},
    {4002, "PFSOLVE_ERROR_FAILED_TO_MAP_MEMORY"},
    {4003, "PFSOLVE_ERROR_FAILED_TO_ALLOCATE_COMMAND_BUFFERS"},
    {4004, "PFSOLVE_ERROR_FAILED_TO_BEGIN_COMMAND_BUFFER"},
    {4005, "PFSOLVE_ERROR_FAILED_TO_END_COMMAND_BUFFER"},
    {4006, "PFSOLVE_ERROR_FAILED_TO_SUBMIT_QUEUE"},
    {4007, "PFSOLVE_ERROR_FAILED_TO_WAIT_FOR_FENCES"},
    {4008, "PFSOLVE_ERROR_FAILED_TO_RESET_FENCES"},
    {4009, "PFSOLVE_ERROR_FAILED_TO_CREATE_DESCRIPTOR_POOL"},
    {4010, "PFSOLVE_ERROR_FAILED_TO_CREATE_DESCRIPTOR_SET_LAYOUT"},
    {4011, "PFSOLVE_ERROR_FAILED_TO_ALLOCATE_DESCRIPTOR_SETS"},
    {4012, "PFSOLVE_ERROR_FAILED_TO_CREATE_PIPELINE_LAYOUT"},
    {4013, "PFSOLVE_ERROR_FAILED_SHADER_PREPROCESS"},
    {4014, "PFSOLVE_ERROR_FAILED_SHADER_PARSE"},
    {4015, "PFSOLVE_ERROR_FAILED_SHADER_LINK"},
    {4016, "PFSOLVE_ERROR_FAILED_SPIRV_GENERATE"},
    {4017, "PFSOLVE_ERROR_FAILED_TO_CREATE_SHADER_MODULE"},
    {4018, "PFSOLVE_ERROR_FAILED_TO_CREATE_INSTANCE"},
    {4019, "PFSOLVE_ERROR_FAILED_TO_SETUP_DEBUG_MESSENGER"},
    {4020, "PFSOLVE_ERROR_FAILED_TO_FIND_PHYSICAL_DEVICE"},
    {4021, "PFSOLVE_ERROR_FAILED_TO_CREATE_DEVICE"},
    {4022, "PFSOLVE_ERROR_FAILED_TO_CREATE_FENCE"},
    {4023, "PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_POOL"},
    {4024, "PFSOLVE_ERROR_FAILED_TO_CREATE_BUFFER"},
    {4025, "PFSOLVE_ERROR_FAILED_TO_ALLOCATE_MEMORY"},
    {4026, "PFSOLVE_ERROR_FAILED_TO_BIND_BUFFER_MEMORY"},
    {4027, "PFSOLVE_ERROR_FAILED_TO_FIND_MEMORY"},
    {4028, "PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE"},
    {4029, "PFSOLVE_ERROR_FAILED_TO_COPY"},
    {4030, "PFSOLVE_ERROR_FAILED_TO_CREATE_PROGRAM"},
    {4031, "PFSOLVE_ERROR_FAILED_TO_COMPILE_PROGRAM"},
    {4032, "PFSOLVE_ERROR_FAILED_TO_GET_CODE_SIZE"},
    {4033, "PFSOLVE_ERROR_FAILED_TO_GET_CODE"},
    {4034, "PFSOLVE_ERROR_FAILED_TO_DESTROY_PROGRAM"},
    {4035, "PFSOLVE_ERROR_FAILED_TO_LOAD_MODULE"},
    {4036, "PFSOLVE_ERROR_FAILED_TO_GET_FUNCTION"},
    {4037, "PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY"},
    {4038, "PFSOLVE_ERROR_FAILED_TO_MODULE_GET_GLOBAL"},
    {4039, "PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL"},
    {4040, "PFSOLVE_ERROR_FAILED_TO_EVENT_RECORD"},
    {4041, "PFSOLVE_ERROR_FAILED_TO_ADD_NAME_EXPRESSION"},
    {4042, "PFSOLVE_ERROR_FAILED_TO_INITIALIZE"},
    {4043, "PFSOLVE_ERROR_FAILED_TO_SET_DEVICE_ID"},
    {4044, "PFSOLVE_ERROR_FAILED_TO_GET_DEVICE"},
    {4045, "PFSOLVE_ERROR_FAILED_TO_CREATE_CONTEXT"},
    {4046, "PFSOLVE_ERROR_FAILED_TO_CREATE_PIPELINE"},
    {4047, "PFSOLVE_ERROR_FAILED_TO_SET_KERNEL_ARG"},
    {4048, "PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_QUEUE"},
    {4049, "PFSOLVE_ERROR_FAILED_TO_RELEASE_COMMAND_QUEUE"},
    {4050, "PFSOLVE_ERROR_FAILED_TO_ENUMERATE_DEVICES"},
    {4051, "PFSOLVE_ERROR_FAILED_TO_GET_ATTRIBUTE"},
    {4052, "PFSOLVE_ERROR_FAILED_TO_CREATE_EVENT"},
    {4053, "PFSOLVE_ERROR_FAILED_TO_CREATE_COMMAND_LIST"},
    {4054, "PFSOLVE_ERROR_FAILED_TO_DESTROY_COMMAND_LIST"},
    {4055, "PFSOLVE_ERROR_FAILED_TO_SUBMIT_BARRIER"},
};

/* Same strings as the reference's switch (pfSolve_Structs.h:545-733), which has no
 * case for PFSOLVE_ERROR_MATH_FAILED. */
const char* getPfSolveErrorString(PfSolveResult result)
{
    for (size_t i = 0; i < sizeof(k_err_names) / sizeof(k_err_names[0]); i++)
        if (k_err_names[i].code == (int)result) return k_err_names[i].name;
    return "Unknown PfSolve error";
}

/* ---- extension introspection ---- */
size_t pfsolve_sizeof_configuration(void) { return sizeof(PfSolveConfiguration); }
size_t pfsolve_sizeof_launch_params(void) { return sizeof(PfSolveLaunchParams); }
size_t pfsolve_sizeof_application(void) { return sizeof(PfSolveApplication); }

#define OFF(S, f) {#f, offsetof(S, f)}
typedef struct { const char* name; size_t off; } field_off;
static const field_off k_cfg_fields[] = {
    OFF(PfSolveConfiguration, FFTdim), OFF(PfSolveConfiguration, size), OFF(PfSolveConfiguration, device),
    OFF(PfSolveConfiguration, stream), OFF(PfSolveConfiguration, num_streams),
    OFF(PfSolveConfiguration, bufferSize), OFF(PfSolveConfiguration, outputBufferSize),
    OFF(PfSolveConfiguration, kernelSize), OFF(PfSolveConfiguration, buffer),
    OFF(PfSolveConfiguration, inputBuffer), OFF(PfSolveConfiguration, outputBuffer),
    OFF(PfSolveConfiguration, kernel), OFF(PfSolveConfiguration, upperBanded),
    OFF(PfSolveConfiguration, jw_type), OFF(PfSolveConfiguration, block), OFF(PfSolveConfiguration, lshift),
    OFF(PfSolveConfiguration, LDA), OFF(PfSolveConfiguration, KU), OFF(PfSolveConfiguration, KL),
    OFF(PfSolveConfiguration, ALT_dgbmv), OFF(PfSolveConfiguration, Msplit),
    OFF(PfSolveConfiguration, jw_control_bitmask), OFF(PfSolveConfiguration, M_size),
    OFF(PfSolveConfiguration, offsetM), OFF(PfSolveConfiguration, offsetV),
    OFF(PfSolveConfiguration, offsetSolution), OFF(PfSolveConfiguration, inputZeropad),
    OFF(PfSolveConfiguration, outputZeropad), OFF(PfSolveConfiguration, scaleC),
    OFF(PfSolveConfiguration, doublePrecision), OFF(PfSolveConfiguration, quadDoubleDoublePrecision),
    OFF(PfSolveConfiguration, quadDoubleDoublePrecisionDoubleMemory), OFF(PfSolveConfiguration, halfPrecision),
    OFF(PfSolveConfiguration, isInputFormatted), OFF(PfSolveConfiguration, isOutputFormatted),
    OFF(PfSolveConfiguration, inputBufferStride), OFF(PfSolveConfiguration, outputBufferStride),
    OFF(PfSolveConfiguration, keepShaderCode), OFF(PfSolveConfiguration, disableCaching),
    OFF(PfSolveConfiguration, numConsecutiveJWIterations), OFF(PfSolveConfiguration, useMultipleInputBuffers),
    OFF(PfSolveConfiguration, fixMaxRadixBluestein), OFF(PfSolveConfiguration, fixMaxRaderPrimeFFT),
    OFF(PfSolveConfiguration, kernelConvolution), OFF(PfSolveConfiguration, registerBoost),
    OFF(PfSolveConfiguration, localPageSize), OFF(PfSolveConfiguration, halfThreads),
    OFF(PfSolveConfiguration, reorderFourStep), OFF(PfSolveConfiguration, useRaderUintLUT),
    OFF(PfSolveConfiguration, vendorID), OFF(PfSolveConfiguration, useStrict32BitAddress),
    OFF(PfSolveConfiguration, offsetT), OFF(PfSolveConfiguration, autotune),
    OFF(PfSolveConfiguration, kernelVariant), OFF(PfSolveConfiguration, inputBlock),
    OFF(PfSolveConfiguration, outputBlock), OFF(PfSolveConfiguration, inputBlockLshift),
    OFF(PfSolveConfiguration, outputBlockLshift), OFF(PfSolveConfiguration, inputBlockScale),
    OFF(PfSolveConfiguration, outputBlockScale), OFF(PfSolveConfiguration, inputBlockOffset),
};
static const field_off k_lp_fields[] = {
    OFF(PfSolveLaunchParams, buffer), OFF(PfSolveLaunchParams, outputBuffer), OFF(PfSolveLaunchParams, kernel),
    OFF(PfSolveLaunchParams, M_size), OFF(PfSolveLaunchParams, offsetM), OFF(PfSolveLaunchParams, offsetV),
    OFF(PfSolveLaunchParams, offsetSolution), OFF(PfSolveLaunchParams, inputZeropad),
    OFF(PfSolveLaunchParams, outputZeropad), OFF(PfSolveLaunchParams, scaleC),
    OFF(PfSolveLaunchParams, inputBufferStride), OFF(PfSolveLaunchParams, outputBufferStride),
    OFF(PfSolveLaunchParams, batchSize), OFF(PfSolveLaunchParams, offsetT),
};

static int64_t find_off(const field_off* t, size_t n, const char* name)
{
    if (!name) return -1;
    for (size_t i = 0; i < n; i++)
        if (strcmp(t[i].name, name) == 0) return (int64_t)t[i].off;
    return -1;
}
int64_t pfsolve_offsetof_configuration(const char* field)
{
    return find_off(k_cfg_fields, sizeof(k_cfg_fields) / sizeof(k_cfg_fields[0]), field);
}
int64_t pfsolve_offsetof_launch_params(const char* field)
{
    return find_off(k_lp_fields, sizeof(k_lp_fields) / sizeof(k_lp_fields[0]), field);
}

PfSolveResult pfsolve_plan_geometry(const PfSolveApplication* app, pfsolve_geometry* g)
{
    if (!app || !app->localFFTPlan || !g) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    const struct pfsolve_plan* p = app->localFFTPlan;
    memset(g, 0, sizeof(*g));
    g->lanes = p->L;
    g->rows_per_lane = p->R;
    g->waves = p->waves;
    g->precision = p->precision;
    g->table_bytes = (int64_t)p->table_bytes;
    g->lds_bytes = p->lds_bytes;
    if (p->long_only) {
        /* long systems: the chunk kernel's geometry, every chunk's table */
        g->lanes = 64;
        g->rows_per_lane = p->long_R;
        g->table_bytes = p->long_ntab * p->long_tab_elems * (p->precision == JW_PREC_F64 ? 8 : 16);
        return PFSOLVE_SUCCESS;
    }
    g->fast_rows_per_lane = p->R64;
    g->fast_waves = p->split64 ? p->split64 : p->waves64;
    if (p->dd64_R) {
        g->fast_rows_per_lane = p->dd64_R;
        g->fast_waves = p->dd64_waves;
    }
    if (p->split64) {
        int c = 0, st = 0;
        if (jwdev_main64s_shape(p->R64, &c, &st)) g->fast_store_waves = st;
    }
    return PFSOLVE_SUCCESS;
}

/* reorder_i0: PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142 */
int64_t pfsolve_reorder_i0(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0)
{
    if (Msplit0 == 0) Msplit0 = 704;
    int64_t size0 = N, size1 = Nalign ? Nalign : N;
    if (N > 3840) { size0 = Msplit0; size1 = Msplit0; }
    int64_t ws = ((size1 + 1023) / 1024) * 64;
    if (size1 < 512) ws = 64;
    int64_t R = (size1 + ws - 1) / ws;
    if (R % 2 == 0) R++;
    int64_t ret_off = (i / size0) * size0, li = i % size0;
    if (((i / size0) + 1) * size0 > N) size0 = N % size0;
    int64_t ceilq = (size0 + R - 1) / R, ret = li / R;
    if ((li % R) < (size0 % R)) ret += (li % R) * ceilq;
    else ret += (size0 % R) * ceilq + ((li % R) - (size0 % R)) * (size0 / R);
    return ret + ret_off;
}

int pfsolve_kernel_supported(int lanes, int rows_per_lane, int precision)
{
    return jwdev_supported(lanes, rows_per_lane, precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD);
}

const char* pfsolve_last_error(void) { return jwdev_last_error(); }

PfSolveResult pfsolve_set_wait_polls(PfSolveApplication* app, int32_t polls)
{
    if (!app || !app->localFFTPlan) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    app->localFFTPlan->wait_polls = polls;
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_set_kernel_events(PfSolveApplication* app, hipEvent_t start, hipEvent_t stop)
{
    if (!app || !app->localFFTPlan) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    app->localFFTPlan->ev_start = start;
    app->localFFTPlan->ev_stop = stop;
    return PFSOLVE_SUCCESS;
}
