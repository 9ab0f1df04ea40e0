/*
 * jw_plan.c -- configuration defaults, validation, geometry and the per-launch
 * resolution of the reference's JW modes into device descriptors.  Plain C,
 * no device calls (unit-tested on the CPU through pfsolve_validate_configuration).
 */
#include "jw_plan.h"

#include <string.h>

#define JW_MAX_N_F64 2048   /* N > 1024: wave-per-system long kernel only */
#define JW_MAX_N_DD 512
#define JW_MAX_N_FUSED_BLOCKS 512   /* fused block I/O runs on the compute / store split kernel */
#define JW_MAX_N_CHAIN 1024         /* chained steps: R <= 16 rows per lane on one wavefront */
#define JW_MAX_N_LONG (1 << 22)     /* long systems: chunks of 512 (FP64) or 256 (DD) rows */

int jw_is_chunked(int64_t N, int precision)
{
    return precision == JW_PREC_F64 ? N > JW_MAX_N_F64 : N > JW_MAX_N_DD;
}

int jw_plan_kind(const PfSolveConfiguration* cfg)
{
    int n = 0, kind = PF_KIND_NONE;
    if (cfg->jw_type >= 10) { n++; kind = PF_KIND_JW; }
    if (cfg->block) { n++; kind = PF_KIND_BLOCK; }
    if (cfg->LDA) { n++; kind = PF_KIND_GBMV; }
    return n == 1 ? kind : (n == 0 ? PF_KIND_NONE : -1);
}

int jw_precision(const PfSolveConfiguration* cfg)
{
    if (cfg->quadDoubleDoublePrecisionDoubleMemory) return JW_PREC_DD_F64MEM;
    if (cfg->quadDoubleDoublePrecision) return JW_PREC_DD;
    if (cfg->doublePrecision) return JW_PREC_F64;
    return 0;
}

/* Defaults of setConfigurationPfSolve (PfSolve/pfSolve_AppManagement/pfSolve_InitializeApp.h:981-1212) */
PfSolveResult jw_configure(const PfSolveConfiguration* in, PfSolveConfiguration* out)
{
    *out = *in;
    out->FFTdim = 1;
    if (in->size[0] == 0) return PFSOLVE_ERROR_EMPTY_size;
    if (in->bufferStride[0] == 0) out->bufferStride[0] = in->size[0];
    if (in->inputBufferStride[0] == 0) out->inputBufferStride[0] = in->size[0];
    if (in->outputBufferStride[0] == 0) out->outputBufferStride[0] = in->size[0];
    for (int i = 1; i < 3; i++) {
        out->size[i] = in->size[i] ? in->size[i] : 1;
        if (in->bufferStride[i] == 0) out->bufferStride[i] = out->bufferStride[i - 1] * out->size[i];
        if (in->inputBufferStride[i] == 0) out->inputBufferStride[i] = out->inputBufferStride[i - 1] * out->size[i];
        if (in->outputBufferStride[i] == 0) out->outputBufferStride[i] = out->outputBufferStride[i - 1] * out->size[i];
    }
    if (out->numConsecutiveJWIterations == 0) out->numConsecutiveJWIterations = 1;
    out->M_size = (int64_t)in->size[0];
    if (out->Msplit[0] == 0) out->Msplit[0] = 704;
    out->warpSize = 64;
    out->vendorID = 0x1002;

    const int kind = jw_plan_kind(out);
    if (kind <= 0) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
    const int prec = jw_precision(out);
    if (prec == 0) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;       /* FP32/FP16 are not built */
    const int64_t N = out->M_size;

    if (kind == PF_KIND_JW) {
        const int alt = (out->jw_type % 1000) / 100 == 1;     /* hundreds digit: ALT */
        const int jwt = alt ? out->jw_type % 100 : out->jw_type % 1000;
        if (jwt < 10 || jwt > 14 || (out->jw_type % 1000) / 100 > 1) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        const int jw = jwt % 10;
        if (alt) {
            /* ALT (even and odd rows two interleaved systems, pfSolve_Plan_JonesWorland.h:265,
             * pfSolve_TridiagonalSolve.h:1622-1646): the JW modes in the parallel layout with baked
             * matrix positions, one iteration; each parity's system on the general kernel */
            if (jw == 4 || out->jw_type > 1000 || out->numConsecutiveJWIterations != 1 ||
                (out->jw_control_bitmask & (RUNTIME_OFFSETM | RUNTIME_OFFSETV)) || out->inputBlock ||
                out->outputBlock || out->kernelVariant == 2)
                return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
            if ((N + 1) / 2 > (prec == JW_PREC_F64 ? 1024 : JW_MAX_N_DD)) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        }
        const int K = out->numConsecutiveJWIterations;
        if (out->upperBanded < 0 || out->upperBanded > 2) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (out->useMultipleInputBuffers != 0 && out->useMultipleInputBuffers != K) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (K != 1) {
            /* chained steps (pfSolve_JonesWorland.h:108-374): the bidiagonal JW modes, FP64,
             * every iteration's system (N -/+ i rows) on one wavefront of the chained kernel */
            if (K < 1 || K > JW_CHAIN_MAX || jw > 2 || out->upperBanded > 1) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
            if (prec != JW_PREC_F64 || out->inputBlock || out->outputBlock || out->kernelVariant == 1)
                return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
            const int64_t nmax = out->upperBanded ? N : N + K - 1;
            const int64_t nmin = out->upperBanded ? N - K + 1 : N;
            if (nmin < 1) return PFSOLVE_ERROR_EMPTY_size;
            if (nmax > JW_MAX_N_CHAIN) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        }
        if (jw != 4 && out->upperBanded == 2 && jw != 3) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (jw == 4) {
            if (out->KL == 0 && out->KU == 0) { out->KL = 2; out->KU = 2; }
            if (out->KL < 0 || out->KU < 0 || out->KL > 2 || out->KU > 2) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        }
        /* beyond one wavefront's rows the step is cut into chunks (the reference's
         * multi-upload path for M_size > 3840, pfSolve_InitializeApp.h:1540-1580) */
        if (!alt && N > JW_MAX_N_LONG) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        if (!alt && jw_is_chunked(N, prec) && out->kernelVariant == 2) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        /* block operators fused around the step: layout 4 on the input (complex in, no
         * accumulate), and on the output layout 7 (complex out, optional accumulate)
         * or none (the step's real systems) */
        if (out->inputBlock || out->outputBlock) {
            const int64_t ib = out->inputBlock, ob = out->outputBlock;
            if (ib <= 0 || ob < 0 || ib % 10 != 4 || (ib / 10) % 10 != 0 || ib / 100 > 4 ||
                (ob && (ob % 10 != 7 || ((ob / 10) % 10 != 0 && (ob / 10) % 10 != 2) || ob / 100 > 4)))
                return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
            if (prec != JW_PREC_F64 || N > JW_MAX_N_FUSED_BLOCKS) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        }
    } else if (out->inputBlock || out->outputBlock) {
        return PFSOLVE_ERROR_UNSUPPORTED_RADIX;   /* only a JW step takes fused blocks */
    } else if (kind == PF_KIND_BLOCK) {
        const int layout = out->block % 10;
        const int op = out->block / 100;
        const int acc = (out->block / 10) % 10;
        if (layout < 1 || layout > 7 || op > 4 || (acc != 0 && acc != 2)) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (prec != JW_PREC_F64) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
    } else {
        if (out->ALT_dgbmv) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (out->LDA != out->KL + out->KU + 1 || out->KL < 0 || out->KU < 0) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (prec != JW_PREC_F64) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
        if (N > 16384) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    }
    return PFSOLVE_SUCCESS;
}

/* Rows per lane R and lanes per system L: the smallest R that fits N on one
 * wavefront, sub-wavefront groups for short systems. */
PfSolveResult jw_select_geometry(int64_t N, int precision, int* L, int* R)
{
    if (N <= 0) return PFSOLVE_ERROR_EMPTY_size;
    if (N <= 64) {
        /* one lane per row keeps the whole batch spread over many wavefronts */
        *L = 64;
        *R = 1;
        return PFSOLVE_SUCCESS;
    }
    const int64_t maxN = precision == JW_PREC_F64 ? JW_MAX_N_F64 : JW_MAX_N_DD;
    if (N > maxN) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    *L = 64;
    *R = (int)((N + 63) / 64);
    if (*R > 16) *R = 32;   /* the general kernel stops at 16 rows per lane */
    return PFSOLVE_SUCCESS;
}

void jw_keep_range(int64_t zp0, int64_t zp1, int runtime, int64_t N, int64_t* lo, int64_t* hi)
{
    /* PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel0/pfSolve_Zeropad.h:306-378: a baked upper
     * bound of 0 means "no upper bound"; runtime bounds are taken literally. */
    int64_t a = zp0, b = zp1;
    if (!runtime && b == 0) b = N;
    if (a < 0) a = 0;
    if (a > N) a = N;
    if (b > N) b = N;
    if (b < a) b = a;
    *lo = a;
    *hi = b;
}

/* Block stride of one diagonal in the ALT storage: the reference's get_ALT_stride
 * (PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:30-64, HIP branch):
 * rows per lane (odd) times lanes of the layout of N rows. */
int64_t jw_alt_stride(int64_t N, uint64_t Msplit0)
{
    int64_t size1 = N > 3840 ? (int64_t)(Msplit0 ? Msplit0 : 704) : N;
    int64_t W = ((size1 + 1023) / 1024) * 64;
    if (size1 < 512) W = 64;
    int64_t R = (size1 + W - 1) / W;
    if (R % 2 == 0) R++;
    return R * W;
}

static int64_t eff(int64_t mask, int64_t bit, int64_t cfgv, uint64_t runv)
{
    return (mask & bit) ? (int64_t)runv : cfgv;
}

static const PfSolveLaunchParams g_zero_lp;

PfSolveResult jw_resolve_launch(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                                int64_t nsys, jw_source_desc* src, jw_run_desc* run)
{
    if (!lp) lp = &g_zero_lp;
    memset(src, 0, sizeof(*src));
    memset(run, 0, sizeof(*run));
    const int64_t N = cfg->M_size;
    const int64_t m = cfg->jw_control_bitmask;
    const int rtM = (m & RUNTIME_OFFSETM) != 0, rtV = (m & RUNTIME_OFFSETV) != 0;
    const int64_t offM = eff(m, RUNTIME_OFFSETM, cfg->offsetM, lp->offsetM);
    const int64_t offV = eff(m, RUNTIME_OFFSETV, cfg->offsetV, lp->offsetV);
    const int64_t offT = eff(m, RUNTIME_OFFSETT, cfg->offsetT, lp->offsetT);
    const int alt = (cfg->jw_type % 1000) / 100 == 1;
    const int jwt = cfg->jw_type % 100;
    const int jw = jwt % 10;
    const int ub = cfg->upperBanded;
    /* baked positions 0, N, 2N; in the ALT variant multiples of its block stride
     * (pfSolve_JonesWorland.h:205-222) */
    const int64_t NB = alt ? jw_alt_stride(N, cfg->Msplit[0]) : N;

    src->N = N;
    src->alt = alt;
    src->mdS = -1;
    src->Msplit0 = (int32_t)cfg->Msplit[0];
    src->permuted = cfg->jw_type < 1000 && !(N <= 64 && cfg->numConsecutiveJWIterations == 1);
    src->mem_f64 = jw_precision(cfg) == JW_PREC_DD_F64MEM;

    /* matrix offsets per mode: PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:117-348 */
    if (jw == 4) {
        src->band_kind = JW_BAND_GB;
        src->offV = offV;
        src->KL = cfg->KL;
        src->KU = cfg->KU;
        src->solve_kind = JW_SOLVE_TRI;
        src->ldS = offT;
        src->mdS = offT + N;
        src->udS = offT + 2 * N;
        src->has_m = 1;
        src->mdM = offM;
        src->udM = offM + N;
    } else if (ub == 2) {
        src->band_kind = JW_BAND_DIAG;
        src->ldS = 0;
        src->mdA = NB;
        src->udS = 2 * NB;
        src->solve_kind = JW_SOLVE_TRI;
    } else {
        if (jw == 0 || jw == 1) {
            if (ub) {
                src->band_kind = JW_BAND_UPPER;
                src->udA = rtM ? offM : NB;
                src->mdA = rtM ? offM + N : 0;
            } else {
                src->band_kind = JW_BAND_LOWER;
                src->mdA = rtM ? offM : NB;
                src->ldA = rtM ? offM + N : 0;
            }
        }
        if (jw == 0 || jw == 2) {
            if (ub) { src->solve_kind = JW_SOLVE_LOWER; src->ldS = rtV ? offV : 2 * NB; }
            else    { src->solve_kind = JW_SOLVE_UPPER; src->udS = rtV ? offV : 2 * NB; }
        }
        if (jw == 3) {
            /* upperBanded toggled for the load (pfSolve_JonesWorland.h:124-127,202-205) */
            src->band_kind = JW_BAND_DIAG;
            if (ub) { src->solve_kind = JW_SOLVE_LOWER; src->mdA = rtM ? offM : NB; src->ldS = rtM ? offM + N : 0; }
            else    { src->solve_kind = JW_SOLVE_UPPER; src->udS = rtM ? offM : NB; src->mdA = rtM ? offM + N : 0; }
        }
    }

    run->nsys = nsys;
    run->N = (int32_t)N;
    run->stride = eff(m, RUNTIME_OUTPUTBUFFERSTRIDE, (int64_t)cfg->outputBufferStride[0], lp->outputBufferStride);
    run->offSol = eff(m, RUNTIME_OFFSETSOLUTION, cfg->offsetSolution, lp->offsetSolution);
    const int izrt = (m & RUNTIME_INPUTZEROPAD) != 0, ozrt = (m & RUNTIME_OUTPUTZEROPAD) != 0;
    jw_keep_range(izrt ? (int64_t)lp->inputZeropad[0] : cfg->inputZeropad[0],
                  izrt ? (int64_t)lp->inputZeropad[1] : cfg->inputZeropad[1], izrt, N, &run->in_lo, &run->in_hi);
    jw_keep_range(ozrt ? (int64_t)lp->outputZeropad[0] : cfg->outputZeropad[0],
                  ozrt ? (int64_t)lp->outputZeropad[1] : cfg->outputZeropad[1], ozrt, N, &run->out_lo, &run->out_hi);

    run->band = src->band_kind != JW_BAND_NONE;
    switch (src->band_kind) {
    case JW_BAND_UPPER: run->ku = 1; break;
    case JW_BAND_LOWER: run->kl = 1; break;
    case JW_BAND_GB: run->kl = src->KL; run->ku = src->KU; break;
    default: break;
    }
    run->fwd = src->solve_kind == JW_SOLVE_LOWER || src->solve_kind == JW_SOLVE_TRI;
    run->b1 = src->solve_kind == JW_SOLVE_UPPER || src->solve_kind == JW_SOLVE_TRI;
    run->b2 = src->has_m;
    /* row-0 scaleC of the solve (pfSolve_TridiagonalSolve.h:1918-1927): applied when the
     * value is runtime or differs from 1; the DD value is {hi, 0} (pfSolve_DispatchPlan.h:196-217) */
    const int rtS = (m & RUNTIME_SCALEC) != 0;
    const double sc = rtS ? (double)lp->scaleC : cfg->scaleC;
    run->apply_scale = (src->solve_kind != JW_SOLVE_NONE) && (rtS || sc != 1.0);
    run->scale_hi = sc;
    run->scale_lo = 0.0;
    run->mem_f64 = src->mem_f64;
    return PFSOLVE_SUCCESS;
}

/* Iteration i of a chained step (numConsecutiveJWIterations = K > 1), for the bidiagonal
 * modes jw 10/11/12 with upperBanded 0/1
 * (PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:108-330):
 *  - N_i = N - i (upperBanded 1) or N + i (upperBanded 0), :110-113; the matvec of
 *    iteration i > 0 has the previous iteration's N_{i-1} columns;
 *  - the matvec diagonal is read and applied in the first (upperBanded 1) or the last
 *    (upperBanded 0) iteration only, an implicit 1 elsewhere (pfSolve_MatVecMul.h:74,98,
 *    pfSolve_ReadWrite.h:1650);
 *  - baked positions: upper [D @0, U @N_i] in iteration 0, then [U @0]; lower [L @0,
 *    D @N_i] in the last iteration, [L @0] before (:159-168, 189-198); the solve at 2 N_i in
 *    the first (lower solve) / last (upper solve) iteration, at N_i otherwise (:291-296,
 *    320-324); runtime offsets are the same every iteration, with N_i for the diagonal;
 *  - storage in reorder_i0 order of N_i rows in the register layout of the largest system
 *    (pfSolve_Plan_JonesWorland.h:277-282). */
PfSolveResult jw_resolve_chain_iteration(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp, int i,
                                         jw_source_desc* src)
{
    if (!lp) lp = &g_zero_lp;
    memset(src, 0, sizeof(*src));
    const int K = cfg->numConsecutiveJWIterations;
    const int64_t N0 = cfg->M_size;
    const int ub = cfg->upperBanded;
    const int jw = (cfg->jw_type % 1000) % 10;
    const int64_t m = cfg->jw_control_bitmask;
    const int rtM = (m & RUNTIME_OFFSETM) != 0, rtV = (m & RUNTIME_OFFSETV) != 0;
    const int64_t offM = eff(m, RUNTIME_OFFSETM, cfg->offsetM, lp->offsetM);
    const int64_t offV = eff(m, RUNTIME_OFFSETV, cfg->offsetV, lp->offsetV);
    const int64_t Ni = ub ? N0 - i : N0 + i;
    src->N = Ni;
    src->Ncols = i == 0 ? N0 : (ub ? Ni + 1 : Ni - 1);
    src->Nalign = ub ? N0 : N0 + K - 1;
    src->mdS = -1;
    src->Msplit0 = (int32_t)cfg->Msplit[0];
    src->permuted = cfg->jw_type < 1000;   /* K > 1: the parallel layout at every size (Plan_JonesWorland.h:305-306) */
    if (jw == 0 || jw == 1) {
        const int diag_here = ub ? i == 0 : i == K - 1;
        src->unit_diag = !diag_here;
        if (ub) {
            src->band_kind = JW_BAND_UPPER;
            src->udA = rtM ? offM : (i == 0 ? Ni : 0);
            src->mdA = rtM ? offM + Ni : 0;
        } else {
            src->band_kind = JW_BAND_LOWER;
            src->ldA = rtM ? offM + Ni : 0;
            src->mdA = rtM ? offM : Ni;
        }
    }
    if (jw == 0 || jw == 2) {
        if (ub) { src->solve_kind = JW_SOLVE_LOWER; src->ldS = rtV ? offV : (i == 0 ? 2 * Ni : Ni); }
        else    { src->solve_kind = JW_SOLVE_UPPER; src->udS = rtV ? offV : (i == K - 1 ? 2 * Ni : Ni); }
    }
    return PFSOLVE_SUCCESS;
}

/* Largest element index the prologue reads from the matrix buffer and the main
 * kernel touches in the RHS buffer (-1: none). */
void jw_index_extent(const jw_source_desc* s, const jw_run_desc* r, int64_t* mat_max, int64_t* rhs_max)
{
    const int64_t N = s->N;
    int64_t m = -1;
#define UPD(base) do { if ((base) + N - 1 > m) m = (base) + N - 1; } while (0)
    switch (s->band_kind) {
    case JW_BAND_DIAG: UPD(s->mdA); break;
    case JW_BAND_UPPER: if (!s->unit_diag) UPD(s->mdA); UPD(s->udA); break;
    case JW_BAND_LOWER: if (!s->unit_diag) UPD(s->mdA); UPD(s->ldA); break;
    case JW_BAND_GB: UPD(s->offV + (int64_t)(s->KL + s->KU) * N); break;
    default: break;
    }
    if (s->solve_kind == JW_SOLVE_LOWER || s->solve_kind == JW_SOLVE_TRI) UPD(s->ldS);
    if (s->solve_kind == JW_SOLVE_UPPER || s->solve_kind == JW_SOLVE_TRI) UPD(s->udS);
    if (s->mdS >= 0) UPD(s->mdS);
    if (s->has_m) { UPD(s->mdM); UPD(s->udM); }
#undef UPD
    *mat_max = m;
    int64_t hi = r->in_hi > r->out_hi ? r->in_hi : r->out_hi;
    *rhs_max = (r->nsys > 0 && hi > 0) ? (r->nsys - 1) * r->stride + r->offSol + hi - 1 : -1;
}

PfSolveResult jw_check_offsets(const jw_source_desc* s, const jw_run_desc* r)
{
    /* matrix positions: every diagonal the step reads starts inside the matrix buffer */
    if (s->mdA < 0 || s->udA < 0 || s->ldA < 0 || s->offV < 0 || s->ldS < 0 || s->udS < 0 ||
        s->mdS < -1 || s->mdM < 0 || s->udM < 0)
        return PFSOLVE_ERROR_EMPTY_bufferSize;
    if (r->stride < 0 || r->offSol < 0) return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    /* systems must not overlap when solved in place */
    if (r->nsys > 1 && r->stride < (r->in_hi > r->out_hi ? r->in_hi : r->out_hi))
        return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    return PFSOLVE_SUCCESS;
}

PfSolveResult jw_resolve_block(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                               int64_t nsys, jw_block_desc* d)
{
    if (!lp) lp = &g_zero_lp;
    memset(d, 0, sizeof(*d));
    const int64_t m = cfg->jw_control_bitmask;
    const int64_t N = (int64_t)cfg->size[0];
    d->N = N;
    d->B = nsys;
    d->layout = cfg->block % 10;
    d->accumulate = ((cfg->block / 10) % 10) == 2;
    d->op = cfg->block / 100;
    d->lshift = cfg->lshift;
    d->inStride = eff(m, RUNTIME_INPUTBUFFERSTRIDE, (int64_t)cfg->inputBufferStride[0], lp->inputBufferStride);
    d->outStride = eff(m, RUNTIME_OUTPUTBUFFERSTRIDE, (int64_t)cfg->outputBufferStride[0], lp->outputBufferStride);
    d->offIn = eff(m, RUNTIME_OFFSETM, cfg->offsetM, lp->offsetM);
    d->offOut = eff(m, RUNTIME_OFFSETSOLUTION, cfg->offsetSolution, lp->offsetSolution);
    const int izrt = (m & RUNTIME_INPUTZEROPAD) != 0, ozrt = (m & RUNTIME_OUTPUTZEROPAD) != 0;
    jw_keep_range(izrt ? (int64_t)lp->inputZeropad[0] : cfg->inputZeropad[0],
                  izrt ? (int64_t)lp->inputZeropad[1] : cfg->inputZeropad[1], izrt, N, &d->in_lo, &d->in_hi);
    jw_keep_range(ozrt ? (int64_t)lp->outputZeropad[0] : cfg->outputZeropad[0],
                  ozrt ? (int64_t)lp->outputZeropad[1] : cfg->outputZeropad[1], ozrt, N, &d->out_lo, &d->out_hi);
    const int rtS = (m & RUNTIME_SCALEC) != 0;
    d->scale = rtS ? (double)lp->scaleC : cfg->scaleC;
    /* baked ScaleC of 0 writes zeros without reading (pfSolve_block.h:82) */
    d->skip_read = (d->op == 1 && !rtS && d->scale == 0.0);
    return PFSOLVE_SUCCESS;
}

PfSolveResult jw_resolve_gbmv(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                              int64_t nsys, jw_gbmv_desc* d)
{
    if (!lp) lp = &g_zero_lp;
    memset(d, 0, sizeof(*d));
    const int64_t m = cfg->jw_control_bitmask;
    const int64_t N = cfg->M_size;
    d->N = N;
    d->nsys = nsys;
    d->LDA = cfg->LDA;
    d->KU = cfg->KU;
    d->KL = cfg->KL;
    d->offM = eff(m, RUNTIME_OFFSETM, cfg->offsetM, lp->offsetM);
    d->offSol = eff(m, RUNTIME_OFFSETSOLUTION, cfg->offsetSolution, lp->offsetSolution);
    d->stride = eff(m, RUNTIME_OUTPUTBUFFERSTRIDE, (int64_t)cfg->outputBufferStride[0], lp->outputBufferStride);
    const int izrt = (m & RUNTIME_INPUTZEROPAD) != 0, ozrt = (m & RUNTIME_OUTPUTZEROPAD) != 0;
    jw_keep_range(izrt ? (int64_t)lp->inputZeropad[0] : cfg->inputZeropad[0],
                  izrt ? (int64_t)lp->inputZeropad[1] : cfg->inputZeropad[1], izrt, N, &d->in_lo, &d->in_hi);
    jw_keep_range(ozrt ? (int64_t)lp->outputZeropad[0] : cfg->outputZeropad[0],
                  ozrt ? (int64_t)lp->outputZeropad[1] : cfg->outputZeropad[1], ozrt, N, &d->out_lo, &d->out_hi);
    return PFSOLVE_SUCCESS;
}
