// jw_main64.hip -- wave-per-system FP64 JW kernels for gfx950 (table format 2).
//
// One wavefront owns a pair of systems (the re/im halves of a complex system
// in XXII layout, or two consecutive real systems); lane t owns rows
// t*R .. t*R+R-1 of both.  Per pair:
//   1. input: the 16-byte chunks (row pairs) inside the input keep-range; the
//      compute / store split kernel loads a lane's own chunks straight into
//      registers one pass ahead, the pair and batch kernels DMA them global ->
//      LDS (global_load_lds_dwordx4, 64 lanes per instruction on one contiguous
//      KiB, XOR-swizzled slots so the blocked read-back is bank-conflict free);
//      an odd range end is read by its owning lane alone;
//   2. banded matvec u' = a.(V x) with neighbour rows by DPP wave shifts;
//   3. three first-order recurrences (forward W, backward Z, backward y),
//      each: lane-local sweep, affine scan across the wavefront, lane-local
//      redo.  Scan: DPP row_shr/row_shl 1,2,4,8 inside 16-lane rows; across
//      rows DPP row_bcast15/row_bcast31 (forward) or readlane of the row
//      starts (backward);
//   4. row-0 scaleC; output written blocked to LDS, read back contiguous and
//      stored with 16-byte nontemporal stores (rows outside the output
//      keep-range untouched; an odd range end is written by its owning lane).
// Coefficients (8 per row, scalings folded) and scan multipliers come from the
// per-matrix-set table (format 2): built by each workgroup of the split kernel
// in its LDS, written by jw_prologue_kernel / jw_prologue_batch_kernel for the
// other kernels and staged once per workgroup (or per work unit) in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "jw_common.h"
#include "jw_prologue.h"

extern "C" void jwdev_set_error(const char* msg);

namespace {

constexpr int kLdsBytesMax = 160 * 1024;

template <int H>
__device__ __forceinline__ int swz(int c)
{
    return c ^ ((c >> 4) & (H - 1));
}

__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }
__host__ __device__ __forceinline__ int64_t max_i64(int64_t a, int64_t b) { return a > b ? a : b; }

// store of a result chunk: written once and not read back by the step, so it is
// marked nontemporal (streamed to memory without allocating in the caches)
__device__ __forceinline__ void store_stream(double2* p, double2 v)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<d2v*>(p));
}

// keeps the scheduler from hoisting one phase's table loads into the previous
// phase, which would raise register pressure past the launch bound
__device__ __forceinline__ void phase_barrier() { __builtin_amdgcn_sched_barrier(0); }

// DPP move of a double with an explicit row mask; rows outside the mask and
// lanes without a source receive 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_rm(double v)
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

// register budget: 512 / waves_per_eu VGPRs per lane (128 for R <= 8, 256 for R = 16)
template <int R>
constexpr int waves_per_eu()
{
    return R <= 8 ? 4 : 2;
}

template <int R>
constexpr int max_threads()
{
    return R <= 4 ? 1024 : (R == 8 ? 768 : 320);
}

// value of row r (which lane `lane` owns when r / R == lane) from the lane's registers
template <int R>
__device__ __forceinline__ double pick_row(const double (&x)[R], int i)
{
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < R; k++) v = (k == i) ? x[k] : v;
    return v;
}

// Everything one pass over a system pair needs besides the table.
struct jw_pass64 {
    double* rhs;
    int64_t nsys, stride, offSol;
    int in_lo, in_hi, out_lo, out_hi;     // keep-ranges, clipped to [0, N)
    int icl, ich, ocl, och;               // their full 16-byte chunks
    bool in_edge_lo, in_edge_hi, out_edge_lo, out_edge_hi;   // odd range ends
    int kl, ku;
    bool fwd, b1, b2, apply_scale;
    double row0_scale;                    // scaleC on lane 0, 1 elsewhere
};

__device__ __forceinline__ jw_pass64 make_pass64(double* rhs, int64_t nsys, int64_t stride, int64_t offSol,
                                                 int in_lo, int in_hi, int out_lo, int out_hi, int kl, int ku,
                                                 bool fwd, bool b1, bool b2, bool apply_scale, double scale,
                                                 int lane)
{
    jw_pass64 g;
    g.rhs = rhs;
    g.nsys = nsys;
    g.stride = stride;
    g.offSol = offSol;
    g.in_lo = in_lo;
    g.in_hi = in_hi;
    g.out_lo = out_lo;
    g.out_hi = out_hi;
    g.icl = (in_lo + 1) >> 1;
    g.ich = in_hi >> 1;
    g.ocl = (out_lo + 1) >> 1;
    g.och = out_hi >> 1;
    g.in_edge_lo = (in_lo & 1) && in_lo < in_hi;
    g.in_edge_hi = (in_hi & 1) && in_hi - 1 >= in_lo;   // never the row of in_edge_lo: lo odd => lo+1 even
    g.out_edge_lo = (out_lo & 1) && out_lo < out_hi;
    g.out_edge_hi = (out_hi & 1) && out_hi - 1 >= out_lo;
    g.kl = kl;
    g.ku = ku;
    g.fwd = fwd;
    g.b1 = b1;
    g.b2 = b2;
    g.apply_scale = apply_scale;
    g.row0_scale = lane == 0 ? scale : 1.0;
    return g;
}

template <int R>
__device__ __forceinline__ void zero_staging(double2* stg, int lane)
{
#pragma unroll
    for (int k = 0; k < R; k++) stg[64 * k + lane] = make_double2(0.0, 0.0);
}

// Leaving a work queue.  A queue is a counter pair {next unit, workgroups that have left};
// the last workgroup to leave puts both back to 0, so the next launch on the stream finds
// an empty queue with no clearing kernel or memset before it.  Every thread calls this
// after its workgroup's last claim; thread 0 counts.  The acquire-release increment orders
// every other workgroup's claims before the reset.
__device__ __forceinline__ void queue_leave(int* counter)
{
    if (threadIdx.x == 0) {
        const int n = __hip_atomic_fetch_add(counter + 1, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (n == (int)gridDim.x - 1) {
            __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(counter + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// The pieces of one pass over systems sys0 and sys0+1 (if it exists).  stg: a
// wavefront's staging for 2 systems (64 R double2); the input slots outside the
// keep-range hold zeros whenever the input DMA is issued.

// issue the input DMA (global -> LDS) of the pair; completion is vmcnt(0)
template <int R>
__device__ __forceinline__ void dma_pair(double2* stg, int lane, const jw_pass64& g, int64_t sys0)
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
    const bool ok1 = sys0 + 1 < g.nsys;
    const double* src0 = g.rhs + sys0 * g.stride + g.offSol;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s == 1 && !ok1) continue;
        const double* src = src0 + s * g.stride;
#pragma unroll
        for (int k = 0; k < H; k++) {
            const int c = swz<H>(64 * k + lane);
            if (c >= g.icl && c < g.ich)
                __builtin_amdgcn_global_load_lds(
                    src + 2 * c, (__attribute__((address_space(3))) void*)(stg + s * SYSC + 64 * k), 16, 0, 0);
        }
    }
}

template <int R>
__device__ __forceinline__ void read_edges(int lane, const jw_pass64& g, int64_t sys0, double (&x0)[R], double (&x1)[R]);

// blocked read-back of the landed input (lane t: rows t R .. t R + R-1), odd range
// ends read from global memory by their owning lane
template <int R>
__device__ __forceinline__ void read_pair(const double2* stg, int lane, const jw_pass64& g, int64_t sys0,
                                          double (&x0)[R], double (&x1)[R])
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
#pragma unroll
    for (int m = 0; m < H; m++) {
        const double2 v0 = stg[swz<H>(lane * H + m)];
        const double2 v1 = stg[SYSC + swz<H>(lane * H + m)];
        x0[2 * m] = v0.x;
        x0[2 * m + 1] = v0.y;
        x1[2 * m] = v1.x;
        x1[2 * m + 1] = v1.y;
    }
    read_edges<R>(lane, g, sys0, x0, x1);
}

// the odd ends of the input keep-range (rows whose 16-byte chunk is only half
// inside), read from global memory by their owning lane
template <int R>
__device__ __forceinline__ void read_edges(int lane, const jw_pass64& g, int64_t sys0, double (&x0)[R], double (&x1)[R])
{
    const bool ok1 = sys0 + 1 < g.nsys;
    const double* src0 = g.rhs + sys0 * g.stride + g.offSol;
    if (g.in_edge_lo || g.in_edge_hi) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const bool on = e == 0 ? g.in_edge_lo : g.in_edge_hi;
            const int r = e == 0 ? g.in_lo : g.in_hi - 1;
            if (on && lane == r / R) {
                const double v0 = src0[r];
                const double v1 = ok1 ? src0[g.stride + r] : 0.0;
#pragma unroll
                for (int k = 0; k < R; k++)
                    if (k == r % R) {
                        x0[k] = v0;
                        x1[k] = v1;
                    }
            }
        }
    }

}

// banded matvec, the three recurrences and row-0 scaleC, in registers.
// ALL3: all three recurrences present (the fused step), decided at compile time.
// G > 1 (R = 2 only): sub-wavefront groups.  The 64 lanes form G groups of LG = 64 / G
// lanes, each solving its own pair of systems of N <= 2 LG rows; lane t is lane t mod LG of
// its group for the table (the set's 64-lane table, whose lanes from N / 2 on are zero
// rows).  What crosses lanes is cut at the group boundaries: the matvec's neighbour rows
// and the recurrences' carries are 0 at a group's ends, and the scans' cross-row stages
// run inside a group only (G = 2: the row_bcast15 stage and the readlane of the group's
// second row start; G = 4: none).  The stages left out are the ones that, for a single
// system of N <= 2 LG rows, only move its zero rows, so a group computes what the pair
// kernel computes for the same pair.
// What a chunk of a long system (64 R of its rows, a pass of its own) takes from outside
// and hands out, per system of the pair.  In: lane 0 holds the two rows before the chunk
// (h1 = row -1, h2 = row -2) and the forward carry cw (W' of row -1); lane 63 holds the two
// rows after it (h1 = row C, h2 = row C + 1) and the backward carries cz, cy (Z' and y of
// row C); every other lane holds zeros.  Out: wl = W' of the lane's last row (the chunk's
// last row on lane 63), zf = Z' of the lane's first row (the chunk's first row on lane 0).
struct jw_edge64 {
    double h1[2], h2[2], cw[2], cz[2], cy[2];
    double wl[2], zf[2];
};

template <int R, bool ALL3, int G = 1, bool EDGE = false>
__device__ __forceinline__ void compute_pair(const double* tab, int lane, const jw_pass64& g, double (&x0)[R],
                                             double (&x1)[R], jw_edge64* e = nullptr)
{
    static_assert(!EDGE || G == 1, "chunk edges: whole wavefronts only");
    static_assert(G == 1 || (R == 2 && (G == 2 || G == 4)), "sub-wavefront groups: R = 2, G = 2 or 4");
    constexpr int SC = 8 * R * 64;            // start of the scan multipliers
    constexpr int LG = 64 / G;
    const int tl = G == 1 ? lane : (lane & (LG - 1));   // lane within the group: the table's lane
    const bool gfirst = G > 1 && tl == 0, glast = G > 1 && tl == LG - 1;
    // ---- 2. banded matvec (coefficients carry the folded row scaling) ----
    {
        double m1a = 0.0, m2a = 0.0, p1a = 0.0, p2a = 0.0;
        double m1b = 0.0, m2b = 0.0, p1b = 0.0, p2b = 0.0;
        if (g.kl >= 1) { m1a = dpp_d<DPP_WAVE_SHR1>(x0[R - 1]); m1b = dpp_d<DPP_WAVE_SHR1>(x1[R - 1]); }
        if (g.kl >= 2) { m2a = dpp_d<DPP_WAVE_SHR1>(x0[R - 2]); m2b = dpp_d<DPP_WAVE_SHR1>(x1[R - 2]); }
        if (g.ku >= 1) { p1a = dpp_d<DPP_WAVE_SHL1>(x0[0]); p1b = dpp_d<DPP_WAVE_SHL1>(x1[0]); }
        if (g.ku >= 2) { p2a = dpp_d<DPP_WAVE_SHL1>(x0[1]); p2b = dpp_d<DPP_WAVE_SHL1>(x1[1]); }
        if (gfirst) { m1a = 0.0; m1b = 0.0; m2a = 0.0; m2b = 0.0; }
        if (glast) { p1a = 0.0; p1b = 0.0; p2a = 0.0; p2b = 0.0; }
        if constexpr (EDGE) {
            // the rows next to the chunk enter where the wave shifts bring zeros
            if (lane == 0) {
                if (g.kl >= 1) { m1a = e->h1[0]; m1b = e->h1[1]; }
                if (g.kl >= 2) { m2a = e->h2[0]; m2b = e->h2[1]; }
            }
            if (lane == 63) {
                if (g.ku >= 1) { p1a = e->h1[0]; p1b = e->h1[1]; }
                if (g.ku >= 2) { p2a = e->h2[0]; p2b = e->h2[1]; }
            }
        }
#pragma unroll
        for (int i = 0; i < R; i++) {
            const double2 c0 = *reinterpret_cast<const double2*>(tab + tab_coef64(0, i, tl, R));
            const double2 c1 = *reinterpret_cast<const double2*>(tab + tab_coef64(1, i, tl, R));
            const double vp2 = tab[tab_coef64(2, i, tl, R)];
            const double ia = x0[i], ib = x1[i];
            const double n1a = (i + 1 < R) ? x0[(i + 1) % R] : p1a;
            const double n1b = (i + 1 < R) ? x1[(i + 1) % R] : p1b;
            const double n2a = (i + 2 < R) ? x0[(i + 2) % R] : ((i + 2 == R) ? p1a : p2a);
            const double n2b = (i + 2 < R) ? x1[(i + 2) % R] : ((i + 2 == R) ? p1b : p2b);
            double ua = c1.x * ia, ub = c1.x * ib;
            ua = __builtin_fma(c0.y, m1a, ua);
            ub = __builtin_fma(c0.y, m1b, ub);
            ua = __builtin_fma(c0.x, m2a, ua);
            ub = __builtin_fma(c0.x, m2b, ub);
            ua = __builtin_fma(c1.y, n1a, ua);
            ub = __builtin_fma(c1.y, n1b, ub);
            ua = __builtin_fma(vp2, n2a, ua);
            ub = __builtin_fma(vp2, n2b, ub);
            m2a = m1a;
            m2b = m1b;
            m1a = ia;
            m1b = ib;
            x0[i] = ua;
            x1[i] = ub;
            if (R >= 16 && (i & 3) == 3) phase_barrier();
        }
    }
    if (R >= 16) phase_barrier();

    // ---- 3a. forward recurrence W_r = u'_r - l'_r W_{r-1} ----
    if (ALL3 || g.fwd) {
        double lr[R];
#pragma unroll
        for (int i = 0; i < R; i++) lr[i] = tab[tab_coef64(2, i, tl, R) + 1];
        if constexpr (EDGE) {
            // W' entering the chunk: its first row becomes u'_0 - l'_0 cw
            if (lane == 0) {
                x0[0] = __builtin_fma(-lr[0], e->cw[0], x0[0]);
                x1[0] = __builtin_fma(-lr[0], e->cw[1], x1[0]);
            }
        }
        double a = x0[0], b = x1[0];
#pragma unroll
        for (int i = 1; i < R; i++) {
            a = __builtin_fma(-lr[i], a, x0[i]);
            b = __builtin_fma(-lr[i], b, x1[i]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double f = tab[SC + k * 64 + tl];
            double sa, sb;
            switch (k) {
            case 0: sa = dpp_d<DPP_ROW_SHR(1)>(a); sb = dpp_d<DPP_ROW_SHR(1)>(b); break;
            case 1: sa = dpp_d<DPP_ROW_SHR(2)>(a); sb = dpp_d<DPP_ROW_SHR(2)>(b); break;
            case 2: sa = dpp_d<DPP_ROW_SHR(4)>(a); sb = dpp_d<DPP_ROW_SHR(4)>(b); break;
            default: sa = dpp_d<DPP_ROW_SHR(8)>(a); sb = dpp_d<DPP_ROW_SHR(8)>(b); break;
            }
            a = __builtin_fma(f, sa, a);
            b = __builtin_fma(f, sb, b);
        }
        if (G <= 2) {
            // rows 1 and 3 take the ends of rows 0 and 2; rows 2 and 3 then take lane 31
            const double fp = tab[SC + 4 * 64 + tl];
            const double sa = dpp_rm<DPP_ROW_BCAST15, 0xa>(a), sb = dpp_rm<DPP_ROW_BCAST15, 0xa>(b);
            a = __builtin_fma(fp, sa, a);
            b = __builtin_fma(fp, sb, b);
            if (G == 1) {
                const double fp2 = tab[SC + 5 * 64 + tl];
                const double ta = dpp_rm<DPP_ROW_BCAST31, 0xc>(a), tb = dpp_rm<DPP_ROW_BCAST31, 0xc>(b);
                a = __builtin_fma(fp2, ta, a);
                b = __builtin_fma(fp2, tb, b);
            }
        }
        double wa = dpp_d<DPP_WAVE_SHR1>(a), wb = dpp_d<DPP_WAVE_SHR1>(b);
        if (gfirst) { wa = 0.0; wb = 0.0; }
#pragma unroll
        for (int i = 0; i < R; i++) {
            wa = __builtin_fma(-lr[i], wa, x0[i]);
            wb = __builtin_fma(-lr[i], wb, x1[i]);
            x0[i] = wa;
            x1[i] = wb;
        }
    }

    if (R >= 16) phase_barrier();
    // ---- 3b/3c. backward recurrences Z_r = W_r - h'_r Z_{r+1}, y_r = Z_r - E_r y_{r+1} ----
#pragma unroll
    for (int stage = 0; stage < 2; stage++) {
        if (R >= 16 && stage == 1) phase_barrier();
        if constexpr (EDGE) {
            if (stage == 0) { e->wl[0] = x0[R - 1]; e->wl[1] = x1[R - 1]; }
            if (stage == 1) { e->zf[0] = x0[0]; e->zf[1] = x1[0]; }
        }
        if (!ALL3 && stage == 0 && !g.b1) continue;
        if (!ALL3 && stage == 1 && !g.b2) continue;
        double hh[R];
#pragma unroll
        for (int i = 0; i < R; i++) hh[i] = tab[tab_coef64(3, i, tl, R) + stage];
        if constexpr (EDGE) {
            // Z' (stage 0) or y (stage 1) entering the chunk from the row after it
            if (lane == 63) {
                const double ca = stage == 0 ? e->cz[0] : e->cy[0], cb = stage == 0 ? e->cz[1] : e->cy[1];
                x0[R - 1] = __builtin_fma(-hh[R - 1], ca, x0[R - 1]);
                x1[R - 1] = __builtin_fma(-hh[R - 1], cb, x1[R - 1]);
            }
        }
        const int j = 1 + stage;
        double a = x0[R - 1], b = x1[R - 1];
#pragma unroll
        for (int i = R - 2; i >= 0; i--) {
            a = __builtin_fma(-hh[i], a, x0[i]);
            b = __builtin_fma(-hh[i], b, x1[i]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double f = tab[SC + (j * 8 + k) * 64 + tl];
            double sa, sb;
            switch (k) {
            case 0: sa = dpp_d<DPP_ROW_SHL(1)>(a); sb = dpp_d<DPP_ROW_SHL(1)>(b); break;
            case 1: sa = dpp_d<DPP_ROW_SHL(2)>(a); sb = dpp_d<DPP_ROW_SHL(2)>(b); break;
            case 2: sa = dpp_d<DPP_ROW_SHL(4)>(a); sb = dpp_d<DPP_ROW_SHL(4)>(b); break;
            default: sa = dpp_d<DPP_ROW_SHL(8)>(a); sb = dpp_d<DPP_ROW_SHL(8)>(b); break;
            }
            a = __builtin_fma(f, sa, a);
            b = __builtin_fma(f, sb, b);
        }
        if (G == 1) {
            // every row adds the in-row results of the later row starts (lanes 16, 32, 48),
            // weighted by matrix-only multipliers (zero for rows at or after that start)
            const double c1 = tab[SC + (j * 8 + 4) * 64 + lane];
            const double c2 = tab[SC + (j * 8 + 5) * 64 + lane];
            const double c3 = tab[SC + (j * 8 + 6) * 64 + lane];
            const double a1 = readlane_d(a, 16), a2 = readlane_d(a, 32), a3 = readlane_d(a, 48);
            const double b1 = readlane_d(b, 16), b2 = readlane_d(b, 32), b3 = readlane_d(b, 48);
            a = __builtin_fma(c1, a1, a);
            b = __builtin_fma(c1, b1, b);
            a = __builtin_fma(c2, a2, a);
            b = __builtin_fma(c2, b2, b);
            a = __builtin_fma(c3, a3, a);
            b = __builtin_fma(c3, b3, b);
        } else if (G == 2) {
            // the group's second row start: lane 16 of group 0, lane 48 of group 1
            const double c1 = tab[SC + (j * 8 + 4) * 64 + tl];
            const double a16 = readlane_d(a, 16), a48 = readlane_d(a, 48);
            const double b16 = readlane_d(b, 16), b48 = readlane_d(b, 48);
            const bool g0 = lane < 32;
            a = __builtin_fma(c1, g0 ? a16 : a48, a);
            b = __builtin_fma(c1, g0 ? b16 : b48, b);
        }
        double za = dpp_d<DPP_WAVE_SHL1>(a), zb = dpp_d<DPP_WAVE_SHL1>(b);
        if (glast) { za = 0.0; zb = 0.0; }
#pragma unroll
        for (int i = R - 1; i >= 0; i--) {
            za = __builtin_fma(-hh[i], za, x0[i]);
            zb = __builtin_fma(-hh[i], zb, x1[i]);
            x0[i] = za;
            x1[i] = zb;
        }
    }

    // ---- row-0 scaleC (pfSolve_TridiagonalSolve.h:1918-1927) ----
    if (g.apply_scale) {
        x0[0] *= g.row0_scale;
        x1[0] *= g.row0_scale;
    }

}

// results written blocked into the staging
template <int R>
__device__ __forceinline__ void write_blocked(double2* stg, int lane, const double (&x0)[R], const double (&x1)[R])
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
    compiler_fence();
#pragma unroll
    for (int m = 0; m < H; m++) {
        stg[swz<H>(lane * H + m)] = make_double2(x0[2 * m], x0[2 * m + 1]);
        stg[SYSC + swz<H>(lane * H + m)] = make_double2(x1[2 * m], x1[2 * m + 1]);
    }
}

// contiguous read of the staging and 16-byte stores of the output keep-range;
// odd range ends stored by lane 0
template <int R>
__device__ __forceinline__ void store_pair(const double2* stg, int lane, const jw_pass64& g, int64_t sys0)
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
    const bool ok1 = sys0 + 1 < g.nsys;
    compiler_fence();
    // all LDS reads first (always in bounds), so that they overlap, then the stores
    double2 v[2][H];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int k = 0; k < H; k++) v[s][k] = stg[s * SYSC + swz<H>(64 * k + lane)];
    // keeps the compiler from sinking each read into its store's branch
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int k = 0; k < H; k++) asm volatile("" : "+v"(v[s][k].x), "+v"(v[s][k].y));
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s == 1 && !ok1) continue;
        // the host only selects this kernel when sys*stride + offSol is even and rhs is 16-byte aligned
        double2* dst2 = reinterpret_cast<double2*>(g.rhs) + (((sys0 + s) * g.stride + g.offSol) >> 1);
#pragma unroll
        for (int k = 0; k < H; k++) {
            const int c = 64 * k + lane;
            if (c >= g.ocl && c < g.och) store_stream(dst2 + c, v[s][k]);
        }
    }
    if (g.out_edge_lo || g.out_edge_hi) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const bool on = e == 0 ? g.out_edge_lo : g.out_edge_hi;
            const int r = e == 0 ? g.out_lo : g.out_hi - 1;
            if (on && lane == 0) {
                double* dst = g.rhs + sys0 * g.stride + g.offSol + r;
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    if (s == 1 && !ok1) continue;
                    const double2 v = stg[s * SYSC + swz<H>(r >> 1)];
                    dst[s * g.stride] = (r & 1) ? v.y : v.x;
                }
            }
        }
    }
}

// zeros back into the slots the input DMA skips
template <int R>
__device__ __forceinline__ void rezero_input(double2* stg, int lane, const jw_pass64& g)
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
#pragma unroll
    for (int s = 0; s < 2; s++) {
#pragma unroll
        for (int k = 0; k < H; k++) {
            const int c = swz<H>(64 * k + lane);
            if (!(c >= g.icl && c < g.ich)) stg[s * SYSC + 64 * k + lane] = make_double2(0.0, 0.0);
        }
    }
}

// One whole pass in one wavefront (input, compute, output through one staging).
template <int R, bool ALL3>
__device__ __forceinline__ void jw_pair64(const double* tab, double2* stg, int lane, const jw_pass64& g,
                                          int64_t sys0)
{
    dma_pair<R>(stg, lane, g, sys0);
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the DMA writes have landed in LDS
    compiler_fence();
    double x0[R], x1[R];
    read_pair<R>(stg, lane, g, sys0, x0, x1);
    compute_pair<R, ALL3>(tab, lane, g, x0, x1);
    write_blocked<R>(stg, lane, x0, x1);
    store_pair<R>(stg, lane, g, sys0);
    rezero_input<R>(stg, lane, g);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): LDS traffic done before the next DMA lands
    compiler_fence();
}

// A set's table (NTAB doubles, whole 1 KiB blocks) from global memory into LDS at `smem`
// One pass of sub-wavefront groups (R = 2): the 2G consecutive systems sys0 .. sys0+2G-1
// (those below sys_end) in one wavefront, group gi solving systems sys0 + 2 gi and
// sys0 + 2 gi + 1.  Staging: system slot k (system sys0 + k) is chunks [k LG, (k+1) LG)
// of the wavefront's 128 chunks (chunk m = rows 2m, 2m+1).  Every chunk is written each
// pass (by the input DMA or as zeros), so the pass needs no zeroed staging.
template <int G>
__device__ __forceinline__ void jw_groups64(const double* tab, double2* stg, int lane, const jw_pass64& g,
                                            int64_t sys0, int64_t sys_end)
{
    constexpr int LG = 64 / G;
    const int tl = lane & (LG - 1), gi = lane / LG;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = 64 * j + lane;
        const int64_t sys = sys0 + c / LG;
        const int m = c & (LG - 1);
        if (sys < sys_end && m >= g.icl && m < g.ich)
            __builtin_amdgcn_global_load_lds(g.rhs + sys * g.stride + g.offSol + 2 * m,
                                             (__attribute__((address_space(3))) void*)(stg + 64 * j), 16, 0, 0);
        else
            stg[c] = make_double2(0.0, 0.0);
    }
    __builtin_amdgcn_s_waitcnt(0);   // the DMA and the zero writes have landed
    compiler_fence();
    const int64_t s0 = sys0 + 2 * gi, s1 = s0 + 1;
    double x0[2], x1[2];
    {
        const double2 v0 = stg[2 * gi * LG + tl], v1 = stg[(2 * gi + 1) * LG + tl];
        x0[0] = v0.x;
        x0[1] = v0.y;
        x1[0] = v1.x;
        x1[1] = v1.y;
    }
    // the odd ends of the input keep-range, by their owning lanes
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const bool on = e == 0 ? g.in_edge_lo : g.in_edge_hi;
        const int r = e == 0 ? g.in_lo : g.in_hi - 1;
        if (on && tl == r / 2) {
            const double v0 = s0 < sys_end ? g.rhs[s0 * g.stride + g.offSol + r] : 0.0;
            const double v1 = s1 < sys_end ? g.rhs[s1 * g.stride + g.offSol + r] : 0.0;
            if (r & 1) { x0[1] = v0; x1[1] = v1; } else { x0[0] = v0; x1[0] = v1; }
        }
    }
    compute_pair<2, true, G>(tab, lane, g, x0, x1);
    compiler_fence();
    stg[2 * gi * LG + tl] = make_double2(x0[0], x0[1]);
    stg[(2 * gi + 1) * LG + tl] = make_double2(x1[0], x1[1]);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): the results are in the staging
    compiler_fence();
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = 64 * j + lane;
        const int64_t sys = sys0 + c / LG;
        const int m = c & (LG - 1);
        const double2 v = stg[c];
        if (sys < sys_end && m >= g.ocl && m < g.och)
            store_stream(reinterpret_cast<double2*>(g.rhs + sys * g.stride + g.offSol) + m, v);
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const bool on = e == 0 ? g.out_edge_lo : g.out_edge_hi;
        const int r = e == 0 ? g.out_lo : g.out_hi - 1;
        if (on && tl == r / 2) {
            if (s0 < sys_end) g.rhs[s0 * g.stride + g.offSol + r] = (r & 1) ? x0[1] : x0[0];
            if (s1 < sys_end) g.rhs[s1 * g.stride + g.offSol + r] = (r & 1) ? x1[1] : x1[0];
        }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): the staging is read before the next pass writes it
    compiler_fence();
}

template <int NTAB>
__device__ __forceinline__ void copy_table_lds(unsigned char* smem, const void* table, int wave, int nwaves, int lane)
{
    static_assert(NTAB * 8 % 1024 == 0, "the table is whole 1 KiB blocks");
    copy_to_lds(smem, table, NTAB * 8, wave, nwaves, lane);
}

template <int R, bool ALL3>
__global__ __launch_bounds__(max_threads<R>()) void jw_main64_kernel(jw_run_desc d)
{
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    copy_table_lds<NTAB>(smem, d.table, wave, nwaves, lane);
    double2* stg = reinterpret_cast<double2*>(smem + NTAB * 8) + (size_t)wave * 64 * R;
    zero_staging<R>(stg, lane);
    __syncthreads();
    const int64_t npass = (d.nsys + 1) >> 1;
    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        // rebuilt every pass from the kernel arguments, so that nothing derived from
        // them stays live in registers across the loop
        const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol,
                                        (int)d.in_lo, (int)d.in_hi, (int)d.out_lo, (int)d.out_hi, d.kl, d.ku,
                                        d.fwd, d.b1, d.b2, d.apply_scale, d.scale_hi, lane);
        jw_pair64<R, ALL3>(tab, stg, lane, g, 2 * q);
    }
}

// ---------------------------------------------------------------------------
// Long FP64 systems on the pair machinery (16-byte aligned systems).  Chunk c of a long
// system (rows c C .. c C + C - 1, C = 512, R = 8) is one wavefront pass over a pair of
// systems like a whole 512-row system, with the chunk's format-2 table (the scalings
// continue across chunks) and the edges of jw_edge64.  The modes of the chunk
// decomposition (jw_long.hip: COEF, REDUCE, FINISH, the reduced solve between REDUCE and
// FINISH) run the same arithmetic, so the carries the reduced solve composes from COEF's
// responses are the ones FINISH applies.
// ---------------------------------------------------------------------------
constexpr int kLong64R = 8;

// reduced data of chunk c, system s: array a (the layout of jw_long_desc)
__device__ __forceinline__ int64_t long_at(const jw_long_desc& d, int a, int64_t c, int64_t s)
{
    return ((int64_t)a * d.red_sys + s) * d.nchunks + c;
}

template <int MODE, bool ALL3>
__device__ __forceinline__ void long64_pass(const double* tab, double2* stg, int lane, const jw_pass64& g,
                                            const jw_long_desc& d, int64_t c, int64_t sys0)
{
    constexpr int R = kLong64R;
    constexpr int C = 64 * R;
    const jw_run_desc& rd = d.run;
    const double* rhs = reinterpret_cast<const double*>(rd.rhs);
    double* red = reinterpret_cast<double*>(d.red);
    const int64_t rb = c * C;
    jw_edge64 e;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        e.h1[s] = 0.0;
        e.h2[s] = 0.0;
        e.cw[s] = 0.0;
        e.cz[s] = 0.0;
        e.cy[s] = 0.0;
    }
    if constexpr (MODE != JW_LONG_COEF) dma_pair<R>(stg, lane, g, sys0);
    // the edges, loaded while the DMA is in flight
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int64_t sys = sys0 + s;
        if constexpr (MODE == JW_LONG_COEF) {
            // unit carries, zero right-hand sides: system 0 takes w, 1 takes z, 2 takes y
            if (lane == 0 && sys == 0) e.cw[s] = 1.0;
            if (lane == 63 && sys == 1) e.cz[s] = 1.0;
            if (lane == 63 && sys == 2) e.cy[s] = 1.0;
        } else if (sys < g.nsys) {
            const int64_t base = sys * rd.stride + rd.offSol;
            if (rd.band) {
                if constexpr (MODE == JW_LONG_REDUCE) {
                    // rows next to the chunk, before any chunk is written: lane 0 rows rb-1,
                    // rb-2, lane 63 rows rb+C, rb+C+1 (inside the input keep-range)
                    const int64_t g1 = lane == 0 ? rb - 1 : rb + C;
                    const int64_t g2 = lane == 0 ? rb - 2 : rb + C + 1;
                    const bool edge = lane == 0 || lane == 63;
                    if (edge && g1 >= rd.in_lo && g1 < rd.in_hi) e.h1[s] = rhs[base + g1];
                    if (edge && g2 >= rd.in_lo && g2 < rd.in_hi) e.h2[s] = rhs[base + g2];
                } else {
                    // the same rows as REDUCE read them (the step runs in place)
                    const bool lo = lane == 0 && c > 0, hi = lane == 63 && c + 1 < d.nchunks;
                    if (lo || hi) {
                        e.h1[s] = red[lo ? long_at(d, 6, c - 1, sys) : long_at(d, 3, c + 1, sys)];
                        e.h2[s] = red[lo ? long_at(d, 5, c - 1, sys) : long_at(d, 4, c + 1, sys)];
                    }
                }
            }
            if constexpr (MODE == JW_LONG_FINISH) {
                if (lane == 0) e.cw[s] = red[long_at(d, 7, c, sys)];
                if (lane == 63) {
                    e.cz[s] = red[long_at(d, 8, c, sys)];
                    e.cy[s] = red[long_at(d, 9, c, sys)];
                }
            }
        }
    }
    double x0[R], x1[R];
    if constexpr (MODE == JW_LONG_COEF) {
#pragma unroll
        for (int i = 0; i < R; i++) {
            x0[i] = 0.0;
            x1[i] = 0.0;
        }
    } else {
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the DMA writes have landed in LDS
        compiler_fence();
        read_pair<R>(stg, lane, g, sys0, x0, x1);
    }
    if constexpr (MODE == JW_LONG_REDUCE) {
        // the rows the neighbouring chunks' banded matvec needs, as read
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            if (sys >= g.nsys) continue;
            if (lane == 0) {
                red[long_at(d, 3, c, sys)] = s == 0 ? x0[0] : x1[0];
                red[long_at(d, 4, c, sys)] = s == 0 ? x0[1] : x1[1];
            }
            if (lane == 63) {
                red[long_at(d, 5, c, sys)] = s == 0 ? x0[R - 2] : x1[R - 2];
                red[long_at(d, 6, c, sys)] = s == 0 ? x0[R - 1] : x1[R - 1];
            }
        }
    }
    compute_pair<R, ALL3, 1, true>(tab, lane, g, x0, x1, &e);
    if constexpr (MODE == JW_LONG_FINISH) {
        write_blocked<R>(stg, lane, x0, x1);
        store_pair<R>(stg, lane, g, sys0);
        rezero_input<R>(stg, lane, g);
        __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): LDS traffic done before the next DMA lands
        compiler_fence();
    } else {
        // boundary values: W' at the chunk's last row, Z' and y at its first
        double* coef = red + (int64_t)JW_LONG_ARRAYS * d.nchunks * d.red_sys;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            if (sys >= g.nsys) continue;
            const double y0 = s == 0 ? x0[0] : x1[0];
            if constexpr (MODE == JW_LONG_COEF) {
                if (lane == 63) coef[9 * c + 3 * sys + 0] = e.wl[s];
                if (lane == 0) {
                    coef[9 * c + 3 * sys + 1] = e.zf[s];
                    coef[9 * c + 3 * sys + 2] = y0;
                }
            } else {
                if (lane == 63) red[long_at(d, 0, c, sys)] = e.wl[s];
                if (lane == 0) {
                    red[long_at(d, 1, c, sys)] = e.zf[s];
                    red[long_at(d, 2, c, sys)] = y0;
                }
            }
        }
    }
}

// A persistent grid: workgroup b takes the units [b per_wg, (b + 1) per_wg) of the
// (chunk, pair of systems) space in chunk order, stages each chunk's table in LDS once and
// spreads the chunk's pairs over its wavefronts.
template <int MODE, bool ALL3>
__global__ __launch_bounds__(max_threads<kLong64R>()) void jw_long64_kernel(jw_long_desc d, int64_t per_wg)
{
    constexpr int R = kLong64R;
    constexpr int C = 64 * R;
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem + NTAB * 8) + (size_t)wave * 64 * R;
    const jw_run_desc& rd = d.run;
    const int64_t nc = d.nchunks;
    const int64_t nsys = MODE == JW_LONG_COEF ? 3 : rd.nsys;
    const int64_t ppc = (nsys + 1) >> 1;
    const int64_t u0 = (int64_t)blockIdx.x * per_wg;
    const int64_t u1 = u0 + per_wg < nc * ppc ? u0 + per_wg : nc * ppc;
    for (int64_t c = u0 / ppc; c < nc && c * ppc < u1; c++) {
        __syncthreads();   // the previous chunk's passes are done: its table may go
        copy_table_lds<NTAB>(smem, reinterpret_cast<const double*>(d.table) + c * d.tab_elems, wave, nwaves, lane);
        zero_staging<R>(stg, lane);   // this chunk's keep-range may be narrower
        __syncthreads();
        // keep-ranges relative to the chunk, inside its rows of the system
        const int64_t rb = c * C;
        const int64_t cn = (int64_t)rd.N - rb < C ? (int64_t)rd.N - rb : C;
        auto clip = [&](int64_t v) -> int {
            const int64_t w = v - rb;
            return (int)(w < 0 ? 0 : (w > cn ? cn : w));
        };
        const int ilo = clip(rd.in_lo), olo = clip(rd.out_lo);
        const int ihi = clip(rd.in_hi) > ilo ? clip(rd.in_hi) : ilo;
        const int ohi = clip(rd.out_hi) > olo ? clip(rd.out_hi) : olo;
        const int64_t q0 = (u0 > c * ppc ? u0 : c * ppc) - c * ppc;
        const int64_t q1 = (u1 < (c + 1) * ppc ? u1 : (c + 1) * ppc) - c * ppc;
        for (int64_t q = q0 + wave; q < q1; q += nwaves) {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(rd.rhs), nsys, rd.stride, rd.offSol + rb, ilo,
                                            ihi, olo, ohi, rd.kl, rd.ku, rd.fwd, rd.b1, rd.b2,
                                            MODE == JW_LONG_FINISH && c == 0 && rd.apply_scale, rd.scale_hi, lane);
            long64_pass<MODE, ALL3>(tab, stg, lane, g, d, c, 2 * q);
        }
    }
}

// The front of a long FP64 step in one launch: the chunks' table builds, then COEF and
// REDUCE of the first slice of systems behind them.  Workgroups take tickets in start
// order.  The first nbuild build the tables, NSUB chunks each (pivots handed from build to
// build), mark the build's tables written (tflags) and leave.  The others drain a queue
// of units in chunk order (unit 0 of a chunk: COEF on two wavefronts; unit j > 0: REDUCE of
// pairs [(j-1) P, j P), two passes per wavefront), waiting for each chunk's tables before
// staging them, so that the reduction runs while later tables are still being built.  A
// worker's ticket comes after every build's, so every build it waits for has started, and
// builds wait only for earlier builds: the launch drains whatever the residency.  All
// waits are bounded (the device error flag is raised when one runs out).
// sync: [0] ticket, [1] queue counter, [kLong64Tflags + b] build b's tables written;
// slots: build b's last pivot (below); both set by jw_long64_clear_kernel before the launch.
constexpr int kLong64Nsub = 6;   // chunks per table build (3072 rows: 384 threads in the pivot scan)
constexpr int kLong64PairsPerUnit = 2 * (max_threads<kLong64R>() / 64);
constexpr int kLong64Tflags = 32;   // the table flags start one 128-byte line after the counters

// Pivot hand-over between the builds through one 64-bit slot per build that holds the bits
// of the build's last pivot, all ones (a NaN pattern never published: a NaN pivot is
// published as the canonical NaN) until then.  The value carries its own readiness, so
// relaxed atomics suffice (the next build needs nothing else the previous one wrote): one
// store and the polls that see it, no flag, no release or acquire.  The wait is bounded:
// when it runs out the device error flag is raised and the build goes on with a unit pivot.
constexpr unsigned long long kPivotEmpty = ~0ull;

struct SlotHooks {
    double* carry;                  // LDS: carry[0] = g entering the build
    unsigned long long* slots;
    int* err;
    int64_t build;
    int chained;
    __device__ void wait_carry(int tid) const
    {
        if (!chained || build == 0 || tid != 0) return;
        int polls = 0;
        unsigned long long v;
        while ((v = __hip_atomic_load(&slots[build - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == kPivotEmpty) {
            if (++polls > kLongMaxPolls) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = (unsigned long long)__double_as_longlong(1.0);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        carry[0] = __longlong_as_double((long long)v);
    }
    __device__ void publish(int tid, const double* sg, int RL) const
    {
        if (!chained || tid != 0) return;
        const double g = sg[RL - 1];
        const unsigned long long v = g != g ? 0x7FF8000000000000ull : (unsigned long long)__double_as_longlong(g);
        __hip_atomic_store(&slots[build], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

__device__ __forceinline__ void wait_table(int* flag, int* err)
{
    int polls = 0;
    // relaxed polls (no cache invalidation per poll), one acquire once the flag is seen
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++polls > kLongMaxPolls) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(16);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <bool ALL3>
__global__ __launch_bounds__(max_threads<kLong64R>()) void jw_long64_front_kernel(jw_source_desc src, const void* mat,
                                                                                 jw_long_desc d, int* sync,
                                                                                 unsigned long long* slots, int* err,
                                                                                 int nbuild, int chained)
{
    constexpr int R = kLong64R;
    constexpr int C = 64 * R;
    constexpr int SR = kLong64Nsub * C;
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    constexpr int P = kLong64PairsPerUnit;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_id;
    int* queue = sync + 1;
    int* tflags = sync + kLong64Tflags;
    double* table = reinterpret_cast<double*>(const_cast<void*>(d.table));
    if (threadIdx.x == 0) s_id = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int tk = s_id;
    if (tk < nbuild) {
        const int64_t rb = (int64_t)tk * SR;
        // the pivot carry follows the build's five row arrays; thread 0 reads the raw
        // super-diagonal entry of the row before the build (the build's first barrier
        // publishes it)
        double* carry = reinterpret_cast<double*>(smem + prologue_lds_bytes(SR, sizeof(double)));
        if (chained && threadIdx.x == 0)
            carry[1] = rb > 0 ? mem_load<double>(mat, src.udS + source_pos(src, rb - 1, src.N), src.mem_f64) : 0.0;
        const SlotHooks hooks{carry, slots, err, tk, chained};
        prologue_body<double, SlotHooks>(src, mat, table + (int64_t)tk * kLong64Nsub * NTAB, 64, R, 2, smem, rb,
                                         chained ? carry : nullptr, kLong64Nsub, hooks);
        // every wavefront waits for its own table stores (written into the L2), then one
        // agent-scope release by thread 0 writes the L2 back and sets the flag (one L2
        // writeback per build: a fence in every wavefront measured 3x longer under the
        // workers' load, and slowed them)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(&tflags[tk], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem + NTAB * 8) + (size_t)wave * 64 * R;
    const jw_run_desc& rd = d.run;
    const int64_t nc = d.nchunks;
    const int64_t ppc = (rd.nsys + 1) >> 1;
    const int upc = 1 + (int)((ppc + P - 1) / P);
    const int nunits = (int)nc * upc;
    int64_t cur = -1;
    for (;;) {
        __syncthreads();   // the previous unit is finished: s_id and the table may change
        if (threadIdx.x == 0) s_id = __hip_atomic_fetch_add(queue, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_id);
        if (u >= nunits) break;
        const int64_t c = u / upc;
        const int j = u - (int)c * upc;
        if (c != cur) {
            if (threadIdx.x == 0) wait_table(&tflags[c / kLong64Nsub], err);
            __syncthreads();
            copy_table_lds<NTAB>(smem, table + c * NTAB, wave, nwaves, lane);
            zero_staging<R>(stg, lane);   // this chunk's keep-range may be narrower
            __syncthreads();
            cur = c;
        }
        const int64_t rb = c * C;
        const int64_t cn = (int64_t)rd.N - rb < C ? (int64_t)rd.N - rb : C;
        auto clip = [&](int64_t v) -> int {
            const int64_t w = v - rb;
            return (int)(w < 0 ? 0 : (w > cn ? cn : w));
        };
        const int ilo = clip(rd.in_lo), olo = clip(rd.out_lo);
        const int ihi = clip(rd.in_hi) > ilo ? clip(rd.in_hi) : ilo;
        const int ohi = clip(rd.out_hi) > olo ? clip(rd.out_hi) : olo;
        if (j == 0) {
            // COEF: the 3 unit-carry systems, passes 0 and 1 on wavefronts 0 and 1
            if (wave < 2) {
                const jw_pass64 g = make_pass64(reinterpret_cast<double*>(rd.rhs), 3, rd.stride, rd.offSol + rb, ilo,
                                                ihi, olo, ohi, rd.kl, rd.ku, rd.fwd, rd.b1, rd.b2, false, 1.0, lane);
                long64_pass<JW_LONG_COEF, ALL3>(tab, stg, lane, g, d, c, 2 * (int64_t)wave);
            }
        } else {
            const int64_t q0 = (int64_t)(j - 1) * P;
            const int64_t q1 = q0 + P < ppc ? q0 + P : ppc;
            for (int64_t q = q0 + wave; q < q1; q += nwaves) {
                const jw_pass64 g = make_pass64(reinterpret_cast<double*>(rd.rhs), rd.nsys, rd.stride, rd.offSol + rb,
                                                ilo, ihi, olo, ohi, rd.kl, rd.ku, rd.fwd, rd.b1, rd.b2, false, 1.0,
                                                lane);
                long64_pass<JW_LONG_REDUCE, ALL3>(tab, stg, lane, g, d, c, 2 * q);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Chained steps (numConsecutiveJWIterations = K): a wavefront owns a pair of systems
// through all K iterations, whose right-hand sides stay in its registers in between (one
// read and one write of HBM per K-step instead of K).  Iteration k applies its own
// format-2 table; the K tables stay in global memory (every wavefront reads them, so they
// are served from L2); LDS holds each wavefront's staging.  The input keep-range applies
// to the first iteration's rows, the output keep-range to the last one's.
// ---------------------------------------------------------------------------
template <int R>
constexpr int chain_waves() { return 4; }

template <int R, bool ALL3>
__global__ __launch_bounds__(256) void jw_chain64_kernel(jw_run_desc d, const double* tables, int64_t tab_elems,
                                                         int K)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem) + (size_t)wave * 64 * R;
    zero_staging<R>(stg, lane);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): zeros written before any DMA lands
    compiler_fence();
    const int64_t npass = (d.nsys + 1) >> 1;
    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol,
                                        (int)d.in_lo, (int)d.in_hi, (int)d.out_lo, (int)d.out_hi, d.kl, d.ku,
                                        d.fwd, d.b1, d.b2, d.apply_scale, d.scale_hi, lane);
        dma_pair<R>(stg, lane, g, 2 * q);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the DMA writes have landed in LDS
        compiler_fence();
        double x0[R], x1[R];
        read_pair<R>(stg, lane, g, 2 * q, x0, x1);
        for (int k = 0; k < K; k++) compute_pair<R, ALL3>(tables + (int64_t)k * tab_elems, lane, g, x0, x1);
        write_blocked<R>(stg, lane, x0, x1);
        store_pair<R>(stg, lane, g, 2 * q);
        rezero_input<R>(stg, lane, g);
        __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): LDS traffic done before the next DMA lands
        compiler_fence();
    }
}

// ---------------------------------------------------------------------------
// Long systems (R = 32 rows per lane, N <= 2048): one system per wavefront pass.
// The 140 KiB table stays in global memory (every wavefront of the grid reads
// it, so it is served from L2) and coefficients are streamed row by row instead
// of held in registers; LDS holds the 16 KiB staging of each wavefront.
// ---------------------------------------------------------------------------
// table reads through a buffer resource: one VGPR lane offset, the row offset in
// an SGPR, no 64-bit address per coefficient
typedef double jw_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 tb2(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    const jw_d2v d = __builtin_bit_cast(jw_d2v, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    return make_double2(d.x, d.y);
}
__device__ __forceinline__ double tb1(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

template <int R, bool ALL3>
__device__ __forceinline__ void compute_single(__amdgpu_buffer_rsrc_t tr, int lane, const jw_pass64& g,
                                               double (&x)[R])
{
    constexpr int SC = 8 * R * 64 * 8;        // byte offset of the scan multipliers
    const int lp = lane * 16;                 // a lane's coefficient pair within a row block
    const int ls = lane * 8;                  // a lane's scan multiplier
    // coefficient pair p of row-in-lane i: byte offset ((p R + i) 64 + lane) 16
    {
        double m1 = 0.0, m2 = 0.0, p1 = 0.0, p2 = 0.0;
        if (g.kl >= 1) m1 = dpp_d<DPP_WAVE_SHR1>(x[R - 1]);
        if (g.kl >= 2) m2 = dpp_d<DPP_WAVE_SHR1>(x[R - 2]);
        if (g.ku >= 1) p1 = dpp_d<DPP_WAVE_SHL1>(x[0]);
        if (g.ku >= 2) p2 = dpp_d<DPP_WAVE_SHL1>(x[1]);
#pragma unroll
        for (int i = 0; i < R; i++) {
            const double2 c0 = tb2(tr, lp, (0 * R + i) * 1024);
            const double2 c1 = tb2(tr, lp, (1 * R + i) * 1024);
            const double vp2 = tb1(tr, lp, (2 * R + i) * 1024);
            const double xi = x[i];
            const double n1 = (i + 1 < R) ? x[(i + 1) % R] : p1;
            const double n2 = (i + 2 < R) ? x[(i + 2) % R] : ((i + 2 == R) ? p1 : p2);
            double u = c1.x * xi;
            u = __builtin_fma(c0.y, m1, u);
            u = __builtin_fma(c0.x, m2, u);
            u = __builtin_fma(c1.y, n1, u);
            u = __builtin_fma(vp2, n2, u);
            m2 = m1;
            m1 = xi;
            x[i] = u;
        }
    }
    if (ALL3 || g.fwd) {
        double a = x[0];
#pragma unroll
        for (int i = 1; i < R; i++) a = __builtin_fma(-tb1(tr, lp + 8, (2 * R + i) * 1024), a, x[i]);
        a = __builtin_fma(tb1(tr, ls, SC + 0 * 512), dpp_d<DPP_ROW_SHR(1)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + 1 * 512), dpp_d<DPP_ROW_SHR(2)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + 2 * 512), dpp_d<DPP_ROW_SHR(4)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + 3 * 512), dpp_d<DPP_ROW_SHR(8)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + 4 * 512), dpp_rm<DPP_ROW_BCAST15, 0xa>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + 5 * 512), dpp_rm<DPP_ROW_BCAST31, 0xc>(a), a);
        double w = dpp_d<DPP_WAVE_SHR1>(a);
#pragma unroll
        for (int i = 0; i < R; i++) {
            w = __builtin_fma(-tb1(tr, lp + 8, (2 * R + i) * 1024), w, x[i]);
            x[i] = w;
        }
    }
#pragma unroll
    for (int stage = 0; stage < 2; stage++) {
        if (!ALL3 && stage == 0 && !g.b1) continue;
        if (!ALL3 && stage == 1 && !g.b2) continue;
        const int j = 1 + stage;
        double a = x[R - 1];
#pragma unroll
        for (int i = R - 2; i >= 0; i--) a = __builtin_fma(-tb1(tr, lp + 8 * stage, (3 * R + i) * 1024), a, x[i]);
        a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 0) * 512), dpp_d<DPP_ROW_SHL(1)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 1) * 512), dpp_d<DPP_ROW_SHL(2)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 2) * 512), dpp_d<DPP_ROW_SHL(4)>(a), a);
        a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 3) * 512), dpp_d<DPP_ROW_SHL(8)>(a), a);
        {
            const double s1 = readlane_d(a, 16), s2 = readlane_d(a, 32), s3 = readlane_d(a, 48);
            a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 4) * 512), s1, a);
            a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 5) * 512), s2, a);
            a = __builtin_fma(tb1(tr, ls, SC + (j * 8 + 6) * 512), s3, a);
        }
        double z = dpp_d<DPP_WAVE_SHL1>(a);
#pragma unroll
        for (int i = R - 1; i >= 0; i--) {
            z = __builtin_fma(-tb1(tr, lp + 8 * stage, (3 * R + i) * 1024), z, x[i]);
            x[i] = z;
        }
    }
    if (g.apply_scale) x[0] *= g.row0_scale;
}

// one system per pass: stg holds 64 R / 2 chunks (the first system region of a pair staging)
template <int R, bool ALL3>
__device__ __forceinline__ void jw_single64(__amdgpu_buffer_rsrc_t tab, double2* stg, int lane, const jw_pass64& g,
                                            int64_t sys)
{
    constexpr int H = R / 2;
    const double* src = g.rhs + sys * g.stride + g.offSol;
#pragma unroll
    for (int k = 0; k < H; k++) {
        const int c = swz<H>(64 * k + lane);
        if (c >= g.icl && c < g.ich)
            __builtin_amdgcn_global_load_lds(src + 2 * c, (__attribute__((address_space(3))) void*)(stg + 64 * k),
                                             16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the DMA writes have landed in LDS
    compiler_fence();
    double x[R];
#pragma unroll
    for (int m = 0; m < H; m++) {
        const double2 v = stg[swz<H>(lane * H + m)];
        x[2 * m] = v.x;
        x[2 * m + 1] = v.y;
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const bool on = e == 0 ? g.in_edge_lo : g.in_edge_hi;
        const int r = e == 0 ? g.in_lo : g.in_hi - 1;
        if (on && lane == r / R) {
            const double v = src[r];
#pragma unroll
            for (int k = 0; k < R; k++)
                if (k == r % R) x[k] = v;
        }
    }
    compute_single<R, ALL3>(tab, lane, g, x);
    compiler_fence();
#pragma unroll
    for (int m = 0; m < H; m++) stg[swz<H>(lane * H + m)] = make_double2(x[2 * m], x[2 * m + 1]);
    compiler_fence();
    double2 v[H];
#pragma unroll
    for (int k = 0; k < H; k++) v[k] = stg[swz<H>(64 * k + lane)];
#pragma unroll
    for (int k = 0; k < H; k++) asm volatile("" : "+v"(v[k].x), "+v"(v[k].y));
    double2* dst2 = reinterpret_cast<double2*>(g.rhs) + ((sys * g.stride + g.offSol) >> 1);
#pragma unroll
    for (int k = 0; k < H; k++) {
        const int c = 64 * k + lane;
        if (c >= g.ocl && c < g.och) store_stream(dst2 + c, v[k]);
    }
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const bool on = e == 0 ? g.out_edge_lo : g.out_edge_hi;
        const int r = e == 0 ? g.out_lo : g.out_hi - 1;
        if (on && lane == 0) {
            const double2 w = stg[swz<H>(r >> 1)];
            g.rhs[sys * g.stride + g.offSol + r] = (r & 1) ? w.y : w.x;
        }
    }
#pragma unroll
    for (int k = 0; k < H; k++) {
        const int c = swz<H>(64 * k + lane);
        if (!(c >= g.icl && c < g.ich)) stg[64 * k + lane] = make_double2(0.0, 0.0);
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): LDS traffic done before the next DMA lands
    compiler_fence();
}

template <int R, bool ALL3>
__global__ __launch_bounds__(256) void jw_main64g_kernel(jw_run_desc d)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // word 3 of the buffer resource: the gfx9 raw-buffer default (data format 32)
    const __amdgpu_buffer_rsrc_t tab = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(d.table), 0,
                                                                         JW_TAB64_ELEMS(R) * 8, 0x00020000);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem) + (size_t)wave * 32 * R;
#pragma unroll
    for (int k = 0; k < R / 2; k++) stg[64 * k + lane] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < d.nsys; q += (int64_t)gridDim.x * nwaves) {
        const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol,
                                        (int)d.in_lo, (int)d.in_hi, (int)d.out_lo, (int)d.out_hi, d.kl, d.ku,
                                        d.fwd, d.b1, d.b2, d.apply_scale, d.scale_hi, lane);
        jw_single64<R, ALL3>(tab, stg, lane, g, q);
    }
}

// Direct blocked loads (compute / store split kernel): lane t fetches its own R
// rows, 16-byte chunks t H .. t H + H-1, of both systems, straight into the
// registers the recurrences use.  One instruction's 64 accesses are spread over
// a 64 x 16H-byte span that its H sibling instructions complete, so every line
// is fetched from memory once and the other accesses hit the cache; no LDS
// transposition is needed.
template <int R>
__device__ __forceinline__ void load_pair_direct(double2 (&pf)[2][R / 2], int lane, const jw_pass64& g, int64_t sys0)
{
    constexpr int H = R / 2;
    const bool ok1 = sys0 + 1 < g.nsys;
    const double2* src0 = reinterpret_cast<const double2*>(g.rhs) + ((sys0 * g.stride + g.offSol) >> 1);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double2* src = src0 + s * (g.stride >> 1);
#pragma unroll
        for (int m = 0; m < H; m++) {
            const int c = lane * H + m;
            double2 v = make_double2(0.0, 0.0);
            if (c >= g.icl && c < g.ich && (s == 0 || ok1)) v = src[c];
            pf[s][m] = v;
        }
    }
}

// the directly loaded chunks as rows, the odd keep-range ends from global memory
template <int R>
__device__ __forceinline__ void unpack_pair(const double2 (&pf)[2][R / 2], int lane, const jw_pass64& g, int64_t sys0,
                                            double (&x0)[R], double (&x1)[R])
{
#pragma unroll
    for (int m = 0; m < R / 2; m++) {
        x0[2 * m] = pf[0][m].x;
        x0[2 * m + 1] = pf[0][m].y;
        x1[2 * m] = pf[1][m].x;
        x1[2 * m + 1] = pf[1][m].y;
    }
    read_edges<R>(lane, g, sys0, x0, x1);
}

// Fused block operators (compute / store split kernel, IOMODE 1 and 2): the pair
// of a pass is the re/im of complex system q of an interleaved (re, im) array.
// Input (block layout 4): rows of the input keep-range, each scaled by the input
// operator, other rows zero; output (IOMODE 1, block layout 7): rows of the
// output keep-range, each scaled by the output operator, then added to the
// complex output when accumulating; IOMODE 2 writes the two real systems (the
// layout-4 packed form) like the plain step.  The arithmetic per element is the
// block kernel's (jw_block_kernel), so the fused step equals the chain of plans.
template <int R>
__device__ __forceinline__ void load_pair_packed(double2 (&pf)[2][R / 2], int lane, const jw_run_desc& d, int64_t q)
{
    constexpr int H = R / 2;
    const double2* src = reinterpret_cast<const double2*>(d.in_c) + (q * d.in_stride_c + d.in_off_c);
    const bool zero_in = d.in_op == 1 && d.in_scale == 0.0;   // baked ScaleC of 0: zeros, input not read
#pragma unroll
    for (int m = 0; m < H; m++) {
        const int r0 = lane * R + 2 * m, r1 = r0 + 1;   // the lane's own rows, as in load_pair_direct
        double2 e0 = make_double2(0.0, 0.0), e1 = make_double2(0.0, 0.0);
        if (!zero_in && r0 >= d.in_lo && r0 < d.in_hi) e0 = src[r0];
        if (!zero_in && r1 >= d.in_lo && r1 < d.in_hi) e1 = src[r1];
        pf[0][m] = make_double2(e0.x, e1.x);
        pf[1][m] = make_double2(e0.y, e1.y);
    }
}

// the input operator on a lane's rows, once they are in registers (rows >= N are zero)
template <int R>
__device__ __forceinline__ void scale_input(double (&x0)[R], double (&x1)[R], int lane, const jw_run_desc& d,
                                            const double* fin)
{
    if (d.in_op < 1 || d.in_op > 4) return;
#pragma unroll
    for (int i = 0; i < R; i++) {
        const double f = fin[lane * R + i];
        x0[i] *= f;
        x1[i] *= f;
    }
}

// per-row factors of the block operators, once per workgroup into LDS (rows >= N
// get 1, so their zeros stay exact zeros even where a factor would be NaN)
__device__ __forceinline__ void block_factors(double* fin, double* fout, int rows, const jw_run_desc& d)
{
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        const bool in = r < d.N;
        fin[r] = in && d.in_op >= 1 && d.in_op <= 4 ? jw_block_scale(d.in_op, r, d.in_lshift, d.in_scale) : 1.0;
        fout[r] = in && d.out_op >= 1 && d.out_op <= 4 ? jw_block_scale(d.out_op, r, d.out_lshift, d.out_scale) : 1.0;
    }
}

template <int R>
__device__ __forceinline__ void store_pair_packed(const double2* stg, int lane, const jw_run_desc& d, int64_t q,
                                                  const double* fout)
{
    constexpr int H = R / 2;
    constexpr int SYSC = 64 * H;
    double2* dst = reinterpret_cast<double2*>(d.out_c) + (q * d.out_stride_c + d.out_off_c);
    const double* s0 = reinterpret_cast<const double*>(stg);
    const double* s1 = reinterpret_cast<const double*>(stg + SYSC);
    const bool zero_out = d.out_op == 1 && d.out_scale == 0.0;
    const bool scaled = d.out_op >= 1 && d.out_op <= 4;
    // the previous output first (all loads in flight together), then the staging
    double2 prev[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int r = 64 * k + lane;
        prev[k] = make_double2(0.0, 0.0);
        if (d.out_acc && r >= d.out_lo && r < d.out_hi) prev[k] = dst[r];
    }
    double re[R], im[R], f[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int r = 64 * k + lane;   // complex row r: chunk r / 2, half r % 2, of both systems
        const int e = 2 * swz<H>(r >> 1) + (r & 1);
        re[k] = s0[e];
        im[k] = s1[e];
        f[k] = scaled ? fout[r] : 1.0;
    }
#pragma unroll
    for (int k = 0; k < R; k++) asm volatile("" : "+v"(re[k]), "+v"(im[k]), "+v"(f[k]));
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int r = 64 * k + lane;
        if (r < d.out_lo || r >= d.out_hi) continue;
        double a = re[k], b = im[k];
        if (zero_out) {
            a = 0.0;
            b = 0.0;
        } else if (scaled) {
            a *= f[k];
            b *= f[k];
        }
        if (d.out_acc) {
            a += prev[k].x;
            b += prev[k].y;
        }
        store_stream(dst + r, make_double2(a, b));
    }
}

// Compute / store split.  The first `ncomp` wavefronts of a workgroup compute,
// the others store.  Every workgroup first builds the coefficient table in its
// LDS.  A compute wavefront owns an output staging; per pass it takes the rows
// of its pair that it loaded straight into registers one pass earlier, issues
// the loads of its next pair at once, computes, waits until a storer has
// emptied its staging, writes the results there and raises its `full` flag.  Storer
// wavefronts poll the flags of their compute wavefronts, stream full stagings
// to global memory (nontemporal stores) and lower the flag.  Compute
// wavefronts issue no global stores, so their vmcnt waits never wait for write
// acknowledgements; stores stay in flight while compute goes on.
// Every wait on a flag is bounded (kSpinLimit polls without progress); a
// wavefront that runs out stops and raises *err, so the grid always drains.
constexpr int kSpinLimit = 1 << 22;

// Flags and pair indices are relaxed workgroup-scope atomics in LDS: they
// compile to plain LDS accesses with no wait on outstanding global loads (a
// volatile access would wait for them and stall the prefetch).  The writer
// waits lgkmcnt(0) for its staging writes before it raises a flag.
__device__ __forceinline__ int lds_flag(int* f)
{
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void lds_set(int* f, int v)
{
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// waits until *f == v; false when `polls` polls ran out
__device__ __forceinline__ bool wait_flag(int* f, int v, int polls = kSpinLimit)
{
    for (int it = 0; it < polls; it++) {
        if (lds_flag(f) == v) return true;
        __builtin_amdgcn_s_sleep(2);
    }
    return false;
}

template <int R>
constexpr int split_threads()
{
    return R <= 4 ? 1024 : (R == 8 ? 768 : 320);
}

template <int R, bool ALL3, int IOMODE>
__global__ __launch_bounds__(split_threads<R>()) void jw_main64s_kernel(jw_run_desc d, int ncomp, int* err,
                                                                        jw_source_desc src, const void* mat)
{
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_full[16], s_done[16];
    __shared__ long long s_pair[16];
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* out_base = reinterpret_cast<double2*>(smem + NTAB * 8);
    const int64_t npass = (d.nsys + 1) >> 1;
    const int64_t step = (int64_t)gridDim.x * ncomp;
    const int polls = d.wait_polls > 0 ? d.wait_polls : (d.wait_polls < 0 ? 0 : kSpinLimit);
    int64_t q = (int64_t)blockIdx.x * ncomp + wave;
    // a compute wavefront's first pair is loaded into registers before the table
    // build, so its latency overlaps the build
    double2 pf[2][R / 2];
    if (wave < ncomp && q < npass) {
        if (IOMODE != 0) {
            load_pair_packed<R>(pf, lane, d, q);
        } else {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol,
                                            (int)d.in_lo, (int)d.in_hi, (int)d.out_lo, (int)d.out_hi, d.kl, d.ku,
                                            d.fwd, d.b1, d.b2, d.apply_scale, d.scale_hi, lane);
            load_pair_direct<R>(pf, lane, g, 2 * q);
        }
    }
    // every workgroup builds the coefficient table in its own LDS from the
    // caller's matrices (no separate prologue launch); the staging area, unused
    // until the barrier below, is the build's scratch
    // (the default per-thread array bounds: smaller ones, 125 instead of 143 VGPRs at R = 8,
    // measured 3-5 % slower on the whole kernel)
    prologue_body<double, NoPivotHooks, true>(src, mat, reinterpret_cast<double*>(smem), 64, R, 2, smem + NTAB * 8);
    // block operator factors past the staging and the build's scratch
    double* fin = reinterpret_cast<double*>(
        smem + NTAB * 8 + max_i64(2LL * 64 * R * 8 * ncomp, (int64_t)prologue_lds_bytes(64 * R, 8)));
    double* fout = fin + 64 * R;
    if (IOMODE != 0) block_factors(fin, fout, 64 * R, d);
    if (wave < ncomp) {
        if (lane == 0) {
            lds_set(&s_full[wave], 0);
            lds_set(&s_done[wave], 0);
        }
    }
    __syncthreads();
    int* vfull = s_full;
    int* vdone = s_done;

    if (wave < ncomp) {
        bool ok = true;
        double2* stg = out_base + (size_t)wave * 64 * R;
        // fused block input: whole systems, zero-filled past N by the packed loads, so
        // the keep-range is [0, N rounded up to even) and read_pair has no edge rows
        const int in_lo = IOMODE != 0 ? 0 : (int)d.in_lo;
        const int in_hi = IOMODE != 0 ? (int)(d.N + (d.N & 1)) : (int)d.in_hi;
        while (q < npass) {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol, in_lo,
                                            in_hi, (int)d.out_lo, (int)d.out_hi, d.kl, d.ku, d.fwd, d.b1, d.b2,
                                            d.apply_scale, d.scale_hi, lane);
            // the rows prefetched a whole pass ago, then the next pair's loads at once
            double x0[R], x1[R];
            unpack_pair<R>(pf, lane, g, 2 * q, x0, x1);
            const int64_t qn = q + step;
            if (qn < npass) {
                if (IOMODE != 0) load_pair_packed<R>(pf, lane, d, qn);
                else load_pair_direct<R>(pf, lane, g, 2 * qn);
            }
            if (IOMODE != 0) scale_input<R>(x0, x1, lane, d, fin);
            compute_pair<R, ALL3>(tab, lane, g, x0, x1);
            if (!wait_flag(vfull + wave, 0, polls)) { ok = false; break; }
            write_blocked<R>(stg, lane, x0, x1);
            if (lane == 0) __hip_atomic_store(&s_pair[wave], (long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): staging and pair index written
            compiler_fence();
            if (lane == 0) lds_set(vfull + wave, 1);
            q = qn;
        }
        if (ok) ok = wait_flag(vfull + wave, 0, polls);
        __builtin_amdgcn_s_waitcnt(0);
        compiler_fence();
        if (lane == 0) {
            lds_set(vdone + wave, 1);
            if (!ok) *err = 1;
        }
        return;
    }

    // storer wavefront j serves compute wavefronts j, j + nstore, ...
    const int nstore = nwaves - ncomp;
    const int j = wave - ncomp;
    int idle = 0;
    for (;;) {
        bool all_done = true, progress = false;
        for (int w = j; w < ncomp; w += nstore) {
            if (lds_flag(vfull + w)) {
                compiler_fence();   // the pair index and staging reads stay after the flag read
                const long long qq = __hip_atomic_load(&s_pair[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const int64_t q = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)qq) |
                                  ((int64_t)__builtin_amdgcn_readfirstlane((int)(qq >> 32)) << 32);
                const jw_pass64 g = make_pass64(reinterpret_cast<double*>(d.rhs), d.nsys, d.stride, d.offSol,
                                                (int)d.in_lo, (int)d.in_hi, (int)d.out_lo, (int)d.out_hi, d.kl,
                                                d.ku, d.fwd, d.b1, d.b2, d.apply_scale, d.scale_hi, lane);
                if (IOMODE == 1) store_pair_packed<R>(out_base + (size_t)w * 64 * R, lane, d, q, fout);
                else store_pair<R>(out_base + (size_t)w * 64 * R, lane, g, 2 * q);
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): the staging has been read
                compiler_fence();
                if (lane == 0) lds_set(vfull + w, 0);
                progress = true;
                all_done = false;
            } else if (!lds_flag(vdone + w)) {
                all_done = false;   // the flag is read before `done`: a raised done with an empty staging is final
            }
        }
        if (all_done) break;
        if (progress) {
            idle = 0;
        } else {
            if (++idle >= polls) {
                if (lane == 0) *err = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// Uneven-size batch: a persistent grid drains a queue of work units (a matrix
// set and a range of its system pairs), in the order the host sorted them
// (largest N first).  A workgroup takes the next unit with one atomic add on
// the queue counter, stages the set's table in LDS when the set changes and
// spreads the unit's pairs over its wavefronts.  Every workgroup leaves the
// loop once the counter passes the last unit.
template <int R>
__global__ __launch_bounds__(max_threads<R>()) void jw_batch64_kernel(const jw_batch_item* items,
                                                                      const jw_batch_unit* units, int nunits,
                                                                      int* counter, int spw)
{
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_unit;
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem + NTAB * 8) + (size_t)wave * 64 * R;
    int cur = -1;
    for (;;) {
        __syncthreads();   // the previous unit is finished: s_unit and the table may change
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= nunits) break;
        const jw_batch_unit un = units[u];
        const jw_batch_item& it = items[un.item];
        if (un.item != cur) {
            copy_table_lds<NTAB>(smem, it.table, wave, nwaves, lane);
            zero_staging<R>(stg, lane);   // the new set's keep-range may be narrower
            __syncthreads();
            cur = un.item;
        }
        if constexpr (R == 2) {
            // sub-wavefront groups: G pairs (2G systems) per wavefront pass
            const int G = spw >= 8 && it.N <= 32 ? 4 : (spw >= 4 && it.N <= 64 ? 2 : 1);
            if (G > 1) {
                const int tl = lane & (64 / G - 1);
                const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0,
                                                it.N, 0, it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0,
                                                it.scale, tl);
                const int64_t sys_end = it.nsys < 2 * (int64_t)un.pair1 ? it.nsys : 2 * (int64_t)un.pair1;
                for (int q = un.pair0 + wave * G; q < un.pair1; q += nwaves * G) {
                    if (G == 2) jw_groups64<2>(tab, stg, lane, g, 2 * (int64_t)q, sys_end);
                    else jw_groups64<4>(tab, stg, lane, g, 2 * (int64_t)q, sys_end);
                }
                continue;
            }
        }
        for (int q = un.pair0 + wave; q < un.pair1; q += nwaves) {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0,
                                            it.N, 0, it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0,
                                            it.scale, lane);
            jw_pair64<R, true>(tab, stg, lane, g, 2 * (int64_t)q);
        }
    }
    queue_leave(counter);
}

// LDS of the build-in-LDS queue kernel: the table, then the wavefronts' stagings, which
// also hold the table build's scratch while a set's table is built
__host__ __device__ constexpr int64_t batch64b_lds_bytes(int R, int waves)
{
    return (int64_t)JW_TAB64_ELEMS(R) * 8 +
           ((int64_t)2 * 64 * R * 8 * waves > (int64_t)(5 * 64 * R) * 8 ? (int64_t)2 * 64 * R * 8 * waves
                                                                         : (int64_t)(5 * 64 * R) * 8);
}

// one set's format-2 table built into LDS (the stagings after it as scratch); kept out of
// line so that the build's registers do not add to the solve loop's
// (the dynamic LDS is named here, not passed in, so the build addresses it as LDS)
template <int R>
__device__ __noinline__ void build_table_lds(const jw_prologue_job* job)
{
    extern __shared__ __align__(16) unsigned char smem_build[];
    prologue_body<double, NoPivotHooks, false, pro_rows(64 * R), pro_rows(64 * R)>(
        job->src, job->matrices, reinterpret_cast<double*>(smem_build), 64, R, 2, smem_build + JW_TAB64_ELEMS(R) * 8);
}

// The batch queue with the table built in LDS: as jw_batch64_kernel, but a workgroup
// whose set changes builds the set's format-2 table from the caller's matrices in its
// own LDS (prologue_body, the staging as scratch) instead of copying a table that a
// separate launch wrote to global memory: no table build launch before the queue, no
// table written and read back through HBM; the build's latency is paid per set change
// by the workgroup, while the other workgroups keep HBM busy.  Needs at least 256
// threads and 4 threads per 16 rows (prologue_body's partition).
template <int R>
__global__ __launch_bounds__(max_threads<R>()) void jw_batch64b_kernel(const jw_batch_item* items,
                                                                       const jw_batch_unit* units, int nunits,
                                                                       int* counter, const jw_prologue_job* jobs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    __shared__ int s_unit;
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem + NTAB * 8) + (size_t)wave * 64 * R;
    int cur = -1;
    for (;;) {
        __syncthreads();   // the previous unit is finished: s_unit, the table and the staging may change
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= nunits) break;
        const jw_batch_unit un = units[u];
        const jw_batch_item& it = items[un.item];
        if (un.item != cur) {
            build_table_lds<R>(jobs + un.item);
            __syncthreads();   // the build's scratch (the stagings) is free again
            zero_staging<R>(stg, lane);
            __syncthreads();
            cur = un.item;
        }
        for (int q = un.pair0 + wave; q < un.pair1; q += nwaves) {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0,
                                            it.N, 0, it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0,
                                            it.scale, lane);
            jw_pair64<R, true>(tab, stg, lane, g, 2 * (int64_t)q);
        }
    }
    queue_leave(counter);
}

// 16-byte stores of a pair straight from the registers (lane t: chunks t H .. t H + H-1
// of both systems, the output keep-range's full chunks), the odd range ends by their
// owning lane
template <int R>
__device__ __forceinline__ void store_pair_direct(int lane, const jw_pass64& g, int64_t sys0, const double (&x0)[R],
                                                  const double (&x1)[R])
{
    constexpr int H = R / 2;
    const bool ok1 = sys0 + 1 < g.nsys;
    double2* dst0 = reinterpret_cast<double2*>(g.rhs) + ((sys0 * g.stride + g.offSol) >> 1);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (s == 1 && !ok1) continue;
        double2* dst = dst0 + s * (g.stride >> 1);
        const double(&x)[R] = s == 0 ? x0 : x1;
#pragma unroll
        for (int m = 0; m < H; m++) {
            const int c = lane * H + m;
            // cached stores: the H stores of a lane each fill part of 64 lines that the
            // sibling stores complete, and L2 merges them before the write-back
            if (c >= g.ocl && c < g.och) dst[c] = make_double2(x[2 * m], x[2 * m + 1]);
        }
    }
    if (g.out_edge_lo || g.out_edge_hi) {
        double* base = g.rhs + sys0 * g.stride + g.offSol;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const bool on = e == 0 ? g.out_edge_lo : g.out_edge_hi;
            const int r = e == 0 ? g.out_lo : g.out_hi - 1;
            if (on && lane == r / R) {
#pragma unroll
                for (int k = 0; k < R; k++)
                    if (k == r % R) {
                        base[r] = x0[k];
                        if (ok1) base[g.stride + r] = x1[k];
                    }
            }
        }
    }
}

// The batch queue with direct I/O: LDS holds the set's table only (no staging), so more
// wavefronts share one table; every wavefront loads its pairs straight into registers
// one pass ahead, computes and stores straight from registers (nontemporal 16-byte
// stores).  Units and the table hand-over as in jw_batch64_kernel.
constexpr int kBatchDirectThreads = 512;
template <int R>
__global__ __launch_bounds__(kBatchDirectThreads) void jw_batch64d_kernel(const jw_batch_item* items,
                                                                          const jw_batch_unit* units, int nunits,
                                                                          int* counter)
{
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_unit;
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    int cur = -1;
    for (;;) {
        __syncthreads();   // the previous unit is finished: s_unit and the table may change
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= nunits) break;
        const jw_batch_unit un = units[u];
        const jw_batch_item& it = items[un.item];
        const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0, it.N, 0,
                                        it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0, it.scale, lane);
        int64_t q = un.pair0 + wave;
        // the first pair's loads go out before the table copy
        double2 pf[2][R / 2];
        if (q < un.pair1) load_pair_direct<R>(pf, lane, g, 2 * q);
        if (un.item != cur) {
            copy_table_lds<NTAB>(smem, it.table, wave, nwaves, lane);
            __syncthreads();
            cur = un.item;
        }
        while (q < un.pair1) {
            const int64_t qn = q + nwaves;
            double x0[R], x1[R];
            unpack_pair<R>(pf, lane, g, 2 * q, x0, x1);
            if (qn < un.pair1) load_pair_direct<R>(pf, lane, g, 2 * qn);
            compute_pair<R, true>(tab, lane, g, x0, x1);
            store_pair_direct<R>(lane, g, 2 * q, x0, x1);
            q = qn;
        }
    }
    queue_leave(counter);
}

// The same queue for long systems (R = 32): one system per wavefront pass, table
// read through a buffer resource; the unit ranges count systems, not pairs.
__global__ __launch_bounds__(256) void jw_batch64g_kernel(const jw_batch_item* items, const jw_batch_unit* units,
                                                          int nunits, int* counter, int /* systems per pass: 1 */)
{
    constexpr int R = 32;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_unit;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* stg = reinterpret_cast<double2*>(smem) + (size_t)wave * 32 * R;
    int cur = -1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= nunits) break;
        const jw_batch_unit un = units[u];
        const jw_batch_item& it = items[un.item];
        if (un.item != cur) {
#pragma unroll
            for (int k = 0; k < R / 2; k++) stg[64 * k + lane] = make_double2(0.0, 0.0);
            cur = un.item;
        }
        const __amdgpu_buffer_rsrc_t tab = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<void*>(it.table), 0, JW_TAB64_ELEMS(R) * 8, 0x00020000);
        for (int q = un.pair0 + wave; q < un.pair1; q += nwaves) {
            const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0,
                                            it.N, 0, it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0,
                                            it.scale, lane);
            jw_single64<R, true>(tab, stg, lane, g, q);
        }
    }
    queue_leave(counter);
}

// The batch queue with the compute / store split (R <= 16): the first `ncomp`
// wavefronts of a workgroup compute, the others store, as in jw_main64s_kernel.
// Per work unit the workgroup stages the set's table in LDS (when the set
// changes), compute wavefronts run the unit's pairs with the next pair's loads
// in registers one pass ahead, and storers stream the results.  Every
// wavefront meets at two barriers per unit: the table and the stagings change
// only while nobody uses them.  A wait that runs out of polls raises *err and
// the workgroup's abort flag; every wavefront then leaves at the next unit.
template <int R>
constexpr int bsplit_threads()
{
    return R <= 4 ? 1024 : (R == 8 ? 768 : 512);
}

template <int R>
__global__ __launch_bounds__(bsplit_threads<R>()) void jw_batch64s_kernel(const jw_batch_item* items,
                                                                          const jw_batch_unit* units, int nunits,
                                                                          int* counter, int ncomp, int* err,
                                                                          int wait_polls)
{
    // polls before a bounded wait gives up: 0 the default, negative none
    const int polls = wait_polls > 0 ? wait_polls : (wait_polls < 0 ? 0 : kSpinLimit);
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    constexpr int STG = 64 * R;   // staging chunks (16 bytes) of a computing wavefront: two systems
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_full[16], s_done[16];
    __shared__ long long s_pair[16];
    __shared__ int s_unit, s_abort;
    const double* tab = reinterpret_cast<const double*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    double2* out_base = reinterpret_cast<double2*>(smem + NTAB * 8);
    if (threadIdx.x == 0) s_abort = 0;
    if (wave < ncomp && lane == 0) lds_set(&s_full[wave], 0);
    int cur = -1;
    for (;;) {
        __syncthreads();   // the previous unit is finished: stagings empty, the table unused
        // the abort flag is only raised inside a unit, so every wavefront sees the same answer
        if (threadIdx.x == 0) s_unit = s_abort ? nunits : atomicAdd(counter, 1);
        if (wave < ncomp && lane == 0) lds_set(&s_done[wave], 0);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= nunits) break;
        const jw_batch_unit un = units[u];
        const jw_batch_item& it = items[un.item];
        const jw_pass64 g = make_pass64(reinterpret_cast<double*>(it.rhs), it.nsys, it.stride, it.offSol, 0, it.N, 0,
                                        it.N, it.kl, it.ku, true, true, true, it.apply_scale != 0, it.scale, lane);
        int64_t q = un.pair0 + wave;
        // a compute wavefront's first pair of the unit is loaded before the table copy
        double2 pf[2][R / 2];
        if (wave < ncomp && q < un.pair1) load_pair_direct<R>(pf, lane, g, 2 * q);
        if (un.item != cur) {
            copy_table_lds<NTAB>(smem, it.table, wave, nwaves, lane);
            __syncthreads();
            cur = un.item;
        }
        if (wave < ncomp) {
            bool ok = true;
            double2* stg = out_base + (size_t)wave * STG;
            while (q < un.pair1) {
                const int64_t qn = q + ncomp;
                double x0[R], x1[R];
                unpack_pair<R>(pf, lane, g, 2 * q, x0, x1);
                if (qn < un.pair1) load_pair_direct<R>(pf, lane, g, 2 * qn);
                compute_pair<R, true>(tab, lane, g, x0, x1);
                if (!wait_flag(&s_full[wave], 0, polls)) { ok = false; break; }
                write_blocked<R>(stg, lane, x0, x1);
                if (lane == 0)
                    __hip_atomic_store(&s_pair[wave], (long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): staging and pair index written
                compiler_fence();
                if (lane == 0) lds_set(&s_full[wave], 1);
                q = qn;
            }
            if (ok) ok = wait_flag(&s_full[wave], 0, polls);
            __builtin_amdgcn_s_waitcnt(0);
            compiler_fence();
            if (lane == 0) {
                lds_set(&s_done[wave], 1);
                if (!ok) {
                    *err = 1;
                    lds_set(&s_abort, 1);
                }
            }
            continue;
        }
        // storer wavefront j serves compute wavefronts j, j + nstore, ...
        const int nstore = nwaves - ncomp;
        const int j = wave - ncomp;
        int idle = 0;
        for (;;) {
            bool all_done = true, progress = false;
            for (int w = j; w < ncomp; w += nstore) {
                if (lds_flag(&s_full[w])) {
                    compiler_fence();   // the pair index and staging reads stay after the flag read
                    const long long qq =
                        __hip_atomic_load(&s_pair[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const int64_t qs = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)qq) |
                                       ((int64_t)__builtin_amdgcn_readfirstlane((int)(qq >> 32)) << 32);
                    store_pair<R>(out_base + (size_t)w * STG, lane, g, 2 * qs);
                    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0): the staging has been read
                    compiler_fence();
                    if (lane == 0) lds_set(&s_full[w], 0);
                    progress = true;
                    all_done = false;
                } else if (!lds_flag(&s_done[w])) {
                    all_done = false;   // `full` is read before `done`: done with an empty staging is final
                }
            }
            if (all_done) break;
            if (progress) {
                idle = 0;
            } else {
                if (++idle >= polls) {
                    if (lane == 0) {
                        *err = 1;
                        lds_set(&s_abort, 1);
                    }
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    queue_leave(counter);
}

template <int R>
const void* kernel_ptr(bool all3)
{
    return all3 ? reinterpret_cast<const void*>(&jw_main64_kernel<R, true>)
                : reinterpret_cast<const void*>(&jw_main64_kernel<R, false>);
}

const void* pick64(int R, bool all3 = true)
{
    switch (R) {
    case 32: return all3 ? reinterpret_cast<const void*>(&jw_main64g_kernel<32, true>)
                         : reinterpret_cast<const void*>(&jw_main64g_kernel<32, false>);
    case 2: return kernel_ptr<2>(all3);
    case 4: return kernel_ptr<4>(all3);
    case 8: return kernel_ptr<8>(all3);
    case 16: return kernel_ptr<16>(all3);
    default: return nullptr;
    }
}

template <int MODE>
const void* long64_ptr(bool all3)
{
    return all3 ? reinterpret_cast<const void*>(&jw_long64_kernel<MODE, true>)
                : reinterpret_cast<const void*>(&jw_long64_kernel<MODE, false>);
}

const void* pick_long64(int mode, bool all3)
{
    switch (mode) {
    case JW_LONG_COEF: return long64_ptr<JW_LONG_COEF>(all3);
    case JW_LONG_REDUCE: return long64_ptr<JW_LONG_REDUCE>(all3);
    case JW_LONG_FINISH: return long64_ptr<JW_LONG_FINISH>(all3);
    default: return nullptr;
    }
}

constexpr int kLong64Waves = max_threads<kLong64R>() / 64;

const void* pick_long64_front(bool all3)
{
    return all3 ? reinterpret_cast<const void*>(&jw_long64_front_kernel<true>)
                : reinterpret_cast<const void*>(&jw_long64_front_kernel<false>);
}

// zeroes the counters and table flags, empties the pivot slots
__global__ void jw_long64_clear_kernel(int* p, int n, unsigned long long* slots, int nslots)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) slots[i] = kPivotEmpty;
}

// clears the work-queue counters (a kernel, so that it can sit in a captured graph)
__global__ void jw_zero_counters_kernel(int* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

template <int R>
const void* batch_ptr()
{
    return reinterpret_cast<const void*>(&jw_batch64_kernel<R>);
}

const void* pick_batch64(int R)
{
    switch (R) {
    case 32: return reinterpret_cast<const void*>(&jw_batch64g_kernel);
    case 2: return batch_ptr<2>();
    case 4: return batch_ptr<4>();
    case 8: return batch_ptr<8>();
    case 16: return batch_ptr<16>();
    default: return nullptr;
    }
}

const void* pick_batch64s(int R)
{
    switch (R) {
    case 2: return reinterpret_cast<const void*>(&jw_batch64s_kernel<2>);
    case 4: return reinterpret_cast<const void*>(&jw_batch64s_kernel<4>);
    case 8: return reinterpret_cast<const void*>(&jw_batch64s_kernel<8>);
    case 16: return reinterpret_cast<const void*>(&jw_batch64s_kernel<16>);
    default: return nullptr;
    }
}

int bsplit_threads_rt(int R) { return R <= 4 ? 1024 : (R == 8 ? 768 : 512); }

template <int R>
const void* split_ptr(bool all3, int io)
{
    if (io == 1)
        return all3 ? reinterpret_cast<const void*>(&jw_main64s_kernel<R, true, 1>)
                    : reinterpret_cast<const void*>(&jw_main64s_kernel<R, false, 1>);
    if (io == 2)
        return all3 ? reinterpret_cast<const void*>(&jw_main64s_kernel<R, true, 2>)
                    : reinterpret_cast<const void*>(&jw_main64s_kernel<R, false, 2>);
    return all3 ? reinterpret_cast<const void*>(&jw_main64s_kernel<R, true, 0>)
                : reinterpret_cast<const void*>(&jw_main64s_kernel<R, false, 0>);
}

// io: 0 plain step, 1 complex in and out (block layouts 4 and 7), 2 complex in
// (layout 4), real systems out
const void* pick_split(int R, bool all3, int io = 0)
{
    switch (R) {
    case 2: return split_ptr<2>(all3, io);
    case 4: return split_ptr<4>(all3, io);
    case 8: return split_ptr<8>(all3, io);
    default: return nullptr;
    }
}

// compute / storer wavefronts of the split kernel: compute wavefronts up to
// what LDS holds (table + output staging each), at most 8 for R = 8 and 12
// below, so that every SIMD holds the same number (2 or 3: a SIMD with one more
// than the others is the slowest), and storers for the rest of the launch bound
// (12 wavefronts for R = 8, 16 below).
bool split_shape(int R, int* ncomp, int* nstore)
{
    const int maxw = R <= 4 ? 16 : (R == 8 ? 12 : 5);
    const int cmax = R <= 4 ? 12 : (R == 8 ? 8 : 2);
    const int64_t table = jwdev_table64_elems(R) * 8;
    const int64_t per = 2LL * 64 * R * 8;   // output staging of 2 systems
    const int64_t avail = kLdsBytesMax - 256 - table - 2LL * 64 * R * 8;   // less the block factors
    int c = (int)(avail / per);
    if (c > cmax) c = cmax;
    if (c < 1 || c >= maxw) return false;
    *ncomp = c;
    *nstore = maxw - c;
    return true;
}

int max_waves64(int R) { return R <= 4 ? 16 : (R == 8 ? 12 : (R == 16 ? 5 : 4)); }

thread_local char g_err64[256];

}  // namespace

extern "C" {

int64_t jwdev_table64_elems(int R) { return JW_TAB64_ELEMS(R); }

/* waves per workgroup: as many as LDS (table + 2 systems of staging per wave)
 * and the kernel's launch bound allow */
/* LDS copy of the table: R <= 16 (R = 32 reads it from global memory) */
static int64_t table_lds_bytes(int R) { return R <= 16 ? jwdev_table64_elems(R) * 8 : 0; }

int jwdev_main64_waves(int R)
{
    if (!pick64(R)) return 0;
    if (R == 32) return 4;
    const int64_t table = table_lds_bytes(R);
    const int64_t per_wave = 2LL * 64 * R * 8;
    int w = (int)((kLdsBytesMax - table) / per_wave);
    if (w > max_waves64(R)) w = max_waves64(R);
    return w < 1 ? 0 : w;
}

int64_t jwdev_main64_lds_bytes(int R, int waves)
{
    if (R == 32) return 64LL * R * 8 * waves;   /* one system per wavefront */
    return table_lds_bytes(R) + 2LL * 64 * R * 8 * waves;
}

/* workgroups per CU, the smaller of the two variants (all recurrences / some) */
int jwdev_main64_occupancy(int R, int waves)
{
    int best = -1;
    for (int v = 0; v < 2; v++) {
        const void* f = pick64(R, v == 0);
        if (!f) return 0;
        const int lds = (int)jwdev_main64_lds_bytes(R, waves);
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds) != hipSuccess) return 0;
        if (best < 0 || nb < best) best = nb;
    }
    return best;
}

int jwdev_main64(const jw_run_desc* r, int R, int waves, int grid, void* stream)
{
    if (!r->rhs || !r->table) { snprintf(g_err64, sizeof(g_err64), "main64: no rhs or table"); return -1; }
    const void* f = pick64(R, r->fwd && r->b1 && r->b2);
    if (!f) { snprintf(g_err64, sizeof(g_err64), "no 64-lane kernel for R=%d", R); return -1; }
    const size_t lds = (size_t)jwdev_main64_lds_bytes(R, waves);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_run_desc*>(r)};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

static const void* pick_chain(int R, bool all3)
{
    switch (R) {
    case 2: return all3 ? reinterpret_cast<const void*>(&jw_chain64_kernel<2, true>)
                        : reinterpret_cast<const void*>(&jw_chain64_kernel<2, false>);
    case 4: return all3 ? reinterpret_cast<const void*>(&jw_chain64_kernel<4, true>)
                        : reinterpret_cast<const void*>(&jw_chain64_kernel<4, false>);
    case 8: return all3 ? reinterpret_cast<const void*>(&jw_chain64_kernel<8, true>)
                        : reinterpret_cast<const void*>(&jw_chain64_kernel<8, false>);
    case 16: return all3 ? reinterpret_cast<const void*>(&jw_chain64_kernel<16, true>)
                         : reinterpret_cast<const void*>(&jw_chain64_kernel<16, false>);
    default: return nullptr;
    }
}

static int64_t chain_lds_bytes(int R) { return 2LL * 64 * R * 8 * 4; }   /* 4 wavefronts' stagings */

/* workgroups (of 4 wavefronts) per CU of the chained-step kernel */
int jwdev_chain64_occupancy(int R)
{
    int best = -1;
    for (int v = 0; v < 2; v++) {
        const void* f = pick_chain(R, v == 0);
        if (!f) return 0;
        const int lds = (int)chain_lds_bytes(R);
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, lds) != hipSuccess) return 0;
        if (best < 0 || nb < best) best = nb;
    }
    return best;
}

int jwdev_chain64(const jw_run_desc* r, const void* tables, int64_t tab_elems, int K, int R, int grid, void* stream)
{
    if (!r->rhs || !tables) { snprintf(g_err64, sizeof(g_err64), "chain64: no rhs or tables"); return -1; }
    const void* f = pick_chain(R, r->fwd && r->b1 && r->b2);
    if (!f || K < 1) { snprintf(g_err64, sizeof(g_err64), "no chained-step kernel for R=%d", R); return -1; }
    const size_t lds = (size_t)chain_lds_bytes(R);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    jw_run_desc rr = *r;
    const double* t = static_cast<const double*>(tables);
    void* args[] = {&rr, &t, &tab_elems, &K};
    e = hipLaunchKernel(f, dim3(grid), dim3(256), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

int jwdev_batch64_occupancy(int R, int waves)
{
    const void* f = pick_batch64(R);
    if (!f || waves < 1 || waves > max_waves64(R)) return 0;
    const int lds = (int)jwdev_main64_lds_bytes(R, waves);
    if (lds > kLdsBytesMax) return 0;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds) != hipSuccess) return 0;
    return nb;
}

int jwdev_batch64(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                  int waves, int grid, int spw, void* stream)
{
    const void* f = pick_batch64(R);
    if (!f) { snprintf(g_err64, sizeof(g_err64), "no batch kernel for R=%d", R); return -1; }
    if (waves < 1 || waves > max_waves64(R) || grid < 1) { snprintf(g_err64, sizeof(g_err64), "bad batch launch"); return -1; }
    const size_t lds = (size_t)jwdev_main64_lds_bytes(R, waves);
    if (lds > (size_t)kLdsBytesMax) { snprintf(g_err64, sizeof(g_err64), "batch LDS %zu too large", lds); return -1; }
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_batch_item**>(&d_items), const_cast<jw_batch_unit**>(&d_units), &nunits,
                    &d_counter, &spw};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

static const void* pick_batch64b(int R)
{
    switch (R) {
    case 2: return reinterpret_cast<const void*>(&jw_batch64b_kernel<2>);
    case 4: return reinterpret_cast<const void*>(&jw_batch64b_kernel<4>);
    case 8: return reinterpret_cast<const void*>(&jw_batch64b_kernel<8>);
    case 16: return reinterpret_cast<const void*>(&jw_batch64b_kernel<16>);
    default: return nullptr;
    }
}

/* the table build needs >= 256 threads and 4 threads per 16 rows */
static bool batch64b_shape_ok(int R, int waves)
{
    return pick_batch64b(R) && waves >= 4 && waves <= max_waves64(R) && 4 * 64 * waves >= 64 * R &&
           batch64b_lds_bytes(R, waves) <= kLdsBytesMax;
}

int jwdev_batch64b_occupancy(int R, int waves)
{
    if (!batch64b_shape_ok(R, waves)) return 0;
    const void* f = pick_batch64b(R);
    const int lds = (int)batch64b_lds_bytes(R, waves);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds) != hipSuccess) return 0;
    return nb;
}

int jwdev_batch64b(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int waves, int grid, const jw_prologue_job* d_jobs, void* stream)
{
    if (!batch64b_shape_ok(R, waves) || grid < 1 || !d_jobs) {
        snprintf(g_err64, sizeof(g_err64), "bad build-in-LDS batch launch (R=%d, waves=%d)", R, waves);
        return -1;
    }
    const void* f = pick_batch64b(R);
    const size_t lds = (size_t)batch64b_lds_bytes(R, waves);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_batch_item**>(&d_items), const_cast<jw_batch_unit**>(&d_units), &nunits,
                    &d_counter, const_cast<jw_prologue_job**>(&d_jobs)};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

static const void* pick_batch64d(int R)
{
    switch (R) {
    case 2: return reinterpret_cast<const void*>(&jw_batch64d_kernel<2>);
    case 4: return reinterpret_cast<const void*>(&jw_batch64d_kernel<4>);
    case 8: return reinterpret_cast<const void*>(&jw_batch64d_kernel<8>);
    case 16: return reinterpret_cast<const void*>(&jw_batch64d_kernel<16>);
    default: return nullptr;
    }
}

int jwdev_batch64d_occupancy(int R, int waves)
{
    const void* f = pick_batch64d(R);
    if (!f || waves < 1 || 64 * waves > kBatchDirectThreads) return 0;
    const int lds = JW_TAB64_ELEMS(R) * 8;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds) != hipSuccess) return 0;
    return nb;
}

int jwdev_batch64d(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int waves, int grid, void* stream)
{
    const void* f = pick_batch64d(R);
    if (!f) { snprintf(g_err64, sizeof(g_err64), "no direct batch kernel for R=%d", R); return -1; }
    if (waves < 1 || 64 * waves > kBatchDirectThreads || grid < 1) {
        snprintf(g_err64, sizeof(g_err64), "bad direct batch launch");
        return -1;
    }
    const size_t lds = (size_t)JW_TAB64_ELEMS(R) * 8;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_batch_item**>(&d_items), const_cast<jw_batch_unit**>(&d_units), &nunits,
                    &d_counter};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

int jwdev_zero_counters(int* d_counters, int n, void* stream)
{
    if (n < 1 || n > 64) return -1;
    hipLaunchKernelGGL(jw_zero_counters_kernel, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                       d_counters, n);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

/* batch queue with the compute / store split (R <= 16).  Compute wavefronts:
 * `ncomp` when 1 <= ncomp and the table plus their stagings fit in LDS with a
 * storer left over, else the default (12 for R <= 4, 8 for R = 8, 4 for
 * R = 16: the same number on every SIMD); storers fill the launch bound. */
int jwdev_batch64s_shape(int R, int ncomp, int* ncomp_out, int* nstore_out)
{
    if (!pick_batch64s(R)) return 0;
    const int nw = bsplit_threads_rt(R) / 64;
    const int64_t per = 2LL * 64 * R * 8;
    const int64_t avail = kLdsBytesMax - 512 - jwdev_table64_elems(R) * 8;
    int cmax = (int)(avail / per);
    if (cmax > nw - 1) cmax = nw - 1;
    const int cdef = R <= 4 ? 12 : (R == 8 ? 8 : 4);
    int c = ncomp >= 1 ? ncomp : cdef;
    if (c > cmax) c = cmax < cdef ? cmax : cdef;
    if (c < 1) return 0;
    *ncomp_out = c;
    *nstore_out = nw - c;
    return 1;
}

int64_t jwdev_batch64s_lds_bytes(int R, int ncomp)
{
    return jwdev_table64_elems(R) * 8 + 2LL * 64 * R * 8 * ncomp;
}

int jwdev_batch64s_occupancy(int R, int ncomp)
{
    int c = 0, st = 0;
    if (!jwdev_batch64s_shape(R, ncomp, &c, &st)) return 0;
    const void* f = pick_batch64s(R);
    const int lds = (int)jwdev_batch64s_lds_bytes(R, c);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * (c + st), lds) != hipSuccess) return 0;
    return nb;
}

int jwdev_batch64s(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int ncomp, int grid, int* d_err, int wait_polls, void* stream)
{
    int c = 0, st = 0;
    if (!jwdev_batch64s_shape(R, ncomp, &c, &st) || grid < 1) {
        snprintf(g_err64, sizeof(g_err64), "no split batch kernel for R=%d", R);
        return -1;
    }
    const void* f = pick_batch64s(R);
    const size_t lds = (size_t)jwdev_batch64s_lds_bytes(R, c);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_batch_item**>(&d_items), const_cast<jw_batch_unit**>(&d_units), &nunits,
                    &d_counter, &c, &d_err, &wait_polls};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * (c + st)), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

/* compute / store split kernel (R <= 8): shape, occupancy, launch */
int jwdev_main64s_shape(int R, int* ncomp, int* nstore)
{
    if (!pick_split(R, true)) return 0;
    return split_shape(R, ncomp, nstore) ? 1 : 0;
}

int64_t jwdev_main64s_lds_bytes(int R, int ncomp)
{
    // output staging of the compute wavefronts, at least the table build's scratch,
    // and the block operators' row factors (in and out) of the fused form
    const int64_t staging = 2LL * 64 * R * 8 * ncomp;
    const int64_t scratch = (int64_t)prologue_lds_bytes(64 * R, 8);
    return jwdev_table64_elems(R) * 8 + max_i64(staging, scratch) + 2LL * 64 * R * 8;
}

int jwdev_main64s_occupancy(int R)
{
    int c = 0, st = 0;
    if (!jwdev_main64s_shape(R, &c, &st)) return 0;
    int best = -1;
    for (int v = 0; v < 6; v++) {
        const void* f = pick_split(R, (v & 1) == 0, v >> 1);
        const int lds = (int)jwdev_main64s_lds_bytes(R, c);
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * (c + st), lds) != hipSuccess) return 0;
        if (best < 0 || nb < best) best = nb;
    }
    return best;
}

int jwdev_main64s(const jw_run_desc* r, const jw_source_desc* src, const void* matrices, int R, int grid, int* d_err,
                  void* stream)
{
    if (!r->rhs || !matrices || !d_err) { snprintf(g_err64, sizeof(g_err64), "main64s: no rhs, matrices or flag"); return -1; }
    int c = 0, st = 0;
    if (!jwdev_main64s_shape(R, &c, &st)) { snprintf(g_err64, sizeof(g_err64), "no split kernel for R=%d", R); return -1; }
    const void* f = pick_split(R, r->fwd && r->b1 && r->b2, (int)r->fused_io);
    const size_t lds = (size_t)jwdev_main64s_lds_bytes(R, c);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return -2;
    void* args[] = {const_cast<jw_run_desc*>(r), &c, &d_err, const_cast<jw_source_desc*>(src), &matrices};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * (c + st)), args, lds, reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : -3;
}

/* Long FP64 systems (16-byte aligned) on the pair machinery: one workgroup of 12
 * wavefronts per CU (the chunk's table and a staging of 2 x 512 rows per wavefront) */
int jwdev_long64_nsub(void) { return kLong64Nsub; }

int64_t jwdev_long64_lds_bytes(void)
{
    return (int64_t)JW_TAB64_ELEMS(kLong64R) * 8 + 2LL * 64 * kLong64R * 8 * kLong64Waves;
}

int jwdev_long64_occupancy(void)
{
    int best = -1;
    for (int m = JW_LONG_COEF; m <= JW_LONG_FINISH + 1; m++) {
        for (int v = 0; v < 2; v++) {
            const void* f = m <= JW_LONG_FINISH ? pick_long64(m, v == 0) : pick_long64_front(v == 0);
            const int lds = (int)jwdev_long64_lds_bytes();
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * kLong64Waves, lds) != hipSuccess) return 0;
            if (best < 0 || nb < best) best = nb;
        }
    }
    return best < 0 ? 0 : best;
}

/* the front of a step (jw_long64_front_kernel): the table builds of every chunk, COEF and
 * REDUCE of the slice in d (its first systems) in one launch, after a fill kernel that
 * zeroes the 32 + nbuild ints of sync and empties the nbuild 8-byte pivot slots at gval */
int jwdev_long64_front(const jw_source_desc* src, const void* matrices, const jw_long_desc* d, int* sync, void* gval,
                       int* err, int cu_count, int occupancy, void* stream)
{
    const int C = 64 * kLong64R;
    const int64_t SR = (int64_t)kLong64Nsub * C;
    if (!matrices || !d->table || !d->red || !d->run.rhs || !sync || !gval || !err) {
        jwdev_set_error("long64 front: no matrices, table, reduced data, rhs or hand-over scratch");
        return -1;
    }
    const int64_t nbuild = (src->N + SR - 1) / SR;
    if (src->N < 1 || nbuild > (1 << 20) || d->nchunks < 1 || (int64_t)d->nchunks * C < d->run.N ||
        d->nchunks > nbuild * kLong64Nsub || d->tab_elems != JW_TAB64_ELEMS(kLong64R) || cu_count < 1 ||
        occupancy < 1 || d->run.nsys < 1 || d->run.nsys > d->red_sys) {
        jwdev_set_error("long64 front: bad size, chunk count, table size, slice or grid");
        return -1;
    }
    if (((uintptr_t)d->run.rhs % 16) != 0 || (d->run.stride % 2) != 0 || (d->run.offSol % 2) != 0) {
        jwdev_set_error("long64 front: systems not 16-byte aligned");
        return -1;
    }
    const int64_t ppc = (d->run.nsys + 1) / 2;
    const int64_t units = (int64_t)d->nchunks * (1 + (ppc + kLong64PairsPerUnit - 1) / kLong64PairsPerUnit);
    if (units > (1LL << 30)) {
        jwdev_set_error("long64 front: too many work units");
        return -1;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int nclear = (int)(kLong64Tflags + nbuild);
    unsigned long long* slots = reinterpret_cast<unsigned long long*>(gval);
    hipLaunchKernelGGL(jw_long64_clear_kernel, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, s, sync, nclear,
                       slots, (int)nbuild);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        jwdev_set_error("long64 front: clear launch failed");
        return -3;
    }
    const void* f = pick_long64_front(d->run.fwd && d->run.b1 && d->run.b2);
    const size_t lds = (size_t)jwdev_long64_lds_bytes();
    if (prologue_lds_bytes((int)SR, sizeof(double)) + 2 * sizeof(double) > lds) {
        jwdev_set_error("long64 front: table build does not fit");
        return -1;
    }
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
        jwdev_set_error("long64 front: hipFuncSetAttribute failed");
        return -2;
    }
    /* every build's workgroup, then a full GPU of workers */
    const int64_t grid = nbuild + (int64_t)cu_count * occupancy;
    jw_source_desc sd = *src;
    const void* mat = matrices;
    jw_long_desc dd = *d;
    int nb = (int)nbuild;
    int chained = src->solve_kind == JW_SOLVE_TRI;
    void* args[] = {&sd, (void*)&mat, &dd, &sync, &slots, &err, &nb, &chained};
    e = hipLaunchKernel(f, dim3((unsigned)grid), dim3(64 * kLong64Waves), args, lds, s);
    if (e != hipSuccess) {
        jwdev_set_error("long64 front: launch failed");
        return -3;
    }
    return 0;
}

int jwdev_long64(const jw_long_desc* d, int cu_count, int occupancy, void* stream)
{
    const int C = 64 * kLong64R;
    if (!d->table || !d->red || (d->mode != JW_LONG_COEF && !d->run.rhs)) {
        jwdev_set_error("long64: no table, reduced data or rhs");
        return -1;
    }
    if (d->mode < JW_LONG_COEF || d->mode > JW_LONG_FINISH || d->nchunks < 1 ||
        (int64_t)d->nchunks * C < d->run.N || d->tab_elems != JW_TAB64_ELEMS(kLong64R) || cu_count < 1 ||
        occupancy < 1) {
        jwdev_set_error("long64: bad mode, chunk count, table size or grid");
        return -1;
    }
    if (d->mode != JW_LONG_COEF && d->run.nsys > d->red_sys) {
        jwdev_set_error("long64: slice exceeds reduced data");
        return -1;
    }
    /* the DMA moves 16-byte row pairs: aligned systems only */
    if (d->mode != JW_LONG_COEF &&
        (((uintptr_t)d->run.rhs % 16) != 0 || (d->run.stride % 2) != 0 || (d->run.offSol % 2) != 0)) {
        jwdev_set_error("long64: systems not 16-byte aligned");
        return -1;
    }
    const int64_t nsys = d->mode == JW_LONG_COEF ? 3 : d->run.nsys;
    if (nsys < 1) return 0;
    const int64_t total = (int64_t)d->nchunks * ((nsys + 1) / 2);
    int64_t grid = (int64_t)cu_count * occupancy;
    const int64_t need = (total + kLong64Waves - 1) / kLong64Waves;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    int64_t per_wg = (total + grid - 1) / grid;
    grid = (total + per_wg - 1) / per_wg;
    const void* f = pick_long64(d->mode, d->run.fwd && d->run.b1 && d->run.b2);
    const size_t lds = (size_t)jwdev_long64_lds_bytes();
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
        jwdev_set_error("long64: hipFuncSetAttribute failed");
        return -2;
    }
    void* args[] = {const_cast<jw_long_desc*>(d), &per_wg};
    e = hipLaunchKernel(f, dim3((unsigned)grid), dim3(64 * kLong64Waves), args, lds,
                        reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        jwdev_set_error("long64: launch failed");
        return -3;
    }
    return 0;
}

}  // extern "C"
