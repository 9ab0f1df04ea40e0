// jw_prologue.h -- the coefficient-table build of the JW step (device code,
// included by jw_device.hip and jw_main64.hip): one workgroup turns one matrix
// set in the reference storage into the table the main kernels apply.
//
// Reference semantics restated (paths relative to the reference root):
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:1561-1733
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_TridiagonalSolve.h:1578-1929
#ifndef PFSOLVE_JW_PROLOGUE_H
#define PFSOLVE_JW_PROLOGUE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jw_common.h"

// ------------------------------------------------------------------------
// table layout (elements of T):
//   coefficient pair p (0..4) of row-in-lane i, lane t:  ((p*R + i)*L + t)*2 + {0,1}
//     p0 = (vm2, vm1)  p1 = (v0, vp1)  p2 = (vp2, l)  p3 = (g, h)  p4 = (D, E)
//   scan multipliers of level k, lane t: SC + (k*L + t)*4 + {0: F, 1: B1, 2: B2}
// ------------------------------------------------------------------------
__host__ __device__ __forceinline__ int64_t tab_coef(int p, int i, int t, int c, int R, int L)
{
    return ((int64_t)(p * R + i) * L + t) * 2 + c;
}
__host__ __device__ __forceinline__ int64_t tab_scan(int k, int t, int j, int R, int L)
{
    return (int64_t)JW_NCOEF * R * L + ((int64_t)k * L + t) * 4 + j;
}
__host__ __device__ __forceinline__ int ilog2_rt(int v)
{
    int r = 0;
    while ((1 << r) < v) r++;
    return r;
}

// ------------------------------------------------------------------------
// prologue: one workgroup (256 threads) per matrix set
//
// Every per-row quantity is staged in LDS (arrays of R*L elements):
//   v[-2..2], l, g, h, D, E   (rows >= N are zero)
// then written in one of two table formats:
//   format 1: the pairs above + Hillis-Steele multipliers over L lanes
//             (jw_main_kernel);
//   format 2 (L = 64): scalings folded in, a_r = D_r g_r,
//       v'_d[r] = a_r v_d[r],  l'_r = a_r l_r / a_{r-1},  h'_r = D_r h_r / D_{r+1},
//     so that the three recurrences are W_r = u'_r - l'_r W_{r-1},
//     Z_r = W_r - h'_r Z_{r+1}, y_r = Z_r - E_r y_{r+1} (W = a w, Z = D z);
//     8 coefficients per row, and per recurrence and lane 8 multipliers of the
//     DPP scan: products over the 2^k lanes a row_shr/row_shl:2^k step
//     combines (k = 0..3), then the cross-row multipliers (forward: row
//     prefix and the row_bcast31 step; backward: one per later row start)
//     (jw_main64_kernel).  Table size JW_TAB64_ELEMS(R).
// Matrix loads are spread over all threads (every load of a thread's rows issued
// at once, the band kept in registers until the fold); the LU scan runs on
// wavefront 0.
// ------------------------------------------------------------------------
constexpr int kPrologueThreads = 512;
constexpr int kProRows = 4;   // rows per thread: a build's rows <= 4 x its threads
constexpr int kMaxRT = 8;     // rows per thread of the pivot scan: a build's rows <= 8 x NP

// rows per thread at most of a build of RL rows on at least 256 threads (row phases and
// pivot scan alike), for the MAXJ / MAXRT bounds of a build whose size is known at compile time
__host__ __device__ constexpr int pro_rows(int RL) { return RL <= 256 ? 1 : (RL + 255) / 256; }

// LDS of prologue_body: 5 arrays of R*L elements (ld -> l -> l', ud -> h -> h', D, E, g)
// five arrays of RL values; the last (g) is at least 16 values long, since the
// pivot scan keeps its wavefront totals there before g is written
__host__ __device__ inline size_t prologue_lds_bytes(int RL, size_t elt)
{
    return (4 * (size_t)RL + (RL > 16 ? (size_t)RL : 16)) * elt;
}

// The band coefficients v_{-2..2} of row r of the matvec (zero outside the matrix),
// P = storage position of row r in the permuted diagonals, in two parts: row_band_load
// issues the row's loads (positions outside the matrix read an in-bounds neighbour: a
// clamped column, or the row's own position), row_band_finish zeroes what lies outside by
// selects and applies the pre-scale.  No per-row branch separates the loads (the compiler
// waits for a load before the join of any branch around it), and a caller can issue the
// loads of several rows before finishing any.
template <typename T>
__device__ __forceinline__ void row_band_load(const jw_source_desc& d, const void* mat, int64_t r, int64_t N,
                                              int64_t P, T (&raw)[6])
{
    using A = Ar<T>;
    const int mf = d.mem_f64;
#pragma unroll
    for (int k = 0; k < 5; k++) raw[k] = A::zero();
    switch (d.band_kind) {
    case JW_BAND_DIAG:
        raw[2] = mem_load<T>(mat, d.mdA + P, mf);
        break;
    case JW_BAND_UPPER:
        raw[3] = mem_load<T>(mat, d.udA + P, mf);
        raw[2] = d.unit_diag ? A::one() : mem_load<T>(mat, d.mdA + P, mf);
        break;
    case JW_BAND_LOWER:
        raw[1] = mem_load<T>(mat, d.ldA + P, mf);
        raw[2] = d.unit_diag ? A::one() : mem_load<T>(mat, d.mdA + P, mf);
        break;
    case JW_BAND_GB:
#pragma unroll
        for (int off = -2; off <= 2; off++) {
            const int64_t col = r + off;
            const int64_t cc = col < 0 ? 0 : (col >= N ? N - 1 : col);
            // band rows outside [-KL, KU] are not in the storage: read the main one
            const int j = (off < -d.KL || off > d.KU) ? d.KU : d.KU - off;
            raw[2 + off] = mem_load<T>(mat, d.offV + (int64_t)j * N + cc, mf);
        }
        break;
    default:
        raw[2] = A::one();
        break;
    }
    raw[5] = d.mdS >= 0 ? mem_load<T>(mat, d.mdS + P, mf) : A::one();
}

template <typename T>
__device__ __forceinline__ void row_band_finish(const jw_source_desc& d, int64_t r, int64_t N, const T (&raw)[6],
                                                T (&v)[5])
{
    using A = Ar<T>;
    // columns of the matvec: N, or in a chained step the previous iteration's rows
    const int64_t ncols = d.Ncols > 0 ? d.Ncols : N;
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] = raw[k];
    switch (d.band_kind) {
    case JW_BAND_UPPER:
        v[3] = A::sel(r + 1 < ncols, raw[3], A::zero());
        break;
    case JW_BAND_LOWER:
        v[1] = A::sel(r >= 1 && r - 1 < ncols, raw[1], A::zero());
        break;
    case JW_BAND_GB:
#pragma unroll
        for (int off = -2; off <= 2; off++) {
            const int64_t col = r + off;
            const bool ok = off >= -d.KL && off <= d.KU && col >= 0 && col < N;
            v[2 + off] = A::sel(ok, raw[2 + off], A::zero());
        }
        break;
    default:
        break;
    }
    if (d.mdS >= 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = A::mul(v[k], raw[5]);
    }
}

// The same in one piece (loads, then selects and the pre-scale), for builds that finish
// each row before loading the next (double-double: with the two-part form the compiler
// keeps the rows' band arrays in scratch memory, measured 3x slower loads).
template <typename T>
__device__ __forceinline__ void row_band(const jw_source_desc& d, const void* mat, int64_t r, int64_t N, int64_t P,
                                         T (&v)[5])
{
    using A = Ar<T>;
    const int mf = d.mem_f64;
    // columns of the matvec: N, or in a chained step the previous iteration's rows
    const int64_t ncols = d.Ncols > 0 ? d.Ncols : N;
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] = A::zero();
    switch (d.band_kind) {
    case JW_BAND_DIAG:
        v[2] = mem_load<T>(mat, d.mdA + P, mf);
        break;
    case JW_BAND_UPPER: {
        const T u = mem_load<T>(mat, d.udA + P, mf);
        v[2] = d.unit_diag ? A::one() : mem_load<T>(mat, d.mdA + P, mf);
        v[3] = A::sel(r + 1 < ncols, u, A::zero());
        break;
    }
    case JW_BAND_LOWER: {
        const T l = mem_load<T>(mat, d.ldA + P, mf);
        v[2] = d.unit_diag ? A::one() : mem_load<T>(mat, d.mdA + P, mf);
        v[1] = A::sel(r >= 1 && r - 1 < ncols, l, A::zero());
        break;
    }
    case JW_BAND_GB: {
        T raw[5];
#pragma unroll
        for (int off = -2; off <= 2; off++) {
            const int64_t col = r + off;
            const int64_t cc = col < 0 ? 0 : (col >= N ? N - 1 : col);
            // band rows outside [-KL, KU] are not in the storage: read the main one
            const int j = (off < -d.KL || off > d.KU) ? d.KU : d.KU - off;
            raw[2 + off] = mem_load<T>(mat, d.offV + (int64_t)j * N + cc, mf);
        }
#pragma unroll
        for (int off = -2; off <= 2; off++) {
            const int64_t col = r + off;
            const bool ok = off >= -d.KL && off <= d.KU && col >= 0 && col < N;
            v[2 + off] = A::sel(ok, raw[2 + off], A::zero());
        }
        break;
    }
    default:
        v[2] = A::one();
        break;
    }
    if (d.mdS >= 0) {
        const T sc = mem_load<T>(mat, d.mdS + P, mf);
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = A::mul(v[k], sc);
    }
}

// storage position of row r of the system the descriptor reads
__device__ __forceinline__ int64_t source_pos(const jw_source_desc& d, int64_t r, int64_t N)
{
    return d.alt ? alt_pos(2 * r + d.alt_parity, d.alt_N)
                 : (d.permuted ? reorder_pos(r, N, d.Nalign, d.Msplit0) : r);
}

// No pivot hand-over between builds.
struct NoPivotHooks {
    __device__ void wait_carry(int) const {}
    template <typename T> __device__ void publish(int, const T*, int) const {}
};

// One table build.  Local row r of the build is row rb + r of the system.  A build of a
// long system's later rows (rb > 0) takes from `carry` the pivot of the row before its
// first (carry[0] = g_{rb-1}) and that row's raw super-diagonal entry (carry[1] =
// ud_{rb-1}), so that its pivots continue the exact recurrence; it writes nsub tables of
// R * L rows each, one after the other (format 1: table_elems apart; format 2, L = 64:
// JW_TAB64_ELEMS(R) apart, the scalings folded across the build's ends too, so that the
// tables of consecutive builds continue each other's scaled recurrences).
// Hooks of a build that hands its pivots on (tridiagonal solves only): wait_carry(tid),
// called by every thread just before the barrier after which carry[0] is first read, and
// publish(tid, sg, RL), called by every thread once the g of every row is final in sg.
// LANE_MAJOR: the rows of a format-2 build go lane-major over the threads (below).  MAXJ,
// MAXRT: rows per thread at most in the row phases (RL <= MAXJ nthr) and in the pivot scan
// (RL <= MAXRT NP); smaller bounds keep a small build's per-thread arrays in fewer registers.
template <typename T, typename H = NoPivotHooks, bool LANE_MAJOR = false, int MAXJ = kProRows, int MAXRT = kMaxRT>
__device__ __forceinline__ void prologue_body(const jw_source_desc& d, const void* mat, T* table, int L, int R, int format,
                              unsigned char* smem_raw, int64_t rb = 0, const T* carry = nullptr, int nsub = 1,
                              const H& hooks = H())
{
    using A = Ar<T>;
    const int RLs = R * L;                 // rows of one table
    const int RL = RLs * nsub;             // rows of this build
    const int64_t tab_elems = (int64_t)JW_NCOEF * RLs + 4 * (int64_t)ilog2_rt(L) * L;
    // format-1 element (p, c) of local row r: table rr / RLs, row-in-lane and lane of rr % RLs
    auto tab1 = [&](int r, int p, int c) -> int64_t {
        const int k = r / RLs, rr = r - k * RLs;
        return k * tab_elems + tab_coef(p, rr % R, rr / R, c, R, L);
    };
    const bool has_carry = rb > 0 && carry != nullptr;
    T* sl = reinterpret_cast<T*>(smem_raw);    // raw ld, then l, then (format 2) l'
    T* sh = sl + RL;                           // raw ud, then h, then (format 2) h'
    T* sD = sh + RL;
    T* sE = sD + RL;
    T* sg = sE + RL;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int64_t N = d.N;
    const int mf = d.mem_f64;
    const bool has_b1 = d.solve_kind == JW_SOLVE_UPPER || d.solve_kind == JW_SOLVE_TRI;
    const bool has_f = d.solve_kind == JW_SOLVE_LOWER || d.solve_kind == JW_SOLVE_TRI;

    // phase 1: raw solve coefficients per row (zeros outside the matrix) into LDS,
    // band coefficients into registers (format 1: straight into the table); row
    // row_of(j) of thread tid: tid + j nthr (consecutive threads read consecutive rows),
    // or with LANE_MAJOR (format 2) lane t of wavefront w takes row t R + i of table k,
    // (k, i) the (w + j nwaves)-th pair, so that a table store of the fold (a coefficient
    // pair of row-in-lane i over the 64 lanes) writes 1 KiB of consecutive memory: without
    // bank conflicts for a table in LDS, at the price of loads that stride R rows over the
    // lanes.  Measured: faster for the split kernel's build (its matrices are read by every
    // workgroup, from the L2), slower for a build that reads its own matrices from HBM
    // (batch queue kernel 4, the long-system builds).  A row of RL or more: no row.
    const int nwv = nthr >> 6;
    const bool lane_major = LANE_MAJOR && format == 2;
    auto row_of = [&](int j) -> int {
        if (!lane_major) return tid + j * nthr;
        const int c = (tid >> 6) + nwv * j;
        const int k = c / R;
        return k * RLs + (tid & 63) * R + (c - k * R);
    };
    T vb[MAXJ][5];
    // FP64: every row's loads first (rows outside the matrix read row 0's positions and
    // are discarded), then the selects, products and LDS stores, so that a thread's rows
    // cost one round of memory latency.  Double-double loads and finishes row by row (the
    // rows' double-double loads at once need more registers than they save time).
    constexpr bool PRE = sizeof(T) == 8;
    T raw[PRE ? MAXJ : 1][6], sr[PRE ? MAXJ : 1][4];
    if constexpr (PRE) {
#pragma unroll
        for (int j = 0; j < MAXJ; j++) {
            const int r = row_of(j);
            const int64_t gr = rb + r;
            // no thread of the wavefront has a row here (wavefront-uniform)
            if (!lane_major ? j * nthr >= RL : (tid >> 6) + nwv * j >= RL / 64) continue;
            const int64_t gq = (r < RL && gr < N) ? gr : 0;
            const int64_t P = source_pos(d, gq, N);
            sr[j][0] = has_f ? mem_load<T>(mat, d.ldS + P, mf) : A::zero();
            sr[j][1] = has_b1 ? mem_load<T>(mat, d.udS + P, mf) : A::zero();
            sr[j][2] = d.has_m ? mem_load<T>(mat, d.mdM + P, mf) : A::one();
            sr[j][3] = d.has_m ? mem_load<T>(mat, d.udM + P, mf) : A::zero();
            row_band_load<T>(d, mat, gq, N, P, raw[j]);
        }
    }
    // format 2 folds D and g of the rows next to the build into its first and last rows:
    // a build of a long system's rows takes D_{rb-1} and D_{rb+RL} from the matrix
    // (every thread reads them: one address each, workgroup-uniform conditions)
    T dprev = A::one(), dnext = A::one();
    if (format == 2 && d.has_m) {
        if (rb >= 1 && rb - 1 < N) dprev = mem_load<T>(mat, d.mdM + source_pos(d, rb - 1, N), mf);
        if (rb + RL < N) dnext = mem_load<T>(mat, d.mdM + source_pos(d, rb + RL, N), mf);
    }
#pragma unroll
    for (int j = 0; j < MAXJ; j++) {
        const int r = row_of(j);
        const int64_t gr = rb + r;
        T ldv = A::zero(), udv = A::zero(), Dv = A::zero(), Ev = A::zero();
#pragma unroll
        for (int k = 0; k < 5; k++) vb[j][k] = A::zero();
        if (r < RL && gr < N) {
            if constexpr (PRE) {
                row_band_finish<T>(d, gr, N, raw[j], vb[j]);
                ldv = A::sel(gr >= 1, sr[j][0], A::zero());
                udv = A::sel(gr + 1 < N, sr[j][1], A::zero());
                Dv = sr[j][2];
                Ev = A::sel(gr + 1 < N, sr[j][3], A::zero());
            } else {
                const int64_t P = source_pos(d, gr, N);
                // the solve diagonals' loads first (the band's products need its own
                // loads), all at the row's position, zeroed by selects where the row has no
                // entry
                T ldr = A::zero(), udr = A::zero(), Dr = A::one(), Er = A::zero();
                if (has_f) ldr = mem_load<T>(mat, d.ldS + P, mf);
                if (has_b1) udr = mem_load<T>(mat, d.udS + P, mf);
                if (d.has_m) {
                    Dr = mem_load<T>(mat, d.mdM + P, mf);
                    Er = mem_load<T>(mat, d.udM + P, mf);
                }
                row_band<T>(d, mat, gr, N, P, vb[j]);
                ldv = A::sel(gr >= 1, ldr, A::zero());
                udv = A::sel(gr + 1 < N, udr, A::zero());
                Dv = Dr;
                Ev = A::sel(gr + 1 < N, Er, A::zero());
            }
        }
        if (r < RL) {
            sl[r] = ldv;
            sh[r] = udv;
            sD[r] = Dv;
            sE[r] = Ev;
            if (format == 1) {
                table[tab1(r, 0, 0)] = vb[j][0];
                table[tab1(r, 0, 1)] = vb[j][1];
                table[tab1(r, 1, 0)] = vb[j][2];
                table[tab1(r, 1, 1)] = vb[j][3];
                table[tab1(r, 2, 0)] = vb[j][4];
            }
        }
    }
    __syncthreads();

    // phase 2: factors of the unit-diagonal solve.  Pivots p_r = 1 - m_r / p_{r-1},
    // m_r = ld_r ud_{r-1}: the Moebius map p -> (p - m)/p is the 2x2 matrix
    // [[1,-m],[1,0]] on (num, den).  Thread t < NP composes the maps of its RT
    // consecutive rows; an inclusive scan over those threads (Hillis-Steele in
    // each wavefront, then the wavefront totals through LDS) gives every thread
    // the map of all rows before its own, applied to (1,1) the pivot of the row
    // before its first; each thread then runs the exact recurrence over its rows.
    // NP depends on the build's rows only (256 up to 2048 rows, one wavefront more per
    // further 512), so the rounding does not depend on the workgroup size (every kernel
    // that builds a table gives the same table, bit for bit).
    if (d.solve_kind == JW_SOLVE_TRI) {
        // participating threads (the workgroup has at least NP)
        const int NP = RL <= 64 * 4 * kMaxRT ? 256 : (RL + 64 * kMaxRT - 1) / (64 * kMaxRT) * 64;
        // the wavefront totals live in g's array, which nothing reads or writes until the
        // per-row pass below (a barrier keeps that pass after every read): no static LDS,
        // so four 256-thread builds of 1024 rows fit one CU
        T* s_tot = sg;
        const int RT = (RL + NP - 1) / NP;
        const int r0 = tid < NP ? tid * RT : RL;   // threads >= NP own no rows
        const int lane = tid & 63, wv = tid >> 6;
        // the entries are rescaled by a power of two after every product (exact)
        auto normalize = [](T& a, T& b, T& c, T& e) {
            double mx = fmax(fmax(fabs(A::hi(a)), fabs(A::hi(b))), fmax(fabs(A::hi(c)), fabs(A::hi(e))));
            if (mx > 0.0) {
                int ex;
                frexp(mx, &ex);
                a = A::ldexp_(a, -ex); b = A::ldexp_(b, -ex); c = A::ldexp_(c, -ex); e = A::ldexp_(e, -ex);
            }
        };
        // X <- X * P (P: the maps of earlier rows, applied first)
        auto compose = [&](T& a, T& b, T& c, T& e, T pa, T pb, T pc, T pe) {
            T na = A::madd(A::mul(a, pa), b, pc);
            T nb = A::madd(A::mul(a, pb), b, pe);
            T nc = A::madd(A::mul(c, pa), e, pc);
            T ne = A::madd(A::mul(c, pb), e, pe);
            a = na; b = nb; c = nc; e = ne;
            normalize(a, b, c, e);
        };
        T a = A::one(), b = A::zero(), c = A::zero(), e = A::one();
        for (int i = 0; i < RT; i++) {
            const int r = r0 + i;
            if (r >= RL) break;
            const int64_t gr = rb + r;
            T m = A::zero();
            if (gr >= 1 && gr < N) m = A::mul(sl[r], r >= 1 ? sh[r - 1] : (has_carry ? carry[1] : A::zero()));
            T na = A::msub(a, m, c), nb = A::msub(b, m, e);
            c = a;
            e = b;
            a = na;
            b = nb;
            normalize(a, b, c, e);
        }
        // ud of the row before this thread's first, read before any thread
        // overwrites its ud with h (after the barrier below)
        T udprev = (r0 >= 1 && r0 < RL) ? sh[r0 - 1] : ((r0 == 0 && has_carry) ? carry[1] : A::zero());
        // the wavefronts past the first NP threads own no rows: they skip the scan
        // (wave-uniform) and only keep the barriers
        const bool part = wv < NP / 64;
        // inclusive scan inside the wavefront
        T xa = A::one(), xb = A::zero(), xc = A::zero(), xe = A::one();
        if (part) {
            for (int k = 0; k < 6; k++) {
                const int sft = 1 << k;
                const int src = lane >= sft ? lane - sft : lane;
                T pa = lane_fetch(a, src), pb = lane_fetch(b, src), pc = lane_fetch(c, src), pe = lane_fetch(e, src);
                if (lane >= sft) compose(a, b, c, e, pa, pb, pc, pe);
            }
            if (lane == 63) {
                s_tot[4 * wv + 0] = a;
                s_tot[4 * wv + 1] = b;
                s_tot[4 * wv + 2] = c;
                s_tot[4 * wv + 3] = e;
            }
            // this thread's exclusive map within the wavefront (lane - 1's inclusive)
            const int src = lane >= 1 ? lane - 1 : lane;
            xa = lane_fetch(a, src);
            xb = lane_fetch(b, src);
            xc = lane_fetch(c, src);
            xe = lane_fetch(e, src);
            if (lane == 0) {
                xa = A::one(); xb = A::zero(); xc = A::zero(); xe = A::one();
            }
        }
        __syncthreads();   // the wavefront totals are in LDS
        if (part) {
            // the maps of all earlier wavefronts, then the in-wavefront part on top (none of
            // it depends on a carry, so a build that waits for one has it ready)
            T wa = A::one(), wb = A::zero(), wc = A::zero(), we = A::one();
            for (int w = 0; w < wv; w++) {
                T ta = s_tot[4 * w + 0], tb = s_tot[4 * w + 1], tc = s_tot[4 * w + 2], te = s_tot[4 * w + 3];
                compose(ta, tb, tc, te, wa, wb, wc, we);
                wa = ta; wb = tb; wc = tc; we = te;
            }
            compose(xa, xb, xc, xe, wa, wb, wc, we);
        }
        hooks.wait_carry(tid);
        __syncthreads();   // the carry is in LDS, the totals are read: g's array is free
        T gprev = A::one();
        if (part) {
            // g_{r-1} = 1 / p_{r-1} of the row before this thread's first: the prefix map
            // applied to (1, 1) (the first row has no sub-diagonal, so any start works), or
            // after a carry to (1, g_in), i.e. p = 1 / g_in
            if (has_carry) {
                const T gin = carry[0];
                gprev = A::div(A::madd(xc, xe, gin), A::madd(xa, xb, gin));
            } else if (r0 >= 1) {
                gprev = A::div(A::add(xc, xe), A::add(xa, xb));
            }
        }
        // first pass: the exact recurrence over this thread's rows from the scanned pivot;
        // the rows' ld / ud stay in registers (nothing reads another thread's raw ld again)
        T pr[MAXRT], ldv[MAXRT], udv[MAXRT];
        const T udp0 = udprev;
        if (part) {
#pragma unroll
            for (int i = 0; i < MAXRT; i++) {
                pr[i] = A::one();
                ldv[i] = A::zero();
                udv[i] = A::zero();
                const int r = r0 + i;
                if (i >= RT || r >= RL) continue;
                ldv[i] = sl[r];
                udv[i] = sh[r];
                T g = A::zero();
                if (rb + r < N) {
                    // p_r = 1 - ld_r ud_{r-1} / p_{r-1},  g_r = 1/p_r
                    const T l = (rb + r >= 1) ? A::mul(ldv[i], gprev) : A::zero();
                    pr[i] = A::sub(A::one(), A::mul(l, udprev));
                    g = A::div(A::one(), pr[i]);
                    gprev = g;
                }
                udprev = udv[i];
                sg[r] = g;
            }
        }
        // sh[0] becomes the refinement flag: its raw ud is in thread 0's registers by now,
        // and the one other reader (the next thread's udprev) read it before the scan
        if (tid == 0) sh[0] = A::zero();
        __syncthreads();   // every first-pass g is in sg

        // Refinement.  Near a parabolic fixed point of the pivot map (m_r -> 1/4, e.g. the
        // Gram matrices of Jones-Worland connection matrices at small degree) the scanned
        // products lose accuracy like eps k^1.5 over k rows, where the sequential recurrence
        // loses eps k^0.5.  The pivots p~ of the first pass have residuals
        //   rho_r = p~_r - (1 - m_r / p~_{r-1})   (rounding inside a thread, the scan's error
        // at its first row), and their errors e_r = p~_r - p_r obey, to first order,
        //   e_r = rho_r + c_r e_{r-1},  c_r = m_r / p~_{r-1}^2,  e_0 = rho_0 = 0:
        // an affine recurrence, solved by one more scan over the same partition (its own
        // rounding is relative to e, far below eps); p_r = p~_r - e_r.
        // Inside a thread rho is one rounding; the scan's error shows at a thread's first
        // row.  The refinement runs only when some first-row residual exceeds a few ulps
        // (the flag is raised by any thread, read by all: the choice is workgroup-uniform
        // and a function of the matrix alone, so every kernel still builds the same table).
        if (part && r0 >= 1 && rb + r0 < N) {
            const double tol = sizeof(T) == 8 ? 0x1p-51 : 0x1p-103;
            const T rho0 = A::add(A::sub(pr[0], A::one()), A::mul(A::mul(ldv[0], udp0), sg[r0 - 1]));
            if (!(fabs(A::hi(rho0)) <= tol * fabs(A::hi(pr[0])))) sh[0] = A::one();
        }
        __syncthreads();
        const bool refine = A::hi(sh[0]) != 0.0;
        // not refined: every first-pass g is final (in sg since the barrier before the
        // check), so the last one goes to the next build at once
        if (!refine) hooks.publish(tid, sg, RL);
        T rho[MAXRT], cc[MAXRT];
        T sc = A::one(), ss = A::zero();      // this thread's map e -> sc e + ss
        if (refine && part) {
            T gm = (r0 >= 1 && r0 < RL) ? sg[r0 - 1] : ((r0 == 0 && has_carry) ? carry[0] : A::one());
            udprev = udp0;
#pragma unroll
            for (int i = 0; i < MAXRT; i++) {
                rho[i] = A::zero();
                cc[i] = A::zero();
                const int r = r0 + i;
                if (i >= RT || rb + r >= N) continue;
                const T m = (rb + r >= 1) ? A::mul(ldv[i], udprev) : A::zero();
                const T mg = A::mul(m, gm);
                rho[i] = A::add(A::sub(pr[i], A::one()), mg);
                cc[i] = A::mul(mg, gm);
                ss = A::madd(rho[i], cc[i], ss);
                sc = A::mul(cc[i], sc);
                gm = sg[r];
                udprev = udv[i];
            }
            // inclusive scan of the maps inside the wavefront, wavefront totals to LDS
            // (sl's first 8 values: every thread holds its raw ld in registers by now)
            for (int k = 0; k < 6; k++) {
                const int sft = 1 << k;
                const int src = lane >= sft ? lane - sft : lane;
                const T pc = lane_fetch(sc, src), ps = lane_fetch(ss, src);
                if (lane >= sft) {
                    ss = A::madd(ss, sc, ps);
                    sc = A::mul(sc, pc);
                }
            }
            if (lane == 63) {
                sl[2 * wv] = sc;
                sl[2 * wv + 1] = ss;
            }
            const int src = lane >= 1 ? lane - 1 : lane;
            const T xs = lane_fetch(ss, src), xc = lane_fetch(sc, src);
            ss = lane == 0 ? A::zero() : xs;
            sc = lane == 0 ? A::one() : xc;
        }
        if (refine) __syncthreads();
        if (refine && part) {
            // error entering this thread's first row: earlier wavefronts' maps, then this
            // wavefront's exclusive part, applied to e = 0
            T ein = A::zero();
            for (int w = 0; w < wv; w++) ein = A::madd(sl[2 * w + 1], sl[2 * w], ein);
            ein = A::madd(ss, sc, ein);
#pragma unroll
            for (int i = 0; i < MAXRT; i++) {
                const int r = r0 + i;
                if (i >= RT || rb + r >= N) continue;
                ein = A::madd(rho[i], cc[i], ein);
                // a correction that is not finite (no dominance, products out of range)
                // is not applied
                const double eh = A::hi(ein);
                const T p = (eh - eh == 0.0) ? A::sub(pr[i], ein) : pr[i];
                sg[r] = A::div(A::one(), p);
            }
        }
        // refined: the corrected g of every row (and of the row before) is in sg; not
        // refined: every thread has read the flag before thread 0 writes h_0 over it
        __syncthreads();
        if (refine) hooks.publish(tid, sg, RL);
        if (part) {
#pragma unroll
            for (int i = 0; i < MAXRT; i++) {
                const int r = r0 + i;
                if (i >= RT || r >= RL) continue;
                // l_r = ld_r g_{r-1},  h_r = ud_r g_r  (zero outside the matrix)
                T l = A::zero(), h = A::zero();
                if (rb + r < N) {
                    if (rb + r >= 1) l = A::mul(ldv[i], r >= 1 ? sg[r - 1] : (has_carry ? carry[0] : A::zero()));
                    h = A::mul(udv[i], sg[r]);
                }
                sl[r] = l;
                sh[r] = h;
            }
        }
    } else {
        for (int r = tid; r < RL; r += nthr) {
            const bool in = rb + r < N;
            if (!(in && d.solve_kind == JW_SOLVE_LOWER)) sl[r] = A::zero();
            sg[r] = in ? A::one() : A::zero();
            if (!(in && d.solve_kind == JW_SOLVE_UPPER)) sh[r] = A::zero();
        }
    }
    __syncthreads();

    if (format == 2) {
        // nsub tables of 64 R rows each, JW_TAB64_ELEMS(R) apart; in a build of a long
        // system's rows the first row's l' and the last row's h' take a and D of the rows
        // next to the build (g_{rb-1} from the carry of a tridiagonal solve, else 1)
        const int64_t te2 = JW_TAB64_ELEMS(R);
        const T aprev = A::mul(dprev, has_carry ? carry[0] : A::one());
        // fold the scalings (row-parallel); l' and h' replace l and h in LDS (a row
        // reads only its own l and h)
#pragma unroll
        for (int j = 0; j < MAXJ; j++) {
            const int r = row_of(j);
            if (r >= RL) continue;
            const int kt = r / RLs, rr = r - kt * RLs;
            const int t = rr / R, i = rr % R;
            const T ar = A::mul(sD[r], sg[r]);
            T cf[8];
#pragma unroll
            for (int k = 0; k < 5; k++) cf[k] = A::mul(ar, vb[j][k]);
            T lp = A::zero(), hp = A::zero();
            if (rb + r >= 1 && rb + r < N)
                lp = A::div(A::mul(ar, sl[r]), r >= 1 ? A::mul(sD[r - 1], sg[r - 1]) : aprev);
            if (rb + r + 1 < N) hp = A::div(A::mul(sD[r], sh[r]), r + 1 < RL ? sD[r + 1] : dnext);
            cf[5] = lp;
            cf[6] = hp;
            cf[7] = sE[r];
            T* tk = table + kt * te2;
#pragma unroll
            for (int p = 0; p < 4; p++) {
                tk[tab_coef64(p, i, t, R)] = cf[2 * p];
                tk[tab_coef64(p, i, t, R) + 1] = cf[2 * p + 1];
            }
            sl[r] = lp;
            sh[r] = hp;
        }
        __syncthreads();
        // scan multipliers: wavefront item (wave-uniform) takes recurrence j = item % 3
        // (0: forward through l', 1: backward through h', 2: backward through E) of
        // table item / 3
        for (int item = tid >> 6; item < 3 * nsub; item += nthr >> 6) {
            const int kt = item / 3, j = item - 3 * kt;
            const int t = tid & 63;
            T* const tabk = table + kt * te2;
            const T* src = (j == 0 ? sl : (j == 1 ? sh : sE)) + kt * RLs;
            T v = A::one();
            for (int i = 0; i < R; i++) v = A::mul(v, A::neg(src[t * R + i]));
            // in-row Hillis-Steele: before step k the value covers the 2^k lanes a
            // shift of 2^k combines; after 4 steps, the row prefix (j = 0) or suffix
            const int tr = t & 15;
            const int64_t SC = (int64_t)8 * R * 64;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int s = 1 << k;
                tabk[SC + (int64_t)(j * 8 + k) * 64 + t] = v;
                if (j == 0) {
                    const T up = lane_fetch(v, tr >= s ? t - s : t);
                    if (tr >= s) v = A::mul(v, up);
                } else {
                    const T dn = lane_fetch(v, tr + s <= 15 ? t + s : t);
                    if (tr + s <= 15) v = A::mul(v, dn);
                }
            }
            const int rw = t >> 4;
            if (j == 0) {
                // forward cross-row combine in two broadcasts: rows 1,3 take the end of
                // rows 0,2 (multiplier: row prefix), then rows 2,3 take lane 31, which
                // row 3 reaches through all of row 2 (prefix of lane 47 times its own)
                const T pre47 = lane_fetch(v, 47);
                T m2 = A::zero();
                if (rw == 2) m2 = v;
                if (rw == 3) m2 = A::mul(pre47, v);
                tabk[SC + (int64_t)4 * 64 + t] = v;
                tabk[SC + (int64_t)5 * 64 + t] = m2;
                tabk[SC + (int64_t)6 * 64 + t] = A::zero();
                tabk[SC + (int64_t)7 * 64 + t] = A::zero();
            } else {
                // backward cross-row combine from the in-row results S_1..S_3 at the
                // row starts (lanes 16, 32, 48): a lane of row r adds
                //   suf_t * F_{r+1},  F_q = S_q + P_q F_{q+1},  P_q = suffix product of lane 16q,
                // i.e. c_k,t * S_k with c_k,t = suf_t * prod_{r<q<k} P_q for k > r, else 0
                const T P1 = lane_fetch(v, 16), P2 = lane_fetch(v, 32);
                T c1 = A::zero(), c2 = A::zero(), c3 = A::zero();
                if (rw == 0) { c1 = v; c2 = A::mul(v, P1); c3 = A::mul(c2, P2); }
                if (rw == 1) { c2 = v; c3 = A::mul(v, P2); }
                if (rw == 2) c3 = v;
                tabk[SC + (int64_t)(j * 8 + 4) * 64 + t] = c1;
                tabk[SC + (int64_t)(j * 8 + 5) * 64 + t] = c2;
                tabk[SC + (int64_t)(j * 8 + 6) * 64 + t] = c3;
                tabk[SC + (int64_t)(j * 8 + 7) * 64 + t] = A::zero();
            }
        }
        return;
    }

    // format 1: pairs (vm2,vm1) (v0,vp1) (vp2,l) (g,h) (D,E) + Hillis-Steele multipliers
    // (the band pairs were written in phase 1)
    for (int r = tid; r < RL; r += nthr) {
        table[tab1(r, 2, 1)] = sl[r];
        table[tab1(r, 3, 0)] = sg[r];
        table[tab1(r, 3, 1)] = sh[r];
        table[tab1(r, 4, 0)] = sD[r];
        table[tab1(r, 4, 1)] = sE[r];
    }
    // scan multipliers: wavefront k (wave-uniform) takes table k
    if ((tid >> 6) < nsub) {
        const int kt = tid >> 6;
        const int t = tid & 63;
        T* const tabk = table + kt * tab_elems;
        const int rbase = kt * RLs;
        const bool active = t < L;
        const int log2L = ilog2_rt(L);
        T bf = A::one(), bb1 = A::one(), bb2 = A::one();
        if (active) {
            for (int i = 0; i < R; i++) {
                const int r = rbase + t * R + i;
                bf = A::mul(bf, A::neg(sl[r]));
                bb1 = A::mul(bb1, A::neg(sh[r]));
                bb2 = A::mul(bb2, A::neg(sE[r]));
            }
        }
        for (int k = 0; k < log2L; k++) {
            const int s = 1 << k;
            if (active) {
                tabk[tab_scan(k, t, 0, R, L)] = bf;
                tabk[tab_scan(k, t, 1, R, L)] = bb1;
                tabk[tab_scan(k, t, 2, R, L)] = bb2;
                tabk[tab_scan(k, t, 3, R, L)] = A::zero();
            }
            const int su = (t >= s) ? t - s : t;
            const int sdn = (t + s < L) ? t + s : t;
            T ubf = lane_fetch(bf, su);
            T ub1 = lane_fetch(bb1, sdn);
            T ub2 = lane_fetch(bb2, sdn);
            if (active && t >= s) bf = A::mul(bf, ubf);
            if (active && t + s < L) {
                bb1 = A::mul(bb1, ub1);
                bb2 = A::mul(bb2, ub2);
            }
        }
    }
}

// Pivot hand-over between the table builds of a long system (tridiagonal solves): build b
// waits for build b-1's last pivot (flags[b-1], acquire) and publishes its own (release).
// The wait is bounded: when it runs out the device error flag is raised and the build goes
// on with a unit pivot, so the grid always drains.
constexpr int kLongMaxPolls = 1 << 22;

template <typename T>
struct ChainHooks {
    T* carry;        // LDS: carry[0] = g entering the build
    int* flags;      // per build: 1 once its last pivot is in gval
    T* gval;         // per build: g of its last row
    int* err;
    int64_t build;
    int chained;
    __device__ void wait_carry(int tid) const
    {
        if (!chained || build == 0 || tid != 0) return;
        int polls = 0;
        bool ok = true;
        while (__hip_atomic_load(&flags[build - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (++polls > kLongMaxPolls) {
                ok = false;
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        carry[0] = ok ? gval[build - 1] : Ar<T>::one();
    }
    __device__ void publish(int tid, const T* sg, int RL) const
    {
        if (!chained || tid != 0) return;
        gval[build] = sg[RL - 1];
        __hip_atomic_store(&flags[build], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
};

#endif  // PFSOLVE_JW_PROLOGUE_H
