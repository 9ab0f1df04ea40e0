// jw_long.hip -- long systems of the JW step on gfx950: FP64 N > 2048 and double-double
// N > 512 (up to 2^22 rows).
//
// The reference cuts such systems into Msplit0-row pieces, reduces every piece to its
// boundary rows in the `kernel` scratch, solves the reduced system and back-substitutes
// over several launches (PfSolve/pfSolve_AppManagement/pfSolve_InitializeApp.h:1540-1580,
// PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:2014-2197,
// PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_TridiagonalSolve.h:1795-1856,
// PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:138-157).  The storage of the matrices
// (Msplit0-row blocks of reorder_i0) is the caller's and is read as it is; the solve is cut
// differently, into chunks of C = 64 R rows, one wavefront per chunk and pair of systems:
//
//   table    every chunk's format-1 coefficient table.  The LU pivots continue from
//            chunk to chunk with the exact recurrence (one workgroup, 2048 or 1024 rows
//            per step, the pivot of the last row carried into the next step); without
//            pivots (bidiagonal solves) every step is its own workgroup.
//   COEF     per chunk, the response of its boundary values to a unit carry: w entering
//            from the row before, z and y entering from the row after (3 virtual systems,
//            zero right-hand side).
//   REDUCE   per chunk and system, the boundary values with zero carries (w at the last
//            row, z and y at the first), and the right-hand side rows the neighbouring
//            chunks' banded matvec needs (the step runs in place).
//   reduced  per system, the carries of every chunk: w forward over the chunks, then z
//            and y backward (each chunk's boundary values are linear in its carries).
//   FINISH   per chunk and system, the chunk again from its carries, stored.
//
// Inside a chunk the work is the general kernel's: banded matvec from neighbour lanes,
// forward and two backward first-order recurrences as lane sweeps plus a 6-step affine
// scan across the 64 lanes.  The right-hand sides are read twice and written once.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "jw_common.h"
#include "jw_prologue.h"

extern "C" void jwdev_set_error(const char* msg);

namespace {

constexpr int kLongThreads = 256;   // 4 wavefronts per chunk workgroup
constexpr int kLongL = 64;
constexpr int kLongLog2L = 6;

// rows per lane of a chunk: FP64 8 (512-row chunks, 52 KiB table in LDS), double-double 4
// (256-row chunks, 64 KiB); NSUB chunks per table build; BOUND: the thread count the chunk
// kernel's register budget is declared for (768: 168 VGPRs, three 256-thread workgroups per
// CU, as many as the FP64 table's LDS allows; 512: two, as for the DD table)
template <typename T> struct LongGeom;
template <> struct LongGeom<double> {
    static constexpr int R = 8, NSUB = 4, BOUND = 768;
};
template <> struct LongGeom<dd> {
    static constexpr int R = 4, NSUB = 4, BOUND = 512;
};

template <typename T>
constexpr int64_t long_tab_elems()
{
    return (int64_t)JW_NCOEF * LongGeom<T>::R * kLongL + 4 * kLongLog2L * kLongL;
}

}  // namespace

// ------------------------------------------------------------------------
// tables: NSUB chunks per build of NSUB * 64 * R rows, one workgroup per build.
// Tridiagonal solves hand the pivot on: build b (taken in dispatch order from a ticket, so
// every build it waits for has started) runs its loads, its pivot-map scan and everything
// else that needs no pivot, then waits for build b-1's last pivot, finishes its pivots by
// the exact recurrence and publishes its own last pivot before it writes its tables.  The
// wait is bounded: when it runs out the device error flag is raised (PfSolveSync reports
// it) and the build goes on with a unit pivot, so the grid always drains.
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kPrologueThreads) void jw_long_prologue_kernel(jw_source_desc d, const void* mat,
                                                                          T* table, int chained, int* sync, T* gval,
                                                                          int* err, int format)
{
    extern __shared__ __align__(16) unsigned char pro_smem[];
    __shared__ int s_build;
    constexpr int R = LongGeom<T>::R;
    constexpr int NSUB = LongGeom<T>::NSUB;
    constexpr int SR = NSUB * R * kLongL;
    // format 1 for the chunk kernel here, format 2 (FP64) for the pair kernel's chunks
    const int64_t TE = format == 2 ? (int64_t)JW_TAB64_ELEMS(R) : long_tab_elems<T>();
    // the pivot carry follows the build's five row arrays
    T* carry = reinterpret_cast<T*>(pro_smem + prologue_lds_bytes(SR, sizeof(T)));
    int64_t build = blockIdx.x;
    if (chained) {
        if (threadIdx.x == 0) s_build = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        build = s_build;
    }
    const int64_t rb = build * SR;
    // the raw super-diagonal entry of the row before the build (the build's first barrier
    // publishes it)
    if (chained && threadIdx.x == 0)
        carry[1] = rb > 0 ? mem_load<T>(mat, d.udS + source_pos(d, rb - 1, d.N), d.mem_f64) : Ar<T>::zero();
    const ChainHooks<T> hooks{carry, sync + 1, gval, err, build, chained};
    prologue_body<T, ChainHooks<T>>(d, mat, table + build * NSUB * TE, kLongL, R, format, pro_smem, rb,
                                    chained ? carry : nullptr, NSUB, hooks);
}

__global__ void jw_long_clear_kernel(int* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------------------
// chunk kernel: grid (workgroups per chunk, chunks)
// ------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(LongGeom<T>::BOUND) void jw_long_chunk_kernel(jw_long_desc d)
{
    using A = Ar<T>;
    constexpr int R = LongGeom<T>::R;
    constexpr int L = kLongL;
    constexpr int LOG2L = kLongLog2L;
    constexpr int C = R * L;
    constexpr int NTAB = (int)long_tab_elems<T>();
    static_assert(NTAB * (int)sizeof(T) % 16 == 0, "the table is whole 16-byte chunks");
    extern __shared__ __align__(16) unsigned char jw_smem[];
    const T* tab = reinterpret_cast<const T*>(jw_smem);

    const int64_t c = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    copy_to_lds(jw_smem, reinterpret_cast<const T*>(d.table) + c * d.tab_elems, NTAB * (int)sizeof(T), wave, nwaves,
                lane);
    __syncthreads();

    const jw_run_desc& rd = d.run;
    const int t = lane;
    const int64_t rb = c * C;
    const int64_t r0 = rb + (int64_t)t * R;
    constexpr int mode = MODE;   // one kernel per mode: no mode branches at run time
    const int mf = rd.mem_f64;
    const int64_t nc = d.nchunks;
    const int64_t nsys = mode == JW_LONG_COEF ? 3 : rd.nsys;
    const int64_t npass = (nsys + 1) / 2;
    T* red = reinterpret_cast<T*>(d.red);
    const int64_t rs = d.red_sys;
    auto at = [&](int a, int64_t cc, int64_t s) -> int64_t { return ((int64_t)a * rs + s) * nc + cc; };
    T* coef = red + (int64_t)JW_LONG_ARRAYS * nc * rs;

    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        const int64_t sys0 = 2 * q;
        T x[2][R];
        // the rows next to the chunk (lane 0: rows rb-1, rb-2; lane 63: rows rb+C, rb+C+1)
        // and the carries (lane 0: w entering; lane 63: z and y entering)
        T h1[2], h2[2], ca[2], cb[2];

        // ---- load: rows outside [in_lo, in_hi) (in_hi <= N) and absent systems read as 0.
        // A system whose chunk lies inside the keep-range (the common case; a wavefront-
        // uniform test) is loaded straight, every load issued before the first use; the
        // others row by row.  Guarded loads start from a zero already in the register, so
        // no path overwrites a register a load is still filling (which would make the
        // compiler wait for every load in flight). ----
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            const bool ok = sys < nsys && mode != JW_LONG_COEF;
            const bool full = ok && rb >= rd.in_lo && rb + C <= rd.in_hi;
            const int64_t base = sys * rd.stride + rd.offSol;
            if (full) {
                if constexpr (sizeof(T) == 8) {
                    const double* pp = reinterpret_cast<const double*>(rd.rhs) + base + r0;
                    if (rd.vec2) {
#pragma unroll
                        for (int i = 0; i < R; i += 2) {
                            const double2 v = *reinterpret_cast<const double2*>(pp + i);
                            x[s][i] = v.x;
                            x[s][i + 1] = v.y;
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < R; i++) x[s][i] = pp[i];
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < R; i++) x[s][i] = mem_load<T>(rd.rhs, base + r0 + i, mf);
                }
            } else {
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const int64_t gr = r0 + i;
                    x[s][i] = A::zero();
                    if (ok && gr >= rd.in_lo && gr < rd.in_hi) x[s][i] = mem_load<T>(rd.rhs, base + gr, mf);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            const bool ok = sys < nsys && mode != JW_LONG_COEF;
            const int64_t base = sys * rd.stride + rd.offSol;
            h1[s] = A::zero();
            h2[s] = A::zero();
            if (rd.band && ok) {
                if (mode == JW_LONG_REDUCE) {
                    // the neighbouring chunks' rows, before any of them is written (the
                    // finishing launch takes them from the reduced data instead): lane 0
                    // rows rb-1, rb-2, lane 63 rows rb+C, rb+C+1
                    const int64_t g1 = t == 0 ? rb - 1 : rb + C;
                    const int64_t g2 = t == 0 ? rb - 2 : rb + C + 1;
                    const bool edge = t == 0 || t == L - 1;
                    if (edge && g1 >= rd.in_lo && g1 < rd.in_hi) h1[s] = mem_load<T>(rd.rhs, base + g1, mf);
                    if (edge && g2 >= rd.in_lo && g2 < rd.in_hi) h2[s] = mem_load<T>(rd.rhs, base + g2, mf);
                } else {
                    const bool lo = t == 0 && c > 0, hi = t == L - 1 && c + 1 < nc;
                    const int64_t i1 = lo ? at(6, c - 1, sys) : at(3, c + 1, sys);
                    const int64_t i2 = lo ? at(5, c - 1, sys) : at(4, c + 1, sys);
                    if (lo || hi) {
                        h1[s] = red[i1];
                        h2[s] = red[i2];
                    }
                }
            }
            ca[s] = A::zero();
            cb[s] = A::zero();
            if (mode == JW_LONG_COEF) {
                if (t == 0 && sys == 0) ca[s] = A::one();
                if (t == L - 1 && sys == 1) ca[s] = A::one();
                if (t == L - 1 && sys == 2) cb[s] = A::one();
            } else if (mode == JW_LONG_FINISH && sys < nsys) {
                // lane 0: w entering; lane 63: z and y entering
                if (t == 0 || t == L - 1) ca[s] = red[at(t == 0 ? 7 : 8, c, sys)];
                if (t == L - 1) cb[s] = red[at(9, c, sys)];
            }
            if (mode == JW_LONG_REDUCE && sys < nsys) {
                if (t == 0) {
                    red[at(3, c, sys)] = x[s][0];
                    red[at(4, c, sys)] = x[s][1];
                }
                if (t == L - 1) {
                    red[at(5, c, sys)] = x[s][R - 2];
                    red[at(6, c, sys)] = x[s][R - 1];
                }
            }
        }

        // ---- banded matvec u = V x (in place) ----
        if (rd.band) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T xm2 = A::zero(), xm1 = A::zero(), xp1 = A::zero(), xp2 = A::zero();
                if (rd.kl >= 1) xm1 = A::sel(t >= 1, lane_fetch(x[s][R - 1], t >= 1 ? lane - 1 : lane), h1[s]);
                if (rd.kl >= 2) xm2 = A::sel(t >= 1, lane_fetch(x[s][R - 2], t >= 1 ? lane - 1 : lane), h2[s]);
                if (rd.ku >= 1) xp1 = A::sel(t + 1 < L, lane_fetch(x[s][0], t + 1 < L ? lane + 1 : lane), h1[s]);
                if (rd.ku >= 2) xp2 = A::sel(t + 1 < L, lane_fetch(x[s][1], t + 1 < L ? lane + 1 : lane), h2[s]);
                T pm2 = xm2, pm1 = xm1;
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const T vm2 = tab[tab_coef(0, i, t, 0, R, L)], vm1 = tab[tab_coef(0, i, t, 1, R, L)];
                    const T v0 = tab[tab_coef(1, i, t, 0, R, L)], vp1 = tab[tab_coef(1, i, t, 1, R, L)];
                    const T vp2 = tab[tab_coef(2, i, t, 0, R, L)];
                    const T xi = x[s][i];
                    const T n1 = (i + 1 < R) ? x[s][(i + 1) % R] : xp1;
                    const T n2 = (i + 2 < R) ? x[s][(i + 2) % R] : ((i + 2 == R) ? xp1 : xp2);
                    T u = A::mul(v0, xi);
                    u = A::madd(u, vm1, pm1);
                    u = A::madd(u, vm2, pm2);
                    u = A::madd(u, vp1, n1);
                    u = A::madd(u, vp2, n2);
                    pm2 = pm1;
                    pm1 = xi;
                    x[s][i] = u;
                }
            }
        }

        // ---- forward recurrence w_r = u_r - l_r w_{r-1}, w entering lane 0 = cw ----
        if (rd.fwd) {
            T lr[R];
#pragma unroll
            for (int i = 0; i < R; i++) lr[i] = tab[tab_coef(2, i, t, 1, R, L)];
            T acc[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T a = ca[s];   // zero on every lane but lane 0
#pragma unroll
                for (int i = 0; i < R; i++) a = A::msub(x[s][i], lr[i], a);
                acc[s] = a;
            }
#pragma unroll
            for (int k = 0; k < LOG2L; k++) {
                const int sft = 1 << k;
                const T bk = tab[tab_scan(k, t, 0, R, L)];
                const int src = t >= sft ? lane - sft : lane;
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const T ap = lane_fetch(acc[s], src);
                    acc[s] = A::sel(t >= sft, A::madd(acc[s], bk, ap), acc[s]);
                }
            }
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const T cprev = lane_fetch(acc[s], t >= 1 ? lane - 1 : lane);
                T w = A::sel(t >= 1, cprev, ca[s]);
#pragma unroll
                for (int i = 0; i < R; i++) {
                    w = A::msub(x[s][i], lr[i], w);
                    x[s][i] = w;
                }
            }
        }
        T wl[2] = {x[0][R - 1], x[1][R - 1]};   // w of the chunk's last row (lane 63)

        // ---- backward recurrences z_r = g_r w_r - h_r z_{r+1} (z entering lane 63 = cz),
        //      then y_r = D_r z_r - E_r y_{r+1} (y entering = cy) ----
        T zf[2] = {x[0][0], x[1][0]};
#pragma unroll
        for (int stage = 0; stage < 2; stage++) {
            const bool on = stage == 0 ? rd.b1 : rd.b2;
            if (on) {
                const int p = stage == 0 ? 3 : 4;
                T gg[R], hh[R];
#pragma unroll
                for (int i = 0; i < R; i++) {
                    gg[i] = tab[tab_coef(p, i, t, 0, R, L)];
                    hh[i] = tab[tab_coef(p, i, t, 1, R, L)];
                }
                T cin[2];   // zero on every lane but lane 63
#pragma unroll
                for (int s = 0; s < 2; s++) cin[s] = A::sel(t == L - 1, stage == 0 ? ca[s] : cb[s], A::zero());
                T acc[2];
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    T a = cin[s];
#pragma unroll
                    for (int i = R - 1; i >= 0; i--) {
                        x[s][i] = A::mul(gg[i], x[s][i]);
                        a = A::msub(x[s][i], hh[i], a);
                    }
                    acc[s] = a;
                }
#pragma unroll
                for (int k = 0; k < LOG2L; k++) {
                    const int sft = 1 << k;
                    const T bk = tab[tab_scan(k, t, 1 + stage, R, L)];
                    const int src = t + sft < L ? lane + sft : lane;
#pragma unroll
                    for (int s = 0; s < 2; s++) {
                        const T ap = lane_fetch(acc[s], src);
                        acc[s] = A::sel(t + sft < L, A::madd(acc[s], bk, ap), acc[s]);
                    }
                }
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const T cnext = lane_fetch(acc[s], t + 1 < L ? lane + 1 : lane);
                    T z = A::sel(t + 1 < L, cnext, cin[s]);
#pragma unroll
                    for (int i = R - 1; i >= 0; i--) {
                        z = A::msub(x[s][i], hh[i], z);
                        x[s][i] = z;
                    }
                }
            }
            if (stage == 0) {
                zf[0] = x[0][0];
                zf[1] = x[1][0];
            }
        }

        if (mode == JW_LONG_FINISH) {
            // ---- row-0 scaleC (pfSolve_TridiagonalSolve.h:1918-1927) ----
            if (rd.apply_scale && c == 0 && t == 0) {
                const T sc = A::from(rd.scale_hi, rd.scale_lo);
#pragma unroll
                for (int s = 0; s < 2; s++) x[s][0] = A::mul(x[s][0], sc);
            }
            // ---- store rows in [out_lo, out_hi) (out_hi <= N): straight when the chunk
            // lies inside the range (wavefront-uniform), else row by row ----
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int64_t sys = sys0 + s;
                if (sys >= nsys) continue;
                const int64_t base = sys * rd.stride + rd.offSol;
                const bool full = rb >= rd.out_lo && rb + C <= rd.out_hi;
                if (full) {
                    if constexpr (sizeof(T) == 8) {
                        double* pp = reinterpret_cast<double*>(rd.rhs) + base + r0;
                        if (rd.vec2) {
#pragma unroll
                            for (int i = 0; i < R; i += 2) {
                                double2 v;
                                v.x = x[s][i];
                                v.y = x[s][i + 1];
                                *reinterpret_cast<double2*>(pp + i) = v;
                            }
                        } else {
#pragma unroll
                            for (int i = 0; i < R; i++) pp[i] = x[s][i];
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < R; i++) mem_store<T>(rd.rhs, base + r0 + i, x[s][i], mf);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < R; i++) {
                        const int64_t gr = r0 + i;
                        if (gr >= rd.out_lo && gr < rd.out_hi) mem_store<T>(rd.rhs, base + gr, x[s][i], mf);
                    }
                }
            }
        } else {
            // ---- boundary values: w at the last row, z and y at the first ----
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int64_t sys = sys0 + s;
                if (sys >= nsys) continue;
                if (mode == JW_LONG_COEF) {
                    if (t == L - 1) coef[9 * c + 3 * sys + 0] = wl[s];
                    if (t == 0) {
                        coef[9 * c + 3 * sys + 1] = zf[s];
                        coef[9 * c + 3 * sys + 2] = x[s][0];
                    }
                } else {
                    if (t == L - 1) red[at(0, c, sys)] = wl[s];
                    if (t == 0) {
                        red[at(1, c, sys)] = zf[s];
                        red[at(2, c, sys)] = x[s][0];
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// reduced solve: one wavefront per system, lanes over 64 chunks at a time.  Forward: the
// maps w -> K_w(c) w + F(c) of the chunks, composed by an inclusive scan across the lanes;
// backward: the maps of (z, y), [[psi, 0], [zeta, omega]] (z, y) + (Z + chi w_in, Y + xi w_in),
// composed from the last chunk down.  The data of one system are contiguous per array.
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void jw_long_reduced_kernel(jw_long_desc d)
{
    using A = Ar<T>;
    const int lane = threadIdx.x & 63;
    const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (s >= d.run.nsys) return;   // wavefront-uniform
    const int64_t nc = d.nchunks, rs = d.red_sys;
    T* red = reinterpret_cast<T*>(d.red);
    const T* F = red + (0 * rs + s) * nc;
    const T* Z = red + (1 * rs + s) * nc;
    const T* Y = red + (2 * rs + s) * nc;
    T* Win = red + (7 * rs + s) * nc;
    T* Zin = red + (8 * rs + s) * nc;
    T* Yin = red + (9 * rs + s) * nc;
    const T* K = red + (int64_t)JW_LONG_ARRAYS * nc * rs;

    // forward: w entering chunk c
    T w = A::zero();
    for (int64_t c0 = 0; c0 < nc; c0 += 64) {
        const int64_t c = c0 + lane;
        const bool on = c < nc;
        T a = on ? K[9 * c] : A::one();
        T b = on ? F[c] : A::zero();
        // inclusive: lane j holds the map of chunks c0..c0+j (later applied after earlier)
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int sft = 1 << k;
            const int src = lane >= sft ? lane - sft : lane;
            const T pa = lane_fetch(a, src), pb = lane_fetch(b, src);
            if (lane >= sft) {
                b = A::madd(b, a, pb);
                a = A::mul(a, pa);
            }
        }
        // w leaving chunk c = a w + b; entering chunk c = leaving chunk c - 1
        const T wout = A::madd(b, a, w);
        const T prev = lane_fetch(wout, lane >= 1 ? lane - 1 : lane);
        if (on) Win[c] = lane >= 1 ? prev : w;
        w = lane_fetch(wout, 63);
    }
    // backward: (z, y) entering chunk c from chunk c + 1
    T z = A::zero(), y = A::zero();
    for (int64_t c1 = nc - 1; c1 >= 0; c1 -= 64) {
        const int64_t c = c1 - lane;
        const bool on = c >= 0;
        // map of chunk c: (z, y) -> (p z + bz, q z + r y + by)
        T p = A::one(), q = A::zero(), r = A::one(), bz = A::zero(), by = A::zero();
        if (on) {
            const T wi = Win[c];
            p = K[9 * c + 4];
            q = K[9 * c + 5];
            r = K[9 * c + 8];
            bz = A::madd(Z[c], K[9 * c + 1], wi);
            by = A::madd(Y[c], K[9 * c + 2], wi);
        }
        // inclusive over chunks c1 .. c (applied in that order)
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int sft = 1 << k;
            const int src = lane >= sft ? lane - sft : lane;
            const T pp = lane_fetch(p, src), pq = lane_fetch(q, src), pr = lane_fetch(r, src);
            const T pbz = lane_fetch(bz, src), pby = lane_fetch(by, src);
            if (lane >= sft) {
                // this map after the earlier one: z'' = p (pp z + pbz) + bz,
                // y'' = q (pp z + pbz) + r (pq z + pr y + pby) + by
                const T nq = A::madd(A::mul(q, pp), r, pq);
                by = A::madd(A::madd(by, q, pbz), r, pby);
                bz = A::madd(bz, p, pbz);
                q = nq;
                p = A::mul(p, pp);
                r = A::mul(r, pr);
            }
        }
        const T zout = A::madd(bz, p, z);
        const T yout = A::madd(A::madd(by, q, z), r, y);
        const T pz = lane_fetch(zout, lane >= 1 ? lane - 1 : lane);
        const T py = lane_fetch(yout, lane >= 1 ? lane - 1 : lane);
        if (on) {
            Zin[c] = lane >= 1 ? pz : z;
            Yin[c] = lane >= 1 ? py : y;
        }
        z = lane_fetch(zout, 63);
        y = lane_fetch(yout, 63);
    }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
namespace {
char g_msg[256];
int fail(const char* what, hipError_t e)
{
    snprintf(g_msg, sizeof(g_msg), "%s: %s", what, hipGetErrorString(e));
    jwdev_set_error(g_msg);
    return -1;
}
int fail(const char* what)
{
    jwdev_set_error(what);
    return -1;
}
template <typename T> size_t chunk_lds() { return (size_t)long_tab_elems<T>() * sizeof(T); }
template <typename T> size_t prologue_lds()
{
    constexpr int SR = LongGeom<T>::NSUB * LongGeom<T>::R * kLongL;
    return prologue_lds_bytes(SR, sizeof(T)) + 2 * sizeof(T);
}
template <typename T>
const void* chunk_fn(int mode)
{
    switch (mode) {
    case JW_LONG_COEF: return reinterpret_cast<const void*>(&jw_long_chunk_kernel<T, JW_LONG_COEF>);
    case JW_LONG_REDUCE: return reinterpret_cast<const void*>(&jw_long_chunk_kernel<T, JW_LONG_REDUCE>);
    default: return reinterpret_cast<const void*>(&jw_long_chunk_kernel<T, JW_LONG_FINISH>);
    }
}

}  // namespace

extern "C" {

int jwdev_long_rows(int precision) { return precision == JW_PREC_F64 ? LongGeom<double>::R : LongGeom<dd>::R; }
int jwdev_long_nsub(int precision) { return precision == JW_PREC_F64 ? LongGeom<double>::NSUB : LongGeom<dd>::NSUB; }

int64_t jwdev_long_lds_bytes(int precision)
{
    return (int64_t)(precision == JW_PREC_F64 ? chunk_lds<double>() : chunk_lds<dd>());
}

int jwdev_long_occupancy(int precision)
{
    const void* f = precision == JW_PREC_F64 ? chunk_fn<double>(JW_LONG_FINISH) : chunk_fn<dd>(JW_LONG_FINISH);
    const size_t lds = (size_t)jwdev_long_lds_bytes(precision);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { fail("hipFuncSetAttribute(long chunks)", e); return 0; }
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kLongThreads, lds);
    if (e != hipSuccess) { fail("occupancy(long chunks)", e); return 0; }
    return nb;
}

int jwdev_long_prologue(const jw_source_desc* src, const void* matrices, void* table, int64_t tab_elems,
                        int precision, int* sync, void* gval, int* err, void* stream)
{
    if (!matrices || !table) return fail("long prologue: no matrices or table");
    const bool f64 = precision == JW_PREC_F64;
    // the table size names the format: format 1 (the chunk kernel), or FP64 format 2 (the
    // pair kernel's chunks, jw_long64_kernel)
    int format = 1;
    if (f64 && tab_elems == (int64_t)JW_TAB64_ELEMS(LongGeom<double>::R))
        format = 2;
    else if (tab_elems != (f64 ? long_tab_elems<double>() : long_tab_elems<dd>()))
        return fail("long prologue: table size mismatch");
    const int SR = f64 ? LongGeom<double>::NSUB * LongGeom<double>::R * kLongL
                       : LongGeom<dd>::NSUB * LongGeom<dd>::R * kLongL;
    const int64_t nit = (src->N + SR - 1) / SR;
    if (nit < 1) return 0;
    // pivots continue from one build to the next (tridiagonal solves): the ticket and the
    // builds' flags start at zero (a fill kernel, so that the launch can be captured)
    const int chained = src->solve_kind == JW_SOLVE_TRI;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (chained) {
        if (!sync || !gval || !err) return fail("long prologue: no hand-over scratch");
        const int n = (int)nit + 1;
        hipLaunchKernelGGL(jw_long_clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sync, n);
        hipError_t ec = hipGetLastError();
        if (ec != hipSuccess) return fail("long clear launch", ec);
    }
    const size_t lds = f64 ? prologue_lds<double>() : prologue_lds<dd>();
    const void* fn = f64 ? reinterpret_cast<const void*>(&jw_long_prologue_kernel<double>)
                         : reinterpret_cast<const void*>(&jw_long_prologue_kernel<dd>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail("hipFuncSetAttribute(long prologue)", e);
    const dim3 grid((unsigned)nit);
    if (f64)
        hipLaunchKernelGGL(jw_long_prologue_kernel<double>, grid, dim3(kPrologueThreads), lds, s, *src, matrices,
                           reinterpret_cast<double*>(table), chained, sync, reinterpret_cast<double*>(gval), err,
                           format);
    else
        hipLaunchKernelGGL(jw_long_prologue_kernel<dd>, grid, dim3(kPrologueThreads), lds, s, *src, matrices,
                           reinterpret_cast<dd*>(table), chained, sync, reinterpret_cast<dd*>(gval), err, format);
    e = hipGetLastError();
    if (e != hipSuccess) return fail("long prologue launch", e);
    return 0;
}

int jwdev_long_chunks(const jw_long_desc* d, int precision, int grid_x, void* stream)
{
    if (!d->table || !d->red || (d->mode != JW_LONG_COEF && !d->run.rhs))
        return fail("long chunks: no table, reduced data or rhs");
    const bool f64 = precision == JW_PREC_F64;
    const int C = (f64 ? LongGeom<double>::R : LongGeom<dd>::R) * kLongL;
    if (d->nchunks < 1 || (int64_t)d->nchunks * C < d->run.N || d->nchunks > 65535 || grid_x < 1)
        return fail("long chunks: bad chunk count or grid");
    if (d->mode != JW_LONG_COEF && d->run.nsys > d->red_sys) return fail("long chunks: slice exceeds reduced data");
    if (d->mode < JW_LONG_COEF || d->mode > JW_LONG_FINISH) return fail("long chunks: bad mode");
    const void* f = f64 ? chunk_fn<double>(d->mode) : chunk_fn<dd>(d->mode);
    const size_t lds = (size_t)jwdev_long_lds_bytes(precision);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail("hipFuncSetAttribute(long chunks)", e);
    void* args[] = {const_cast<jw_long_desc*>(d)};
    e = hipLaunchKernel(f, dim3((unsigned)grid_x, (unsigned)d->nchunks), dim3(kLongThreads), args, lds,
                        reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail("long chunks launch", e);
    return 0;
}

int jwdev_long_reduced(const jw_long_desc* d, int precision, void* stream)
{
    if (!d->red) return fail("long reduced: no reduced data");
    if (d->run.nsys > d->red_sys) return fail("long reduced: slice exceeds reduced data");
    if (d->run.nsys < 1) return 0;
    // one wavefront per system, four per workgroup
    const unsigned grid = (unsigned)((d->run.nsys + 3) / 4);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (precision == JW_PREC_F64)
        hipLaunchKernelGGL(jw_long_reduced_kernel<double>, dim3(grid), dim3(256), 0, s, *d);
    else
        hipLaunchKernelGGL(jw_long_reduced_kernel<dd>, dim3(grid), dim3(256), 0, s, *d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("long reduced launch", e);
    return 0;
}

}  // extern "C"
