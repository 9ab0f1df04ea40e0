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

constexpr int kLongThreads = 512;   // 8 wavefronts per chunk workgroup
constexpr int kLongL = 64;
constexpr int kLongLog2L = 6;

// rows per lane of a chunk: FP64 8 (512-row chunks, 52 KiB table in LDS), double-double 4
// (256-row chunks, 64 KiB)
template <typename T> struct LongGeom;
template <> struct LongGeom<double> { static constexpr int R = 8; static constexpr int NSUB = 4; };
template <> struct LongGeom<dd> { static constexpr int R = 4; static constexpr int NSUB = 4; };

template <typename T>
constexpr int64_t long_tab_elems()
{
    return (int64_t)JW_NCOEF * LongGeom<T>::R * kLongL + 4 * kLongLog2L * kLongL;
}

}  // namespace

// ------------------------------------------------------------------------
// tables: NSUB chunks per build of nsub * 64 * R rows
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kPrologueThreads) void jw_long_prologue_kernel(jw_source_desc d, const void* mat,
                                                                          T* table, int sequential)
{
    extern __shared__ __align__(16) unsigned char pro_smem[];
    constexpr int R = LongGeom<T>::R;
    constexpr int NSUB = LongGeom<T>::NSUB;
    constexpr int SR = NSUB * R * kLongL;
    constexpr int64_t TE = long_tab_elems<T>();
    // the pivot carry follows the build's five row arrays; g's array is the fifth
    T* carry = reinterpret_cast<T*>(pro_smem + prologue_lds_bytes(SR, sizeof(T)));
    const T* sg = reinterpret_cast<const T*>(pro_smem) + 4 * (int64_t)SR;
    const int64_t nit = (d.N + SR - 1) / SR;
    const int64_t it0 = sequential ? 0 : (int64_t)blockIdx.x;
    const int64_t it1 = sequential ? nit : it0 + 1;
    for (int64_t it = it0; it < it1; it++) {
        const int64_t rb = it * SR;
        prologue_body<T>(d, mat, table + it * NSUB * TE, kLongL, R, 1, pro_smem, rb, sequential ? carry : nullptr,
                         NSUB);
        if (it + 1 < it1) {
            // every thread is done with this build's LDS and carry; the next build starts
            // from the final pivot of the last row and that row's raw super-diagonal entry
            // (the row has a successor: a later build exists)
            __syncthreads();
            if (threadIdx.x == 0) {
                const int64_t r = rb + SR - 1;
                carry[0] = sg[SR - 1];
                carry[1] = mem_load<T>(mat, d.udS + source_pos(d, r, d.N), d.mem_f64);
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------
// chunk kernel: grid (workgroups per chunk, chunks)
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kLongThreads) void jw_long_chunk_kernel(jw_long_desc d)
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
    const int mode = d.mode;
    const int mf = rd.mem_f64;
    const int64_t nc = d.nchunks;
    const int64_t nsys = mode == JW_LONG_COEF ? 3 : rd.nsys;
    const int64_t npass = (nsys + 1) / 2;
    T* red = reinterpret_cast<T*>(d.red);
    const int64_t rs = d.red_sys;
    auto at = [&](int a, int64_t cc, int64_t s) -> int64_t { return ((int64_t)a * nc + cc) * rs + s; };
    T* coef = red + (int64_t)JW_LONG_ARRAYS * nc * rs;

    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        const int64_t sys0 = 2 * q;
        T x[2][R];
        T hm2[2], hm1[2], hp1[2], hp2[2], cw[2], cz[2], cy[2];

        // ---- load: rows outside [in_lo, in_hi) (in_hi <= N) and absent systems read as 0 ----
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            const bool ok = sys < nsys && mode != JW_LONG_COEF;
            const int64_t base = sys * rd.stride + rd.offSol;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int64_t gr = r0 + i;
                const bool k = ok && gr >= rd.in_lo && gr < rd.in_hi;
                x[s][i] = k ? mem_load<T>(rd.rhs, base + gr, mf) : A::zero();
            }
            hm2[s] = hm1[s] = hp1[s] = hp2[s] = A::zero();
            if (rd.band && ok) {
                if (mode == JW_LONG_REDUCE) {
                    // the neighbouring chunks' rows, before any of them is written (the
                    // finishing launch takes them from the reduced data instead)
                    auto rhs_at = [&](int64_t gr) -> T {
                        return (gr >= rd.in_lo && gr < rd.in_hi) ? mem_load<T>(rd.rhs, base + gr, mf) : A::zero();
                    };
                    if (t == 0) {
                        hm2[s] = rhs_at(rb - 2);
                        hm1[s] = rhs_at(rb - 1);
                    }
                    if (t == L - 1) {
                        hp1[s] = rhs_at(rb + C);
                        hp2[s] = rhs_at(rb + C + 1);
                    }
                } else {
                    if (t == 0 && c > 0) {
                        hm2[s] = red[at(5, c - 1, sys)];
                        hm1[s] = red[at(6, c - 1, sys)];
                    }
                    if (t == L - 1 && c + 1 < nc) {
                        hp1[s] = red[at(3, c + 1, sys)];
                        hp2[s] = red[at(4, c + 1, sys)];
                    }
                }
            }
            cw[s] = cz[s] = cy[s] = A::zero();
            if (mode == JW_LONG_COEF) {
                if (sys == 0) cw[s] = A::one();
                if (sys == 1) cz[s] = A::one();
                if (sys == 2) cy[s] = A::one();
            } else if (mode == JW_LONG_FINISH && sys < nsys) {
                cw[s] = red[at(7, c, sys)];
                cz[s] = red[at(8, c, sys)];
                cy[s] = red[at(9, c, sys)];
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
                if (rd.kl >= 1) xm1 = A::sel(t >= 1, lane_fetch(x[s][R - 1], t >= 1 ? lane - 1 : lane), hm1[s]);
                if (rd.kl >= 2) xm2 = A::sel(t >= 1, lane_fetch(x[s][R - 2], t >= 1 ? lane - 1 : lane), hm2[s]);
                if (rd.ku >= 1) xp1 = A::sel(t + 1 < L, lane_fetch(x[s][0], t + 1 < L ? lane + 1 : lane), hp1[s]);
                if (rd.ku >= 2) xp2 = A::sel(t + 1 < L, lane_fetch(x[s][1], t + 1 < L ? lane + 1 : lane), hp2[s]);
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
                T a = A::sel(t == 0, cw[s], A::zero());
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
                T w = A::sel(t >= 1, cprev, cw[s]);
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
                T cin[2];
#pragma unroll
                for (int s = 0; s < 2; s++) cin[s] = stage == 0 ? cz[s] : cy[s];
                T acc[2];
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    T a = A::sel(t == L - 1, cin[s], A::zero());
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
            // ---- store rows in [out_lo, out_hi) (out_hi <= N) ----
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int64_t sys = sys0 + s;
                if (sys >= nsys) continue;
                const int64_t base = sys * rd.stride + rd.offSol;
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const int64_t gr = r0 + i;
                    if (gr >= rd.out_lo && gr < rd.out_hi) mem_store<T>(rd.rhs, base + gr, x[s][i], mf);
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
// reduced solve: one thread per system, the chunks in blocks of B (loads of a block issued
// together, then the dependent updates)
// ------------------------------------------------------------------------
template <typename T, int B>
__global__ __launch_bounds__(256) void jw_long_reduced_kernel(jw_long_desc d)
{
    using A = Ar<T>;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.run.nsys) return;
    const int64_t nc = d.nchunks, rs = d.red_sys;
    const T* __restrict__ in = reinterpret_cast<const T*>(d.red);                        // F, Z, Y
    T* __restrict__ car = reinterpret_cast<T*>(d.red) + 7 * nc * rs;                     // w_in, z_in, y_in
    const T* __restrict__ K = reinterpret_cast<const T*>(d.red) + JW_LONG_ARRAYS * nc * rs;

    // w_in of chunk c is w at the last row of chunk c - 1: F_{c-1} + K_w(c-1) w_in(c-1)
    T w = A::zero();
    for (int64_t c0 = 0; c0 < nc; c0 += B) {
        T f[B], kw[B];
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t c = c0 + j;
            f[j] = A::zero();
            kw[j] = A::zero();
            if (c < nc) {
                f[j] = in[c * rs + s];
                kw[j] = K[9 * c];
            }
        }
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t c = c0 + j;
            if (c < nc) {
                car[c * rs + s] = w;
                w = A::madd(f[j], kw[j], w);
            }
        }
    }
    // z_in, y_in of chunk c are z, y at the first row of chunk c + 1
    T z = A::zero(), y = A::zero();
    for (int64_t c1 = nc - 1; c1 >= 0; c1 -= B) {
        T zc[B], yc[B], wi[B], k1[B], k2[B], k4[B], k5[B], k8[B];
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t c = c1 - j;
            zc[j] = yc[j] = wi[j] = k1[j] = k2[j] = k4[j] = k5[j] = k8[j] = A::zero();
            if (c >= 0) {
                zc[j] = in[(nc + c) * rs + s];
                yc[j] = in[(2 * nc + c) * rs + s];
                wi[j] = car[c * rs + s];
                k1[j] = K[9 * c + 1];
                k2[j] = K[9 * c + 2];
                k4[j] = K[9 * c + 4];
                k5[j] = K[9 * c + 5];
                k8[j] = K[9 * c + 8];
            }
        }
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int64_t c = c1 - j;
            if (c >= 0) {
                car[(nc + c) * rs + s] = z;
                car[(2 * nc + c) * rs + s] = y;
                const T zn = A::madd(A::madd(zc[j], k1[j], wi[j]), k4[j], z);
                const T yn = A::madd(A::madd(A::madd(yc[j], k2[j], wi[j]), k5[j], z), k8[j], y);
                z = zn;
                y = yn;
            }
        }
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
    const void* f = precision == JW_PREC_F64 ? reinterpret_cast<const void*>(&jw_long_chunk_kernel<double>)
                                             : reinterpret_cast<const void*>(&jw_long_chunk_kernel<dd>);
    const size_t lds = (size_t)jwdev_long_lds_bytes(precision);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { fail("hipFuncSetAttribute(long chunks)", e); return 0; }
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kLongThreads, lds);
    if (e != hipSuccess) { fail("occupancy(long chunks)", e); return 0; }
    return nb;
}

int jwdev_long_prologue(const jw_source_desc* src, const void* matrices, void* table, int64_t tab_elems,
                        int precision, void* stream)
{
    if (!matrices || !table) return fail("long prologue: no matrices or table");
    const bool f64 = precision == JW_PREC_F64;
    if (tab_elems != (f64 ? long_tab_elems<double>() : long_tab_elems<dd>()))
        return fail("long prologue: table size mismatch");
    const int SR = f64 ? LongGeom<double>::NSUB * LongGeom<double>::R * kLongL
                       : LongGeom<dd>::NSUB * LongGeom<dd>::R * kLongL;
    const int64_t nit = (src->N + SR - 1) / SR;
    if (nit < 1) return 0;
    // pivots continue from one build to the next: one workgroup, in order
    const int sequential = src->solve_kind == JW_SOLVE_TRI;
    const size_t lds = f64 ? prologue_lds<double>() : prologue_lds<dd>();
    const void* fn = f64 ? reinterpret_cast<const void*>(&jw_long_prologue_kernel<double>)
                         : reinterpret_cast<const void*>(&jw_long_prologue_kernel<dd>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail("hipFuncSetAttribute(long prologue)", e);
    const dim3 grid((unsigned)(sequential ? 1 : nit));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (f64)
        hipLaunchKernelGGL(jw_long_prologue_kernel<double>, grid, dim3(kPrologueThreads), lds, s, *src, matrices,
                           reinterpret_cast<double*>(table), sequential);
    else
        hipLaunchKernelGGL(jw_long_prologue_kernel<dd>, grid, dim3(kPrologueThreads), lds, s, *src, matrices,
                           reinterpret_cast<dd*>(table), sequential);
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
    const void* f = f64 ? reinterpret_cast<const void*>(&jw_long_chunk_kernel<double>)
                        : reinterpret_cast<const void*>(&jw_long_chunk_kernel<dd>);
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
    const unsigned grid = (unsigned)((d->run.nsys + 255) / 256);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (precision == JW_PREC_F64)
        hipLaunchKernelGGL((jw_long_reduced_kernel<double, 8>), dim3(grid), dim3(256), 0, s, *d);
    else
        hipLaunchKernelGGL((jw_long_reduced_kernel<dd, 4>), dim3(grid), dim3(256), 0, s, *d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("long reduced launch", e);
    return 0;
}

}  // extern "C"
