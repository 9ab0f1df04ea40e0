// jw_dd64.hip -- the double-double JW step on the wave-per-system DPP kernel family.
//
// One wavefront owns a pair of systems; lane t owns rows t*R .. t*R+R-1 of both (R in
// {2, 4, 8}: N <= 512).  The coefficient table is format 2 in double-double (8
// coefficients per row with the row scalings folded in, then per recurrence the DPP scan
// multipliers), built by jw_prologue_kernel<dd> and staged once per workgroup in LDS.
// Per pair, in registers: banded matvec with the neighbour rows by DPP wave shifts; the
// forward and the two backward first-order recurrences, each a lane-local sweep, an
// affine scan across the wavefront (DPP row shifts 1, 2, 4, 8 inside 16-lane rows, then
// across rows row_bcast15/row_bcast31 forward or readlane of the row starts backward)
// and a lane-local redo; row-0 scaleC.  Every DPP move of a double-double moves its two
// doubles (four 32-bit moves).  Arithmetic: the reference's double-double forms
// (dd_mul_r / dd_madd_r of jw_common.h, after
// PfSolve/pfSolve_CodeGen/pfSolve_MathUtils/pfSolve_MathUtils.h:1093-1196,2398-2412).
// Storage: DD (16 bytes per value) or FP64 promoted with a zero low word (mem_f64); rows
// sys*stride + offSol + row, one 16-byte (or 8-byte) access per row and lane.
//
// Reference semantics restated (paths relative to the reference root):
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:38-408
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_TridiagonalSolve.h:1578-1929
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "jw_common.h"
#include "jw_prologue.h"

extern "C" void jwdev_set_error(const char* msg);

namespace {

constexpr int kDdThreads = 512;
constexpr int DPP_BCAST15 = 0x142;
constexpr int DPP_BCAST31 = 0x143;
constexpr int SHR(int n) { return 0x110 + n; }
constexpr int SHL(int n) { return 0x100 + n; }
constexpr int WAVE_SHL1 = 0x130, WAVE_SHR1 = 0x138;

template <int CTRL>
__device__ __forceinline__ dd dpp_dd(dd v)
{
    return dd_make(dpp_d<CTRL>(v.hi), dpp_d<CTRL>(v.lo));
}

// DPP move with a row mask: rows outside the mask and lanes without a source get 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_rm_d(double v)
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ dd dpp_rm_dd(dd v)
{
    return dd_make(dpp_rm_d<CTRL, ROWMASK>(v.hi), dpp_rm_d<CTRL, ROWMASK>(v.lo));
}

__device__ __forceinline__ dd readlane_dd(dd v, int l) { return dd_make(readlane_d(v.hi, l), readlane_d(v.lo, l)); }

template <int R>
__device__ __forceinline__ dd tabv(const dd* tab, int p, int i, int lane, int c)
{
    return tab[tab_coef64(p, i, lane, R) + c];
}

// keeps the scheduler from hoisting one phase's table loads into the previous phase
__device__ __forceinline__ void phase_barrier() { __builtin_amdgcn_sched_barrier(0); }

// the pair's matvec, the recurrences present (fwd, b1, b2: wave-uniform) and row-0 scale
template <int R>
__device__ __forceinline__ void compute_pair_dd(const dd* tab, int lane, int kl, int ku, bool fwd, bool b1, bool b2,
                                                bool apply_scale, dd scale, dd (&x0)[R], dd (&x1)[R])
{
    using A = Ar<dd>;
    constexpr int SC = 8 * R * 64;
    // ---- banded matvec (coefficients carry the folded row scaling) ----
    {
        dd m1a = A::zero(), m2a = A::zero(), p1a = A::zero(), p2a = A::zero();
        dd m1b = A::zero(), m2b = A::zero(), p1b = A::zero(), p2b = A::zero();
        if (kl >= 1) { m1a = dpp_dd<WAVE_SHR1>(x0[R - 1]); m1b = dpp_dd<WAVE_SHR1>(x1[R - 1]); }
        if (kl >= 2) { m2a = dpp_dd<WAVE_SHR1>(x0[R - 2]); m2b = dpp_dd<WAVE_SHR1>(x1[R - 2]); }
        if (ku >= 1) { p1a = dpp_dd<WAVE_SHL1>(x0[0]); p1b = dpp_dd<WAVE_SHL1>(x1[0]); }
        if (ku >= 2) { p2a = dpp_dd<WAVE_SHL1>(x0[1]); p2b = dpp_dd<WAVE_SHL1>(x1[1]); }
#pragma unroll
        for (int i = 0; i < R; i++) {
            const dd vm2 = tabv<R>(tab, 0, i, lane, 0), vm1 = tabv<R>(tab, 0, i, lane, 1);
            const dd v0 = tabv<R>(tab, 1, i, lane, 0), vp1 = tabv<R>(tab, 1, i, lane, 1);
            const dd vp2 = tabv<R>(tab, 2, i, lane, 0);
            const dd ia = x0[i], ib = x1[i];
            const dd n1a = (i + 1 < R) ? x0[(i + 1) % R] : p1a;
            const dd n1b = (i + 1 < R) ? x1[(i + 1) % R] : p1b;
            const dd n2a = (i + 2 < R) ? x0[(i + 2) % R] : ((i + 2 == R) ? p1a : p2a);
            const dd n2b = (i + 2 < R) ? x1[(i + 2) % R] : ((i + 2 == R) ? p1b : p2b);
            dd ua = dd_mul_r(v0, ia), ub = dd_mul_r(v0, ib);
            ua = dd_madd_r(ua, vm1, m1a);
            ub = dd_madd_r(ub, vm1, m1b);
            ua = dd_madd_r(ua, vm2, m2a);
            ub = dd_madd_r(ub, vm2, m2b);
            ua = dd_madd_r(ua, vp1, n1a);
            ub = dd_madd_r(ub, vp1, n1b);
            ua = dd_madd_r(ua, vp2, n2a);
            ub = dd_madd_r(ub, vp2, n2b);
            m2a = m1a;
            m2b = m1b;
            m1a = ia;
            m1b = ib;
            x0[i] = ua;
            x1[i] = ub;
        }
    }
    phase_barrier();
    // ---- forward recurrence W_r = u'_r - l'_r W_{r-1} ----
    if (fwd) {
        dd lr[R];
#pragma unroll
        for (int i = 0; i < R; i++) lr[i] = tabv<R>(tab, 2, i, lane, 1);
        dd a = x0[0], b = x1[0];
#pragma unroll
        for (int i = 1; i < R; i++) {
            a = dd_msub_r(x0[i], lr[i], a);
            b = dd_msub_r(x1[i], lr[i], b);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const dd f = tab[SC + k * 64 + lane];
            dd sa, sb;
            switch (k) {
            case 0: sa = dpp_dd<SHR(1)>(a); sb = dpp_dd<SHR(1)>(b); break;
            case 1: sa = dpp_dd<SHR(2)>(a); sb = dpp_dd<SHR(2)>(b); break;
            case 2: sa = dpp_dd<SHR(4)>(a); sb = dpp_dd<SHR(4)>(b); break;
            default: sa = dpp_dd<SHR(8)>(a); sb = dpp_dd<SHR(8)>(b); break;
            }
            a = dd_madd_r(a, f, sa);
            b = dd_madd_r(b, f, sb);
        }
        {
            // rows 1 and 3 take the ends of rows 0 and 2; rows 2 and 3 then take lane 31
            const dd fp = tab[SC + 4 * 64 + lane], fp2 = tab[SC + 5 * 64 + lane];
            const dd sa = dpp_rm_dd<DPP_BCAST15, 0xa>(a), sb = dpp_rm_dd<DPP_BCAST15, 0xa>(b);
            a = dd_madd_r(a, fp, sa);
            b = dd_madd_r(b, fp, sb);
            const dd ta = dpp_rm_dd<DPP_BCAST31, 0xc>(a), tb = dpp_rm_dd<DPP_BCAST31, 0xc>(b);
            a = dd_madd_r(a, fp2, ta);
            b = dd_madd_r(b, fp2, tb);
        }
        dd wa = dpp_dd<WAVE_SHR1>(a), wb = dpp_dd<WAVE_SHR1>(b);
#pragma unroll
        for (int i = 0; i < R; i++) {
            wa = dd_msub_r(x0[i], lr[i], wa);
            wb = dd_msub_r(x1[i], lr[i], wb);
            x0[i] = wa;
            x1[i] = wb;
        }
    }
    phase_barrier();
    // ---- backward recurrences Z_r = W_r - h'_r Z_{r+1}, y_r = Z_r - E_r y_{r+1} ----
#pragma unroll
    for (int stage = 0; stage < 2; stage++) {
        if (stage == 1) phase_barrier();
        if ((stage == 0 && !b1) || (stage == 1 && !b2)) continue;
        dd hh[R];
#pragma unroll
        for (int i = 0; i < R; i++) hh[i] = tabv<R>(tab, 3, i, lane, stage);
        const int j = 1 + stage;
        dd a = x0[R - 1], b = x1[R - 1];
#pragma unroll
        for (int i = R - 2; i >= 0; i--) {
            a = dd_msub_r(x0[i], hh[i], a);
            b = dd_msub_r(x1[i], hh[i], b);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const dd f = tab[SC + (j * 8 + k) * 64 + lane];
            dd sa, sb;
            switch (k) {
            case 0: sa = dpp_dd<SHL(1)>(a); sb = dpp_dd<SHL(1)>(b); break;
            case 1: sa = dpp_dd<SHL(2)>(a); sb = dpp_dd<SHL(2)>(b); break;
            case 2: sa = dpp_dd<SHL(4)>(a); sb = dpp_dd<SHL(4)>(b); break;
            default: sa = dpp_dd<SHL(8)>(a); sb = dpp_dd<SHL(8)>(b); break;
            }
            a = dd_madd_r(a, f, sa);
            b = dd_madd_r(b, f, sb);
        }
        {
            // every row adds the in-row results of the later row starts (lanes 16, 32, 48),
            // weighted by matrix-only multipliers (zero for rows at or after that start)
            const dd c1 = tab[SC + (j * 8 + 4) * 64 + lane];
            const dd c2 = tab[SC + (j * 8 + 5) * 64 + lane];
            const dd c3 = tab[SC + (j * 8 + 6) * 64 + lane];
            const dd a1 = readlane_dd(a, 16), a2 = readlane_dd(a, 32), a3 = readlane_dd(a, 48);
            const dd q1 = readlane_dd(b, 16), q2 = readlane_dd(b, 32), q3 = readlane_dd(b, 48);
            a = dd_madd_r(a, c1, a1);
            b = dd_madd_r(b, c1, q1);
            a = dd_madd_r(a, c2, a2);
            b = dd_madd_r(b, c2, q2);
            a = dd_madd_r(a, c3, a3);
            b = dd_madd_r(b, c3, q3);
        }
        dd za = dpp_dd<WAVE_SHL1>(a), zb = dpp_dd<WAVE_SHL1>(b);
#pragma unroll
        for (int i = R - 1; i >= 0; i--) {
            za = dd_msub_r(x0[i], hh[i], za);
            zb = dd_msub_r(x1[i], hh[i], zb);
            x0[i] = za;
            x1[i] = zb;
        }
    }
    // ---- row-0 scaleC (pfSolve_TridiagonalSolve.h:1918-1927) ----
    if (apply_scale && lane == 0) {
        x0[0] = dd_mul_r(x0[0], scale);
        x1[0] = dd_mul_r(x1[0], scale);
    }
}

// The DD pair kernel: the table staged in LDS once per workgroup, then pairs of
// systems over the grid.  Rows outside [in_lo, in_hi) read as 0; rows in
// [out_lo, out_hi) are written.
template <int R>
__global__ __launch_bounds__(kDdThreads) void jw_dd64_kernel(jw_run_desc d)
{
    using A = Ar<dd>;
    constexpr int NTAB = JW_TAB64_ELEMS(R);
    extern __shared__ __align__(16) unsigned char smem[];
    const dd* tab = reinterpret_cast<const dd*>(smem);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    copy_to_lds(smem, d.table, NTAB * 16, wave, nwaves, lane);
    __syncthreads();
    const int mf = d.mem_f64;
    const int r0 = lane * R;
    const dd scale = A::from(d.scale_hi, d.scale_lo);
    const int64_t npass = (d.nsys + 1) / 2;
    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        const int64_t sys0 = 2 * q;
        const bool ok1 = sys0 + 1 < d.nsys;
        const int64_t base0 = sys0 * d.stride + d.offSol, base1 = base0 + d.stride;
        dd x0[R], x1[R];
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int r = r0 + i;
            const bool in = r >= d.in_lo && r < d.in_hi;
            x0[i] = A::zero();
            x1[i] = A::zero();
            if (in) x0[i] = mem_load<dd>(d.rhs, base0 + r, mf);
            if (in && ok1) x1[i] = mem_load<dd>(d.rhs, base1 + r, mf);
        }
        compute_pair_dd<R>(tab, lane, d.kl, d.ku, d.fwd != 0, d.b1 != 0, d.b2 != 0, d.apply_scale != 0, scale, x0, x1);
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int r = r0 + i;
            if (r >= d.out_lo && r < d.out_hi) {
                mem_store<dd>(d.rhs, base0 + r, x0[i], mf);
                if (ok1) mem_store<dd>(d.rhs, base1 + r, x1[i], mf);
            }
        }
    }
}

const void* pick_dd64(int R)
{
    switch (R) {
    case 2: return reinterpret_cast<const void*>(&jw_dd64_kernel<2>);
    case 4: return reinterpret_cast<const void*>(&jw_dd64_kernel<4>);
    case 8: return reinterpret_cast<const void*>(&jw_dd64_kernel<8>);
    default: return nullptr;
    }
}

int64_t dd64_lds(int R) { return (int64_t)JW_TAB64_ELEMS(R) * 16; }

}  // namespace

extern "C" {

/* workgroups of `waves` wavefronts per CU (0: no kernel for R) */
int jwdev_dd64_occupancy(int R, int waves)
{
    const void* f = pick_dd64(R);
    if (!f || waves < 1 || 64 * waves > kDdThreads) return 0;
    const int lds = (int)dd64_lds(R);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds) != hipSuccess) return 0;
    return nb;
}

int64_t jwdev_dd64_table_elems(int R) { return pick_dd64(R) ? (int64_t)JW_TAB64_ELEMS(R) : 0; }

int jwdev_dd64(const jw_run_desc* r, int R, int waves, int grid, void* stream)
{
    if (!r->rhs || !r->table) { jwdev_set_error("dd64: no rhs or table"); return -1; }
    const void* f = pick_dd64(R);
    if (!f || waves < 1 || 64 * waves > kDdThreads || grid < 1) { jwdev_set_error("dd64: bad launch"); return -1; }
    const size_t lds = (size_t)dd64_lds(R);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return -2;
    void* args[] = {const_cast<jw_run_desc*>(r)};
    const hipError_t e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) { jwdev_set_error("dd64 launch failed"); return -3; }
    return 0;
}

}  // extern "C"
