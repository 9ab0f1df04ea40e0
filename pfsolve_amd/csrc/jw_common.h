// jw_common.h -- device helpers shared by the gfx950 JW kernels: double-double
// arithmetic, arithmetic traits (one kernel source for FP64 and DD), cross-lane
// moves (ds_bpermute, DPP row/wave shifts, readlane), caller-memory access and
// the reference's storage permutation.
#ifndef PFSOLVE_JW_COMMON_H
#define PFSOLVE_JW_COMMON_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jw_internal.h"

// ------------------------------------------------------------------------
// double-double arithmetic (Dekker 1971; Knuth TwoSum; Joldes, Muller,
// Popescu, ACM TOMS 44(2) 2017, Alg. 6 (add), 12 (mul), 17 (div)).
// This file is compiled with -ffp-contract=off so that the error-free
// transforms stay exact; FMAs are written explicitly.
// ------------------------------------------------------------------------
struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd dd_make(double h, double l) { dd r; r.hi = h; r.lo = l; return r; }

__device__ __forceinline__ dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    return dd_make(s, (a - (s - bb)) + (b - bb));
}

__device__ __forceinline__ dd fast_two_sum(double a, double b)
{
    double s = a + b;
    return dd_make(s, b - (s - a));
}

__device__ __forceinline__ dd two_prod(double a, double b)
{
    double p = a * b;
    return dd_make(p, __builtin_fma(a, b, -p));
}

__device__ __forceinline__ dd dd_add(dd x, dd y)
{
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    dd v = fast_two_sum(s.hi, c);
    double w = t.lo + v.lo;
    return fast_two_sum(v.hi, w);
}

__device__ __forceinline__ dd dd_neg(dd x) { return dd_make(-x.hi, -x.lo); }
__device__ __forceinline__ dd dd_sub(dd x, dd y) { return dd_add(x, dd_neg(y)); }

__device__ __forceinline__ dd dd_mul(dd x, dd y)
{
    dd c = two_prod(x.hi, y.hi);
    double tl0 = x.lo * y.lo;
    double tl1 = __builtin_fma(x.hi, y.lo, tl0);
    double cl2 = __builtin_fma(x.lo, y.hi, tl1);
    double cl3 = c.lo + cl2;
    return fast_two_sum(c.hi, cl3);
}

__device__ __forceinline__ dd dd_div(dd x, dd y)
{
    double th = x.hi / y.hi;
    dd c = two_prod(y.hi, th);
    double cl = __builtin_fma(y.lo, th, c.lo);
    dd r = fast_two_sum(c.hi, cl);
    double pih = x.hi - r.hi;
    double dl = x.lo - r.lo;
    double d = pih + dl;
    double tl = d / y.hi;
    return fast_two_sum(th, tl);
}

// ------------------------------------------------------------------------
// double-double as the reference's generated code forms it
// (PfSolve/pfSolve_CodeGen/pfSolve_MathUtils/pfSolve_MathUtils.h): a sum is TwoSum of the
// high words, the low words added to its error term, one quick renormalisation
// (PfQuadSum / PfQuadQuickSum, :1093-1152, double-double branch :1188-1196); a product is
// TwoProd of the high words, the two cross terms added by FMA, one quick renormalisation
// (:2398-2412).  dd_madd_r fuses a + b*c: the product's two words go into the sum without
// the product's own renormalisation (15 FP64 operations; the accurate forms above take 29).
// ------------------------------------------------------------------------
__device__ __forceinline__ dd dd_mul_r(dd x, dd y)
{
    const double p = x.hi * y.hi;
    double e = __builtin_fma(x.hi, y.hi, -p);
    e = __builtin_fma(x.hi, y.lo, e);
    e = __builtin_fma(x.lo, y.hi, e);
    return fast_two_sum(p, e);
}

__device__ __forceinline__ dd dd_madd_r(dd a, dd b, dd c)
{
    const double p = b.hi * c.hi;
    double e = __builtin_fma(b.hi, c.hi, -p);
    e = __builtin_fma(b.hi, c.lo, e);
    e = __builtin_fma(b.lo, c.hi, e);
    const dd s = two_sum(a.hi, p);
    return fast_two_sum(s.hi, s.lo + (a.lo + e));
}

// a - b*c
__device__ __forceinline__ dd dd_msub_r(dd a, dd b, dd c) { return dd_madd_r(a, dd_neg(b), c); }

// ------------------------------------------------------------------------
// cross-lane moves (ds_bpermute: any lane of the wavefront, 32 bits per op)
// ------------------------------------------------------------------------
__device__ __forceinline__ double lane_fetch(double v, int src_lane)
{
    int lo = __double2loint(v);
    int hi = __double2hiint(v);
    lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, lo);
    hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, hi);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ dd lane_fetch(dd v, int src_lane)
{
    return dd_make(lane_fetch(v.hi, src_lane), lane_fetch(v.lo, src_lane));
}

// ------------------------------------------------------------------------
// arithmetic traits: one kernel source for FP64 and double-double
// ------------------------------------------------------------------------
template <typename T> struct Ar;

template <> struct Ar<double> {
    __device__ static __forceinline__ double zero() { return 0.0; }
    __device__ static __forceinline__ double one() { return 1.0; }
    __device__ static __forceinline__ double from(double h, double) { return h; }
    __device__ static __forceinline__ double add(double a, double b) { return a + b; }
    __device__ static __forceinline__ double sub(double a, double b) { return a - b; }
    __device__ static __forceinline__ double mul(double a, double b) { return a * b; }
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
    __device__ static __forceinline__ double neg(double a) { return -a; }
    // a + b*c and a - b*c as single roundings
    __device__ static __forceinline__ double madd(double a, double b, double c) { return __builtin_fma(b, c, a); }
    __device__ static __forceinline__ double msub(double a, double b, double c) { return __builtin_fma(-b, c, a); }
    __device__ static __forceinline__ double hi(double a) { return a; }
    __device__ static __forceinline__ double ldexp_(double a, int e) { return ldexp(a, e); }
    __device__ static __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
};

template <> struct Ar<dd> {
    __device__ static __forceinline__ dd zero() { return dd_make(0.0, 0.0); }
    __device__ static __forceinline__ dd one() { return dd_make(1.0, 0.0); }
    __device__ static __forceinline__ dd from(double h, double l) { return dd_make(h, l); }
    __device__ static __forceinline__ dd add(dd a, dd b) { return dd_add(a, b); }
    __device__ static __forceinline__ dd sub(dd a, dd b) { return dd_sub(a, b); }
    // products and multiply-adds in the reference's forms (dd_mul_r, dd_madd_r); sums,
    // differences and quotients in the accurate forms
    __device__ static __forceinline__ dd mul(dd a, dd b) { return dd_mul_r(a, b); }
    __device__ static __forceinline__ dd div(dd a, dd b) { return dd_div(a, b); }
    __device__ static __forceinline__ dd neg(dd a) { return dd_neg(a); }
    __device__ static __forceinline__ dd madd(dd a, dd b, dd c) { return dd_madd_r(a, b, c); }
    __device__ static __forceinline__ dd msub(dd a, dd b, dd c) { return dd_msub_r(a, b, c); }
    __device__ static __forceinline__ double hi(dd a) { return a.hi; }
    __device__ static __forceinline__ dd ldexp_(dd a, int e) { return dd_make(ldexp(a.hi, e), ldexp(a.lo, e)); }
    __device__ static __forceinline__ dd sel(bool c, dd a, dd b) { return c ? a : b; }
};

// element access of caller memory: T storage, or FP64 storage promoted to T
template <typename T> __device__ __forceinline__ T mem_load(const void* base, int64_t idx, int mem_f64)
{
    if (mem_f64) return Ar<T>::from(reinterpret_cast<const double*>(base)[idx], 0.0);
    return reinterpret_cast<const T*>(base)[idx];
}
// double-double without a branch on the storage: the two doubles of element idx (DD
// storage), or the double at idx twice with the low word then 0 (FP64 storage).  A caller's
// loads then stay in flight together (a branch around each load makes the compiler wait for
// it before the join).
template <> __device__ __forceinline__ dd mem_load<dd>(const void* base, int64_t idx, int mem_f64)
{
    const double* p = reinterpret_cast<const double*>(base);
    const int64_t i = mem_f64 ? idx : 2 * idx;
    const double h = p[i];
    const double l = p[mem_f64 ? i : i + 1];
    return dd_make(h, mem_f64 ? 0.0 : l);
}
template <> __device__ __forceinline__ double mem_load<double>(const void* base, int64_t idx, int)
{
    return reinterpret_cast<const double*>(base)[idx];
}
template <typename T> __device__ __forceinline__ void mem_store(void* base, int64_t idx, T v, int mem_f64)
{
    if (mem_f64) {
        // round DD to FP64 storage
        reinterpret_cast<double*>(base)[idx] = Ar<T>::hi(Ar<T>::add(v, Ar<T>::zero()));
        return;
    }
    reinterpret_cast<T*>(base)[idx] = v;
}
template <> __device__ __forceinline__ void mem_store<double>(void* base, int64_t idx, double v, int)
{
    reinterpret_cast<double*>(base)[idx] = v;
}

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v / 2); }

// reorder_i0 (storage position of logical row i,
// PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142, HIP branch).
// Up to 3840 rows the whole diagonal is one block; longer diagonals are blocks of Msplit0
// rows (the last one shorter), each permuted on its own.  Inside a block of n rows laid out
// as R rows per lane (R odd, from the block size or Nalign), row k is held in register
// k % R of lane k / R, and registers are stored one after the other, register j holding
// ceil(n/R) rows for j < n % R and floor(n/R) after.
__device__ __forceinline__ int64_t reorder_pos(int64_t i, int64_t N, int64_t Nalign = 0, int64_t Msplit0 = 704)
{
    // 32-bit arithmetic (systems have fewer than 2^31 rows): the divisions are the cost of
    // this function in the table builds
    const uint32_t ui = (uint32_t)i, uN = (uint32_t)N;
    uint32_t blk = uN;
    uint32_t NA = Nalign > 0 ? (uint32_t)Nalign : uN;
    if (uN > 3840u) {
        blk = Msplit0 > 0 ? (uint32_t)Msplit0 : 704u;
        NA = blk;
    }
    // rows per thread of the register layout: from Nalign (the largest system of a
    // chained step) or the block size, the split of the stored rows over them: from N
    uint32_t ws = ((NA + 1023u) / 1024u) * 64u;
    if (NA < 512u) ws = 64u;
    uint32_t R = (NA + ws - 1u) / ws;
    if ((R & 1u) == 0u) R++;
    const uint32_t base = blk == uN ? 0u : (ui / blk) * blk;
    const uint32_t n = base + blk > uN ? uN - base : blk;
    const uint32_t li = ui - base;
    const uint32_t ceilq = (n + R - 1u) / R;
    const uint32_t q = li / R, k = li - q * R;
    const uint32_t nq = n / R, rem = n - nq * R;
    uint32_t ret = q;
    if (k < rem)
        ret += k * ceilq;
    else
        ret += rem * ceilq + (k - rem) * nq;
    return (int64_t)(base + ret);
}


// Storage position of row i in the ALT layout (the reference's reorder_i_alt,
// PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:144-219, N <= 3840):
// row i is row k = i/2 of the subsystem of its parity, held in register k % R by lane k / R
// of the parity's half wavefront (W lanes, R rows per lane of the layout of N rows).
__device__ __forceinline__ int64_t alt_pos(int64_t i, int64_t N)
{
    int64_t W = ((N + 1023) / 1024) * 64;
    if (N < 512) W = 64;
    int64_t R = (N + W - 1) / W;
    if ((R & 1) == 0) R++;
    const int64_t k = i >> 1;
    return (k % R) * W + (i & 1) * (W >> 1) + k / R;
}

// ------------------------------------------------------------------------
// DPP moves of a double (two 32-bit v_mov_b32_dpp); lanes without a source
// receive 0 (bound_ctrl).  CTRL: row_shl:n = 0x100+n, row_shr:n = 0x110+n,
// wave_shl:1 = 0x130, wave_shr:1 = 0x138.
// ------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
constexpr int DPP_ROW_SHL(int n) { return 0x100 + n; }
constexpr int DPP_ROW_SHR(int n) { return 0x110 + n; }
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}


// `bytes` (a multiple of 16) from global memory at `src` (16-byte aligned) into LDS
// at `lds`: global_load_lds moves 64 x 16 bytes per instruction, every wavefront
// issues all of its blocks before one wait (a copy through registers would wait
// for each load in turn); the caller's barrier then publishes the data.  `wave`
// must be wavefront-uniform.
__device__ __forceinline__ void copy_to_lds(void* lds, const void* src, int bytes, int wave, int nwaves, int lane)
{
    const int n16 = bytes / 16;
    const double2* s = reinterpret_cast<const double2*>(src);
    double2* d = reinterpret_cast<double2*>(lds);
    for (int k = wave; 64 * k < n16; k += nwaves) {
        const int c = 64 * k + lane;
        if (c < n16)
            __builtin_amdgcn_global_load_lds(s + c, (__attribute__((address_space(3))) void*)(d + 64 * k), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wavefront's blocks have landed
}

// coefficient table of format 2 (64 lanes): pair p of row-in-lane i, lane t
__host__ __device__ __forceinline__ int tab_coef64(int p, int i, int t, int R)
{
    return ((p * R + i) * 64 + t) * 2;
}

// block scale operators of row n (PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_RegistersScaleC.h:27-201,
// runtime-scaleC branches): 1 ScaleC, 2 ScaleD, 3 SphLaplA, 4 SphLaplB; anything else 1
__device__ __forceinline__ double jw_block_scale(int op, int64_t n, int lshift, double s)
{
    const double x = (double)n, l = (double)lshift;
    switch (op) {
    case 1: return s;
    case 2: return 2.0 * (x + l + 1.0) * sqrt((x + 1.0) / (x + s + l + 1.5));
    case 3: return 4.0 * (x + l + 2.0) * (x + l + 3.0) *
                   sqrt((x + 1.0) * (x + 2.0) / ((x + s + l + 2.5) * (x + s + l + 3.5)));
    case 4: return (x + l + 1.0) * (4.0 * l + 6.0) * sqrt((x + 1.0) / (x + s + l + 1.5));
    default: return 1.0;
    }
}

#endif
