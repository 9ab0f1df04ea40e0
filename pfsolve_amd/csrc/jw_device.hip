// jw_device.hip -- gfx950 (CDNA4, wave64) kernels of the Jones-Worland step.
//
// Kernels
//   jw_prologue_kernel  one workgroup per matrix set: reads the caller's
//                       matrices in the reference storage (any JW mode or the
//                       fused V/T/M step), factors the unit-diagonal solve
//                       (LU of I+L+U, pivots by a Moebius-map scan across
//                       lanes), and writes the per-row coefficient table and
//                       the per-lane scan multipliers.  jw_prologue_batch_kernel
//                       does the same for many sets, one workgroup each.
//   jw_main_kernel<T,R,L>
//                       persistent: each workgroup stages one table in LDS,
//                       each wavefront owns 64/L groups of L lanes, each group
//                       owns a pair of systems (the re/im halves of a complex
//                       system in XXII layout), each lane owns R consecutive
//                       rows.  Per pair: banded matvec (neighbour rows by
//                       ds_bpermute), one forward and two backward first-order
//                       recurrences, each as lane-local sweep + log2(L)-step
//                       affine scan across lanes + lane-local redo, row-0
//                       scaleC, store.  FP64 or double-double.
//   jw_block_kernel     block plan: layout conversion + scale operator.
//   jw_gbmv_kernel      dgbmv plan: general band matvec, x staged in LDS.
//
// Reference semantics restated (paths relative to the reference root):
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:38-408
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_MatVecMul.h:28-133,383-446
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_TridiagonalSolve.h:1578-1929
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:1561-1733,1906-2012,2275-2371
//   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_RegistersScaleC.h:27-201
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "jw_common.h"
#include "jw_prologue.h"

// SMALL: a build of at most kPrologueThreads rows (one row per thread, at most 2 per thread
// in the pivot scan), whose per-thread arrays then fit in fewer registers (the double-double
// build spills otherwise)
template <typename T, bool SMALL = false>
__global__ __launch_bounds__(kPrologueThreads) void jw_prologue_kernel(jw_source_desc d, const void* mat, T* table,
                                                                       int L, int R, int format)
{
    extern __shared__ __align__(16) unsigned char pro_smem[];
    if constexpr (SMALL)
        prologue_body<T, NoPivotHooks, false, 1, 2>(d, mat, table, L, R, format, pro_smem);
    else
        prologue_body<T>(d, mat, table, L, R, format, pro_smem);
}

// batched: one workgroup per matrix set; workgroup 0 also clears the work-queue counters.
// NT = 256 when every set has at most 1024 rows: four such workgroups share a CU
// (LDS and registers), so 1024 sets need one round instead of two.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void jw_prologue_batch_kernel(const jw_prologue_job* jobs, int L_unused,
                                                                             int* counters, int ncounters)
{
    extern __shared__ __align__(16) unsigned char pro_smem[];
    if (blockIdx.x == 0 && counters && (int)threadIdx.x < ncounters) counters[threadIdx.x] = 0;
    const jw_prologue_job& j = jobs[blockIdx.x];
    prologue_body<T>(j.src, j.matrices, reinterpret_cast<T*>(j.table), j.L, j.R, j.format, pro_smem);
}

// chained step: one workgroup per iteration, each building its format-2 table (64 lanes,
// R rows per lane, R <= 16 so 256 threads cover the rows); the jobs come by value
__global__ __launch_bounds__(256) void jw_prologue_chain_kernel(jw_chain_jobs j)
{
    extern __shared__ __align__(16) unsigned char pro_smem[];
    const int k = blockIdx.x;
    double* table = reinterpret_cast<double*>(static_cast<unsigned char*>(j.table) + (int64_t)k * j.tab_stride);
    prologue_body<double>(j.src[k], j.matrices[k], table, 64, j.R, 2, pro_smem);
}

// ------------------------------------------------------------------------
// main kernel
// ------------------------------------------------------------------------
template <typename T> struct Pair { T a, b; };

template <typename T, int R, int L>
__device__ __forceinline__ Pair<T> lds_pair(const T* tab, int p, int i, int t)
{
    Pair<T> r;
    const T* q = tab + tab_coef(p, i, t, 0, R, L);
    r.a = q[0];
    r.b = q[1];
    return r;
}

template <typename T, int R, int L>
__global__ __launch_bounds__(512) void jw_main_kernel(jw_run_desc d)
{
    using A = Ar<T>;
    extern __shared__ __align__(16) unsigned char jw_smem[];
    T* tab = reinterpret_cast<T*>(jw_smem);
    constexpr int LOG2L = ilog2_c(L);
    constexpr int NTAB = JW_NCOEF * R * L + 4 * LOG2L * L;
    constexpr int G = 64 / L;

    // stage the coefficient table in LDS
    static_assert(NTAB * (int)sizeof(T) % 16 == 0, "the table is whole 16-byte chunks");
    copy_to_lds(jw_smem, d.table, NTAB * (int)sizeof(T), __builtin_amdgcn_readfirstlane(threadIdx.x >> 6),
                blockDim.x >> 6, threadIdx.x & 63);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int t = lane & (L - 1);
    const int g = lane / L;
    const int r0 = t * R;
    const int mf = d.mem_f64;
    const int64_t npass = (d.nsys + 2 * G - 1) / (2 * G);

    for (int64_t q = (int64_t)blockIdx.x * nwaves + wave; q < npass; q += (int64_t)gridDim.x * nwaves) {
        const int64_t sys0 = (q * G + g) * 2;
        T x[2][R];

        // ---- load (rows outside [in_lo, in_hi) and absent systems read as 0) ----
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            const bool ok = sys < d.nsys;
            const int64_t base = sys * d.stride + d.offSol;
            if constexpr (sizeof(T) == 8 && (R % 2) == 0) {
                if (d.vec2) {
#pragma unroll
                    for (int i = 0; i < R; i += 2) {
                        const int r = r0 + i;
                        const bool k0 = ok && r >= d.in_lo && r < d.in_hi;
                        const bool k1 = ok && r + 1 >= d.in_lo && r + 1 < d.in_hi;
                        const double* p = reinterpret_cast<const double*>(d.rhs) + base + r;
                        if (k0 && k1) {
                            double2 v = *reinterpret_cast<const double2*>(p);
                            x[s][i] = v.x;
                            x[s][i + 1] = v.y;
                        } else {
                            x[s][i] = k0 ? p[0] : 0.0;
                            x[s][i + 1] = k1 ? p[1] : 0.0;
                        }
                    }
                    continue;
                }
            }
            const int64_t rs = d.row_step > 1 ? d.row_step : 1;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = r0 + i;
                const bool k = ok && r >= d.in_lo && r < d.in_hi;
                x[s][i] = k ? mem_load<T>(d.rhs, base + r * rs, mf) : A::zero();
            }
        }

        // ---- banded matvec u = V x (in place) ----
        if (d.band) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                // x[r0-2], x[r0-1], x[r0+R], x[r0+R+1] from the neighbour lanes of the group
                T xm2 = A::zero(), xm1 = A::zero(), xp1 = A::zero(), xp2 = A::zero();
                if (d.kl >= 1) {
                    T f = lane_fetch(x[s][R - 1], t >= 1 ? lane - 1 : lane);
                    xm1 = A::sel(t >= 1, f, A::zero());
                }
                if (d.kl >= 2) {
                    if constexpr (R >= 2) {
                        T f = lane_fetch(x[s][R - 2], t >= 1 ? lane - 1 : lane);
                        xm2 = A::sel(t >= 1, f, A::zero());
                    } else {
                        T f = lane_fetch(x[s][0], t >= 2 ? lane - 2 : lane);
                        xm2 = A::sel(t >= 2, f, A::zero());
                    }
                }
                if (d.ku >= 1) {
                    T f = lane_fetch(x[s][0], t + 1 < L ? lane + 1 : lane);
                    xp1 = A::sel(t + 1 < L, f, A::zero());
                }
                if (d.ku >= 2) {
                    if constexpr (R >= 2) {
                        T f = lane_fetch(x[s][1], t + 1 < L ? lane + 1 : lane);
                        xp2 = A::sel(t + 1 < L, f, A::zero());
                    } else {
                        T f = lane_fetch(x[s][0], t + 2 < L ? lane + 2 : lane);
                        xp2 = A::sel(t + 2 < L, f, A::zero());
                    }
                }
                T pm2 = xm2, pm1 = xm1;
#pragma unroll
                for (int i = 0; i < R; i++) {
                    Pair<T> c0 = lds_pair<T, R, L>(tab, 0, i, t);   // vm2, vm1
                    Pair<T> c1 = lds_pair<T, R, L>(tab, 1, i, t);   // v0, vp1
                    T vp2 = tab[tab_coef(2, i, t, 0, R, L)];
                    const T xi = x[s][i];
                    const T n1 = (i + 1 < R) ? x[s][(i + 1) % R] : xp1;
                    const T n2 = (i + 2 < R) ? x[s][(i + 2) % R] : ((i + 2 == R) ? xp1 : xp2);
                    T u = A::mul(c1.a, xi);
                    u = A::madd(u, c0.b, pm1);
                    u = A::madd(u, c0.a, pm2);
                    u = A::madd(u, c1.b, n1);
                    u = A::madd(u, vp2, n2);
                    pm2 = pm1;
                    pm1 = xi;
                    x[s][i] = u;
                }
            }
        }

        // ---- forward recurrence w_r = u_r - l_r w_{r-1} ----
        if (d.fwd) {
            T lr[R];
#pragma unroll
            for (int i = 0; i < R; i++) lr[i] = tab[tab_coef(2, i, t, 1, R, L)];
            T acc[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T a = A::zero();
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
                    T ap = lane_fetch(acc[s], src);
                    acc[s] = A::sel(t >= sft, A::madd(acc[s], bk, ap), acc[s]);
                }
            }
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T c = lane_fetch(acc[s], t >= 1 ? lane - 1 : lane);
                T w = A::sel(t >= 1, c, A::zero());
#pragma unroll
                for (int i = 0; i < R; i++) {
                    w = A::msub(x[s][i], lr[i], w);
                    x[s][i] = w;
                }
            }
        }

        // ---- backward recurrences z_r = g_r w_r - h_r z_{r+1}, then y_r = D_r z_r - E_r y_{r+1} ----
#pragma unroll
        for (int stage = 0; stage < 2; stage++) {
            if (stage == 0 && !d.b1) continue;
            if (stage == 1 && !d.b2) continue;
            const int p = stage == 0 ? 3 : 4;
            T gg[R], hh[R];
#pragma unroll
            for (int i = 0; i < R; i++) {
                Pair<T> c = lds_pair<T, R, L>(tab, p, i, t);
                gg[i] = c.a;
                hh[i] = c.b;
            }
            T acc[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T a = A::zero();
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
                    T ap = lane_fetch(acc[s], src);
                    acc[s] = A::sel(t + sft < L, A::madd(acc[s], bk, ap), acc[s]);
                }
            }
#pragma unroll
            for (int s = 0; s < 2; s++) {
                T c = lane_fetch(acc[s], t + 1 < L ? lane + 1 : lane);
                T z = A::sel(t + 1 < L, c, A::zero());
#pragma unroll
                for (int i = R - 1; i >= 0; i--) {
                    z = A::msub(x[s][i], hh[i], z);
                    x[s][i] = z;
                }
            }
        }

        // ---- row-0 scaleC (pfSolve_TridiagonalSolve.h:1918-1927) ----
        if (d.apply_scale && t == 0) {
            const T sc = A::from(d.scale_hi, d.scale_lo);
#pragma unroll
            for (int s = 0; s < 2; s++) x[s][0] = A::mul(x[s][0], sc);
        }

        // ---- store (rows in [out_lo, out_hi) only) ----
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int64_t sys = sys0 + s;
            const bool ok = sys < d.nsys;
            const int64_t base = sys * d.stride + d.offSol;
            if constexpr (sizeof(T) == 8 && (R % 2) == 0) {
                if (d.vec2) {
#pragma unroll
                    for (int i = 0; i < R; i += 2) {
                        const int r = r0 + i;
                        const bool k0 = ok && r >= d.out_lo && r < d.out_hi;
                        const bool k1 = ok && r + 1 >= d.out_lo && r + 1 < d.out_hi;
                        double* p = reinterpret_cast<double*>(d.rhs) + base + r;
                        if (k0 && k1) {
                            double2 v;
                            v.x = x[s][i];
                            v.y = x[s][i + 1];
                            *reinterpret_cast<double2*>(p) = v;
                        } else {
                            if (k0) p[0] = x[s][i];
                            if (k1) p[1] = x[s][i + 1];
                        }
                    }
                    continue;
                }
            }
            const int64_t rs = d.row_step > 1 ? d.row_step : 1;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = r0 + i;
                if (ok && r >= d.out_lo && r < d.out_hi) mem_store<T>(d.rhs, base + r * rs, x[s][i], mf);
            }
        }
    }
}

// ------------------------------------------------------------------------
// block plan kernel (FP64): PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_block.h:37-115
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void jw_block_kernel(jw_block_desc d)
{
    const double* in = reinterpret_cast<const double*>(d.in);
    double* out = reinterpret_cast<double*>(d.out);
    const int64_t total = d.N * d.B;
    const int L = d.layout;
    const bool in_complex = (L == 2 || L == 3 || L == 4);
    const bool out_complex = (L == 5 || L == 6 || L == 7);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t xr = e % d.N, y = e / d.N;
        double re = 0.0, im = 0.0;
        if (!d.skip_read && xr >= d.in_lo && xr < d.in_hi) {
            int64_t idx = y * d.inStride + d.offIn;
            if (L == 7) idx *= 2;
            idx += xr;
            if (in_complex) {
                double2 v = reinterpret_cast<const double2*>(in)[idx];
                re = v.x;
                im = v.y;
            } else if (L == 6) {
                im = in[idx];
            } else if (L == 7) {
                re = in[idx];
                im = in[idx + d.inStride];
            } else {
                re = in[idx];
            }
        }
        if (!d.skip_read && d.op >= 1 && d.op <= 4) {
            const double f = jw_block_scale(d.op, xr, d.lshift, d.scale);
            re *= f;
            im *= f;
        }
        if (!(xr >= d.out_lo && xr < d.out_hi)) continue;
        int64_t o = y * d.outStride + d.offOut;
        if (L == 4) o *= 2;
        o += xr;
        if (L == 2) {
            if (d.accumulate) re += out[o];
            out[o] = re;
        } else if (L == 3) {
            if (d.accumulate) im += out[o];
            out[o] = im;
        } else if (L == 4) {
            if (d.accumulate) {
                re += out[o];
                im += out[o + d.outStride];
            }
            out[o] = re;
            out[o + d.outStride] = im;
        } else if (out_complex) {
            double2* oc = reinterpret_cast<double2*>(out);
            if (d.accumulate) {
                double2 prev = oc[o];
                re += prev.x;
                im += prev.y;
            }
            double2 v;
            v.x = re;
            v.y = im;
            oc[o] = v;
        } else {
            if (d.accumulate) re += out[o];
            out[o] = re;
        }
    }
}

// ------------------------------------------------------------------------
// dgbmv plan kernel (FP64): y = A_band x in place
// PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_dgbmv.h:38-112,
// PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_MatVecMul.h:383-446
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void jw_gbmv_kernel(jw_gbmv_desc d)
{
    extern __shared__ double xs[];
    const double* A = reinterpret_cast<const double*>(d.mat);
    double* rhs = reinterpret_cast<double*>(d.rhs);
    for (int64_t sys = blockIdx.x; sys < d.nsys; sys += gridDim.x) {
        const int64_t base = sys * d.stride + d.offSol;
        __syncthreads();
        for (int64_t r = threadIdx.x; r < d.N; r += blockDim.x)
            xs[r] = (r >= d.in_lo && r < d.in_hi) ? rhs[base + r] : 0.0;
        __syncthreads();
        for (int64_t r = threadIdx.x; r < d.N; r += blockDim.x) {
            double acc = 0.0;
            for (int j = 0; j < d.LDA; j++) {
                const int64_t col = r + d.KU - j;
                if (col < 0 || col >= d.N) continue;
                acc = __builtin_fma(xs[col], A[d.offM + (int64_t)j * d.N + col], acc);
            }
            if (r >= d.out_lo && r < d.out_hi) rhs[base + r] = acc;
        }
    }
}

// ------------------------------------------------------------------------
// host-side launchers (extern "C")
// ------------------------------------------------------------------------
static thread_local char g_err[256];
static void set_err(const char* what, hipError_t e)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
}

typedef void (*main_fn)(jw_run_desc);

#define JW_F64_CASE(LL, RR) \
    if (L == LL && R == RR) return reinterpret_cast<const void*>(&jw_main_kernel<double, RR, LL>);
#define JW_DD_CASE(LL, RR) \
    if (L == LL && R == RR) return reinterpret_cast<const void*>(&jw_main_kernel<dd, RR, LL>);

static const void* pick_main(int prec, int L, int R)
{
    if (prec == JW_PREC_F64) {
        JW_F64_CASE(64, 1) JW_F64_CASE(64, 2) JW_F64_CASE(64, 3) JW_F64_CASE(64, 4)
        JW_F64_CASE(64, 5) JW_F64_CASE(64, 6) JW_F64_CASE(64, 7) JW_F64_CASE(64, 8)
        JW_F64_CASE(64, 9) JW_F64_CASE(64, 10) JW_F64_CASE(64, 11) JW_F64_CASE(64, 12)
        JW_F64_CASE(64, 13) JW_F64_CASE(64, 14) JW_F64_CASE(64, 15) JW_F64_CASE(64, 16)
        JW_F64_CASE(32, 1) JW_F64_CASE(32, 2) JW_F64_CASE(32, 4) JW_F64_CASE(32, 8)
        JW_F64_CASE(16, 1) JW_F64_CASE(16, 2) JW_F64_CASE(16, 4)
        JW_F64_CASE(8, 1) JW_F64_CASE(8, 2) JW_F64_CASE(8, 4) JW_F64_CASE(8, 8)
    } else {
        JW_DD_CASE(64, 1) JW_DD_CASE(64, 2) JW_DD_CASE(64, 3) JW_DD_CASE(64, 4)
        JW_DD_CASE(64, 5) JW_DD_CASE(64, 6) JW_DD_CASE(64, 7) JW_DD_CASE(64, 8)
        JW_DD_CASE(32, 1) JW_DD_CASE(32, 2) JW_DD_CASE(32, 4) JW_DD_CASE(32, 8)
        JW_DD_CASE(16, 1) JW_DD_CASE(16, 2) JW_DD_CASE(16, 4)
        JW_DD_CASE(8, 1) JW_DD_CASE(8, 2) JW_DD_CASE(8, 4) JW_DD_CASE(8, 8)
    }
    return nullptr;
}

// ------------------------------------------------------------------------
// device scratch of the allocation-free entry point (pfsolve_append_noalloc):
// the coefficient table of one launch and its device error flag, one copy per
// GPU for the life of the process.  Host code never passes `scratch` itself to
// a kernel: its device address per GPU comes from hipGetSymbolAddress.
// ------------------------------------------------------------------------
constexpr size_t kNoAllocTableBytes = 144 * 1024;   // largest table: R = 32, 140 KiB
constexpr size_t kNoAllocBytes = kNoAllocTableBytes + 256;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static void* g_scratch_ptr[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

extern "C" {

/* Device address of this GPU's scratch (the current device), resolved once per device. */
int jwdev_noalloc_scratch(void** table, int** err, size_t* table_bytes)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_err("hipGetDevice", e); return -1; }
    if (dev < 0 || dev >= kNoAllocMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) { set_err("hipGetSymbolAddress(scratch)", e); return -2; }
        g_scratch_ptr[dev] = ptr;
    }
    *table = g_scratch_ptr[dev];
    *err = reinterpret_cast<int*>(static_cast<unsigned char*>(g_scratch_ptr[dev]) + kNoAllocTableBytes);
    *table_bytes = kNoAllocTableBytes;
    return 0;
}

const char* jwdev_last_error(void) { return g_err; }
void jwdev_set_error(const char* msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

int jwdev_supported(int L, int R, int precision) { return pick_main(precision, L, R) != nullptr; }

int64_t jwdev_table_elems(int L, int R)
{
    return (int64_t)JW_NCOEF * R * L + 4LL * ilog2_rt(L) * L;
}

int64_t jwdev_lds_bytes(int L, int R, int precision)
{
    return jwdev_table_elems(L, R) * (precision == JW_PREC_F64 ? 8 : 16);
}

int jwdev_device_info(int* cu_count, int* lds_per_block, int* arch_major, int* arch_minor)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_err("hipGetDevice", e); return -1; }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) { set_err("hipGetDeviceProperties", e); return -1; }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (lds_per_block) *lds_per_block = (int)prop.sharedMemPerBlock;
    if (arch_major) *arch_major = prop.major;
    if (arch_minor) *arch_minor = prop.minor;
    return 0;
}

int jwdev_prologue(const jw_source_desc* d, const void* matrices, void* table, int L, int R,
                   int precision, int format, void* stream)
{
    const int RL = R * L;
    const size_t elt = precision == JW_PREC_F64 ? 8 : 16;
    const size_t lds = prologue_lds_bytes(RL, elt);
    if (lds > 160 * 1024 || RL > kProRows * kPrologueThreads) {
        snprintf(g_err, sizeof(g_err), "prologue: too many rows");
        return -1;
    }
    if (format == 2 && L != 64) { snprintf(g_err, sizeof(g_err), "format 2 needs 64 lanes"); return -1; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // a build of at most one row per thread takes the small-array variant (same table)
    const bool small = RL <= kPrologueThreads;
    const void* fn = precision == JW_PREC_F64
                         ? (small ? reinterpret_cast<const void*>(&jw_prologue_kernel<double, true>)
                                  : reinterpret_cast<const void*>(&jw_prologue_kernel<double, false>))
                         : (small ? reinterpret_cast<const void*>(&jw_prologue_kernel<dd, true>)
                                  : reinterpret_cast<const void*>(&jw_prologue_kernel<dd, false>));
    hipError_t ea = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (ea != hipSuccess) { set_err("hipFuncSetAttribute(prologue)", ea); return -2; }
    void* args[] = {const_cast<jw_source_desc*>(d), const_cast<void**>(&matrices), &table, &L, &R, &format};
    hipError_t el = hipLaunchKernel(fn, dim3(1), dim3(kPrologueThreads), args, lds, s);
    if (el != hipSuccess) { set_err("prologue launch", el); return -1; }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_err("prologue launch", e); return -1; }
    return 0;
}

int jwdev_prologue_chain(const jw_chain_jobs* jobs, void* stream)
{
    const int RL = jobs->R * 64;
    if (jobs->K < 1 || jobs->K > JW_CHAIN_MAX || RL > kProRows * 256) {
        snprintf(g_err, sizeof(g_err), "chain prologue: bad job count or rows");
        return -1;
    }
    const size_t lds = prologue_lds_bytes(RL, 8);
    const void* fn = reinterpret_cast<const void*>(&jw_prologue_chain_kernel);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { set_err("hipFuncSetAttribute(chain prologue)", e); return -2; }
    hipLaunchKernelGGL(jw_prologue_chain_kernel, dim3(jobs->K), dim3(256), lds, reinterpret_cast<hipStream_t>(stream),
                       *jobs);
    e = hipGetLastError();
    if (e != hipSuccess) { set_err("chain prologue launch", e); return -1; }
    return 0;
}

int jwdev_prologue_batch(const jw_prologue_job* d_jobs, int njobs, int max_RL, int precision, int* d_counters,
                         int ncounters, void* stream)
{
    const size_t elt = precision == JW_PREC_F64 ? 8 : 16;
    const size_t lds = prologue_lds_bytes(max_RL, elt);
    if (lds > 160 * 1024 || max_RL > kProRows * kPrologueThreads) {
        snprintf(g_err, sizeof(g_err), "prologue: too many rows");
        return -1;
    }
    if (njobs < 1) return 0;
    const bool narrow = max_RL <= kProRows * 256;
    const int nt = narrow ? 256 : kPrologueThreads;
    if (ncounters > nt) { snprintf(g_err, sizeof(g_err), "too many counters"); return -1; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const void* fn;
    if (precision == JW_PREC_F64)
        fn = narrow ? reinterpret_cast<const void*>(&jw_prologue_batch_kernel<double, 256>)
                    : reinterpret_cast<const void*>(&jw_prologue_batch_kernel<double, kPrologueThreads>);
    else
        fn = narrow ? reinterpret_cast<const void*>(&jw_prologue_batch_kernel<dd, 256>)
                    : reinterpret_cast<const void*>(&jw_prologue_batch_kernel<dd, kPrologueThreads>);
    hipError_t ea = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (ea != hipSuccess) { set_err("hipFuncSetAttribute(prologue batch)", ea); return -2; }
    const jw_prologue_job* jobs = d_jobs;
    int zero = 0;
    void* args[] = {const_cast<jw_prologue_job**>(&jobs), &zero, &d_counters, &ncounters};
    hipError_t el = hipLaunchKernel(fn, dim3(njobs), dim3(nt), args, lds, s);
    if (el != hipSuccess) { set_err("prologue batch launch", el); return -1; }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_err("prologue batch launch", e); return -1; }
    return 0;
}

int jwdev_main(const jw_run_desc* r, int L, int R, int precision, int waves, int grid, void* stream)
{
    /* a launch without its right-hand sides or table is a host bug: an error, not a fault */
    if (!r->rhs || !r->table) { snprintf(g_err, sizeof(g_err), "main: no rhs or table"); return -1; }
    const void* f = pick_main(precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD, L, R);
    if (!f) { snprintf(g_err, sizeof(g_err), "no kernel for L=%d R=%d prec=%d", L, R, precision); return -1; }
    const size_t lds = (size_t)jwdev_lds_bytes(L, R, precision);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { set_err("hipFuncSetAttribute", e); return -2; }
    if (waves < 1 || waves > 8) waves = 4;
    void* args[] = {const_cast<jw_run_desc*>(r)};
    e = hipLaunchKernel(f, dim3(grid), dim3(64 * waves), args, lds, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) { set_err("main launch", e); return -3; }
    return 0;
}

int jwdev_main_occupancy(int L, int R, int precision, int waves)
{
    const void* f = pick_main(precision == JW_PREC_F64 ? JW_PREC_F64 : JW_PREC_DD, L, R);
    if (!f) return 0;
    const size_t lds = (size_t)jwdev_lds_bytes(L, R, precision);
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * waves, lds);
    if (e != hipSuccess) { set_err("occupancy", e); return 0; }
    return nb;
}

int jwdev_block(const jw_block_desc* d, void* stream)
{
    const int64_t total = d->N * d->B;
    int64_t grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(jw_block_kernel, dim3((unsigned)grid), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), *d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_err("block launch", e); return -1; }
    return 0;
}

int jwdev_gbmv(const jw_gbmv_desc* d, void* stream)
{
    int64_t grid = d->nsys < 8192 ? d->nsys : 8192;
    if (grid < 1) grid = 1;
    const size_t lds = (size_t)d->N * 8;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&jw_gbmv_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { set_err("hipFuncSetAttribute(gbmv)", e); return -2; }
    hipLaunchKernelGGL(jw_gbmv_kernel, dim3((unsigned)grid), dim3(256), lds,
                       reinterpret_cast<hipStream_t>(stream), *d);
    e = hipGetLastError();
    if (e != hipSuccess) { set_err("gbmv launch", e); return -1; }
    return 0;
}

}  // extern "C"
