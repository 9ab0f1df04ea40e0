/*
 * pfsolve_oracle.h -- CPU oracle for the Jones-Worland (JW) step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU, what the
 * reference (PfSolve) computes on the JW path, in FP64, in __float128 and in
 * double-double.  It is the checker that the HIP path is compared against.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it; the product library (libpfsolve_gfx950.so) never links it and
 * never falls back to it.
 *
 * Reference files restated (paths relative to the reference root):
 *   PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142   reorder_i0
 *   PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:301-359  mode flags, geometry
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:117-348      matrix offsets per mode
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:1561-1733       matrix load (permuted)
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:1906-2012       RHS load/store
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel0/pfSolve_Zeropad.h:306-378           keep-range zero pad
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_MatVecMul.h:28-133          bidiagonal matvec
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_MatVecMul.h:383-446         band matvec (dgbmv)
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_TridiagonalSolve.h:1578-1929 unit-diagonal solve, row-0 scaleC
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:2275-2371       block layouts
 *   PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_RegistersScaleC.h:27-201    block scale operators
 *   benchmark_scripts/vkFFT_scripts/src/sample_0_benchmark_VkFFT_single.cpp:437-458  FP64 Thomas order
 */
#ifndef PFSOLVE_ORACLE_H
#define PFSOLVE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* double-double number: value = hi + lo, |lo| <= ulp(hi)/2 */
typedef struct pfo_dd {
    double hi;
    double lo;
} pfo_dd;

/* ---- random numbers: SplitMix64, U(a,b) from 53-bit draws ---- */
uint64_t pfo_splitmix64(uint64_t* state);
double pfo_uniform(uint64_t* state, double a, double b);
void pfo_fill_uniform(uint64_t seed, double a, double b, double* out, int64_t n);

/* ---- layout maths ---- */
int64_t pfo_reorder_i0(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0);
/* ALT layout (jw_type hundreds digit 1): block stride of a diagonal and position of row i */
int64_t pfo_alt_stride(int64_t Nalign, int64_t Msplit0);
int64_t pfo_reorder_i_alt(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0);
int pfo_jw_registers_per_thread(int64_t N, int K, int upperBanded);
int pfo_jw_uses_parallel_thomas(int64_t N, int K, int jw_type);
/* 1 when row is read/written; runtime=1 means the [zp0,zp1) pair came from the launch params */
int pfo_zeropad_keep(int64_t row, int64_t zp0, int64_t zp1, int runtime);

/* ---- reference JW problem description (one plan + one launch) ---- */
#define PFO_RUNTIME_OFFSETM 2
#define PFO_RUNTIME_OFFSETV 4
#define PFO_RUNTIME_OFFSETSOLUTION 8
#define PFO_RUNTIME_INPUTZEROPAD 16
#define PFO_RUNTIME_OUTPUTZEROPAD 32
#define PFO_RUNTIME_SCALEC 64
#define PFO_RUNTIME_OUTPUTBUFFERSTRIDE 256
#define PFO_RUNTIME_OFFSETT 512

typedef struct pfo_jw_problem {
    int jw_type;      /* 10..14, +1000 = natural matrix order */
    int upperBanded;  /* 0, 1, 2 */
    int64_t N;        /* rows per system */
    int64_t B;        /* number of real systems */
    int64_t mask;     /* jw_control_bitmask */
    int Msplit0;      /* 0 = HIP default 704 */
    int KL, KU;       /* band of V for the fused step (LDA = KL+KU+1) */
    /* plan-time values */
    int64_t cfg_offsetM, cfg_offsetV, cfg_offsetSolution, cfg_offsetT;
    int64_t cfg_izp[2], cfg_ozp[2];
    double cfg_scaleC;
    int64_t cfg_outStride;
    /* launch-time values (used where the mask bit is set) */
    int64_t run_offsetM, run_offsetV, run_offsetSolution, run_offsetT;
    int64_t run_izp[2], run_ozp[2];
    double run_scaleC;
    int64_t run_outStride;
} pfo_jw_problem;

/* Reference JW step, one chained iteration.  buffer = matrices (the plan's
 * `buffer`), out = RHS buffer (`outputBuffer`), solved in place.
 * Returns 0 on success, a negative value for an unsupported mode. */
int pfo_jw_step_f64(const pfo_jw_problem* p, const double* buffer, double* out);
/* Same operator evaluated in __float128 on the same stored FP64 inputs.
 * out receives the rounded FP64 result; out_lo (optional) the remainder. */
int pfo_jw_step_f128(const pfo_jw_problem* p, const double* buffer, double* out, double* out_lo);
/* Double-double arithmetic restatement, DD memory (16 bytes per element). */
int pfo_jw_step_dd(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out);
/* __float128 evaluation of DD inputs; the result is rounded to DD. */
int pfo_jw_step_f128_dd(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out);

/* Chained step (numConsecutiveJWIterations = K, jw 10/11/12, upperBanded 0/1): iteration i
 * reads its matrices from buffers[i] (the same pointer K times without
 * useMultipleInputBuffers); p->N is the first system's size. */
int pfo_jw_chain_f64(const pfo_jw_problem* p, int K, const double* const* buffers, double* out);
int pfo_jw_chain_f128(const pfo_jw_problem* p, int K, const double* const* buffers, double* out);

/* dgbmv plan: y = A_band x in place; A at offsetM (band-major, LDA rows). */
int pfo_dgbmv_f64(const pfo_jw_problem* p, int LDA, int KU, int KL, const double* buffer, double* out);
int pfo_dgbmv_f128(const pfo_jw_problem* p, int LDA, int KU, int KL, const double* buffer, double* out);

/* block plan: layout conversion + scale operator (+ accumulate), FP64.
 * Buffers are addressed in doubles; complex values are (re, im) pairs. */
typedef struct pfo_block_problem {
    int block;          /* e.g. 104 */
    int lshift;
    int64_t N;          /* size[0] */
    int64_t B;          /* size[1] */
    int64_t mask;
    int64_t offsetM, offsetSolution;   /* effective values */
    int64_t izp[2], ozp[2];
    int izp_runtime, ozp_runtime;
    double scaleC;
    int scaleC_runtime;
    int64_t inStride, outStride;
} pfo_block_problem;
int pfo_block_f64(const pfo_block_problem* p, const double* in, double* out);
/* scale factor of a block operator for row n (for closed-form tests) */
double pfo_block_scale_factor(int op, int64_t n, int lshift, double s);
long double pfo_block_scale_factor_ld(int op, int64_t n, int lshift, long double s);

/* ---- double-double primitives (exposed for known-answer tests) ---- */
pfo_dd pfo_dd_add(pfo_dd a, pfo_dd b);
pfo_dd pfo_dd_sub(pfo_dd a, pfo_dd b);
pfo_dd pfo_dd_mul(pfo_dd a, pfo_dd b);
pfo_dd pfo_dd_div(pfo_dd a, pfo_dd b);
/* exact error-free transforms */
pfo_dd pfo_two_sum(double a, double b);
pfo_dd pfo_two_prod(double a, double b);
/* __float128 helpers: r = a op b evaluated in f128, rounded to DD */
pfo_dd pfo_q_add(pfo_dd a, pfo_dd b);
pfo_dd pfo_q_mul(pfo_dd a, pfo_dd b);
pfo_dd pfo_q_div(pfo_dd a, pfo_dd b);
/* |a-b|/|b| evaluated in f128 */
double pfo_dd_relerr(pfo_dd a, pfo_dd b);

/* ---- synthetic inputs of the JW benchmark workloads ---- */
/* raw matrices: V band-major (KL+KU+1) x N (AB[j*N+col] = A(col-KU+j.. ), T sub/diag/super,
 * M diag/super, as drawn from the stated distributions */
void pfo_gen_raw_matrices(uint64_t seed, int64_t N, int KL, int KU,
                          double* Vband, double* Ta, double* Tb, double* Tc,
                          double* Md, double* Me);
/* pack raw matrices into the fused-step storage:
 *   V  -> buffer[offV + j*N + col]                       (dgbmv band-major, natural order)
 *   T  -> buffer[offT + {0,N,2N} + P(r)] = {a/b, 1/b, c/b} (JW tridiagonal block)
 *   M  -> buffer[offM + {0,N} + P(r)]    = {1/d, e/d}      (JW bidiagonal block)
 * with P = reorder_i0 when permuted != 0, identity otherwise. */
void pfo_pack_fused(int64_t N, int KL, int KU, const double* Vband, const double* Ta,
                    const double* Tb, const double* Tc, const double* Md, const double* Me,
                    int permuted, int Msplit0, int64_t offV, int64_t offT, int64_t offM,
                    double* buffer);
/* DD version of the packing: raw values are DD, normalisation in f128 rounded to DD */
void pfo_pack_fused_dd(int64_t N, int KL, int KU, const pfo_dd* Vband, const pfo_dd* Ta,
                       const pfo_dd* Tb, const pfo_dd* Tc, const pfo_dd* Md, const pfo_dd* Me,
                       int permuted, int Msplit0, int64_t offV, int64_t offT, int64_t offM,
                       pfo_dd* buffer);

/* y = M^{-1} T^{-1} V x on RAW matrices in __float128 (independent of storage
 * format and packing), for one system; x and y have N entries. */
void pfo_fused_raw_f128(int64_t N, int KL, int KU, const double* Vband, const double* Ta,
                        const double* Tb, const double* Tc, const double* Md, const double* Me,
                        const double* x, double* y);

/* dense tiny-N helpers for brute-force checks (N <= 64) */
void pfo_dense_solve_f128(int64_t N, const double* A /*row-major NxN*/, const double* b, double* x);

/* ---- timed CPU baseline: the FP64 restatement over many systems ---- */
/* Runs pfo_jw_step_f64 on systems [b0,b1) only; threads = OpenMP threads (0 = all). */
int pfo_jw_step_f64_range(const pfo_jw_problem* p, const double* buffer, double* out,
                          int64_t b0, int64_t b1, int threads);
int pfo_jw_step_dd_range(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out, int64_t b0, int64_t b1,
                         int threads);
int pfo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
