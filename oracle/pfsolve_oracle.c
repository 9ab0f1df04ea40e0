/*
 * pfsolve_oracle.c -- CPU oracle for the JW step (TEST INFRASTRUCTURE ONLY).
 * See pfsolve_oracle.h for the reference files each part restates.
 *
 * Parity pins (tests/test_oracle.py): reorder_i0 against the reference's own
 * header compiled under oracle/_ref; Thomas against dense __float128 solves,
 * manufactured solutions and the tridiag(-1,2,-1) closed form; DD primitives
 * against exact error-free-transform identities and __float128; block scale
 * operators against their closed forms in long double.
 */
#include "pfsolve_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <quadmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef __float128 f128;

/* ------------------------------------------------------------------ */
/* random numbers                                                      */
/* ------------------------------------------------------------------ */
uint64_t pfo_splitmix64(uint64_t* state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double pfo_uniform(uint64_t* state, double a, double b)
{
    double u = (double)(pfo_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
    return a + (b - a) * u;
}

void pfo_fill_uniform(uint64_t seed, double a, double b, double* out, int64_t n)
{
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) out[i] = pfo_uniform(&s, a, b);
}

/* ------------------------------------------------------------------ */
/* layout maths                                                        */
/* ------------------------------------------------------------------ */
/* restates reorder_i0, PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:66-142 (HIP branch) */
int64_t pfo_reorder_i0(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0)
{
    if (Msplit0 == 0) Msplit0 = 704;
    int64_t size0 = N;
    int64_t size1 = Nalign ? Nalign : N;
    if (N > 3840) {
        size0 = Msplit0;
        size1 = Msplit0;
    }
    int64_t ws = ((size1 + 1023) / 1024) * 64;
    if (size1 < 512) ws = 64;
    if (ws < 64) ws = 64;
    if (ws % 64) ws = 64 * ((ws + 63) / 64);
    int64_t R = (size1 + ws - 1) / ws;
    if (R % 2 == 0) R++;
    int64_t ret_off = (i / size0) * size0;
    int64_t li = i % size0;
    if (((i / size0) + 1) * size0 > N) size0 = N % size0;
    int64_t ret = li / R;
    int64_t ceilq = (size0 + R - 1) / R;
    if ((li % R) < (size0 % R))
        ret += (li % R) * ceilq;
    else
        ret += (size0 % R) * ceilq + ((li % R) - (size0 % R)) * (size0 / R);
    return ret + ret_off;
}

/* rows per lane and lanes of the ALT layout (HIP branch of get_ALT_stride and
 * reorder_i_alt, PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:30-64,144-219) */
static void alt_geometry(int64_t size1, int64_t* W, int64_t* R)
{
    int64_t ws = ((size1 + 1023) / 1024) * 64;
    if (size1 < 512) ws = 64;
    if (ws < 64) ws = 64;
    if (ws % 64) ws = 64 * ((ws + 63) / 64);
    int64_t r = (size1 + ws - 1) / ws;
    if (r % 2 == 0) r++;
    *W = ws;
    *R = r;
}

/* block stride of one diagonal in the ALT storage: rows per lane x lanes */
int64_t pfo_alt_stride(int64_t Nalign, int64_t Msplit0)
{
    if (Msplit0 == 0) Msplit0 = 704;
    int64_t size1 = Nalign > 3840 ? Msplit0 : Nalign;
    int64_t W, R;
    alt_geometry(size1, &W, &R);
    return W * R;
}

/* storage position of row i in the ALT layout: row i belongs to the subsystem of its
 * parity (lanes [0, W/2) even, [W/2, W) odd), as row k = i/2 of that subsystem, held by
 * lane k / R of the half in register k % R */
int64_t pfo_reorder_i_alt(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0)
{
    if (Msplit0 == 0) Msplit0 = 704;
    int64_t size0 = N, size1 = Nalign ? Nalign : N;
    if (N > 3840) { size0 = Msplit0; size1 = Msplit0; }
    int64_t W, R;
    alt_geometry(size1, &W, &R);
    const int64_t ret_off = (i / size0) * size0;
    const int64_t li = i % size0;
    const int64_t k = li / 2;
    return (k % R) * W + (li % 2) * (W / 2) + k / R + ret_off;
}

/* PfSolve/pfSolve_PlanManagement/pfSolve_Plans/pfSolve_Plan_JonesWorland.h:274-283,311-347 */
int pfo_jw_registers_per_thread(int64_t N, int K, int upperBanded)
{
    int64_t tempM = N;
    if (!upperBanded) tempM += K - 1;
    int64_t lws = ((tempM + 1023) / 1024) * 64;
    if (tempM < 512) lws = 64;
    if (lws < 64) lws = 64;
    int64_t numWarps = lws / 64;
    if (numWarps > 16) numWarps = 16;
    int64_t R = (tempM + 64 * numWarps - 1) / (64 * numWarps);
    if (R % 2 == 0) R++;
    return (int)R;
}

/* pfSolve_Plan_JonesWorland.h:305-306 */
int pfo_jw_uses_parallel_thomas(int64_t N, int K, int jw_type)
{
    if (jw_type > 1000) return 0;
    if (N <= 64 && K == 1) return 0;
    return 1;
}

/* PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel0/pfSolve_Zeropad.h:306-378 */
int pfo_zeropad_keep(int64_t row, int64_t zp0, int64_t zp1, int runtime)
{
    if (runtime) return (uint64_t)row >= (uint64_t)zp0 && (uint64_t)row < (uint64_t)zp1;
    if (zp0) {
        if (row < zp0) return 0;
        if (zp1 && row >= zp1) return 0;
        return 1;
    }
    if (zp1) return row < zp1;
    return 1;
}

/* ------------------------------------------------------------------ */
/* problem resolution: which diagonal sits where, per mode             */
/* ------------------------------------------------------------------ */
enum { MV_NONE = 0, MV_DIAG, MV_UPPER, MV_LOWER };
enum { SV_NONE = 0, SV_LOWER, SV_UPPER, SV_TRI };

typedef struct jw_resolved {
    int mv, sv, band;
    int64_t mdA, udA, ldA;          /* matvec diagonals */
    int64_t ldS, udS, mdS;          /* solve diagonals (mdS: T pre-scale, fused) */
    int64_t offV;                   /* band V (fused) */
    int64_t mdM, udM;               /* bidiagonal M (fused) */
    int64_t offSol, stride;
    int64_t izp0, izp1, ozp0, ozp1;
    int izp_rt, ozp_rt, scale_rt;
    double scaleC;
    int64_t* P;                     /* storage permutation of matrix diagonals */
    int alt;                        /* ALT: rows couple with rows +-2 (two interleaved systems) */
} jw_resolved;

static void free_resolved(jw_resolved* r) { free(r->P); r->P = NULL; }

/* Offsets per mode: PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:117-348 */
static int resolve_problem(const pfo_jw_problem* p, jw_resolved* rs)
{
    memset(rs, 0, sizeof(*rs));
    const int64_t N = p->N;
    if (N < 1) return -1;
    int natural = p->jw_type > 1000;
    int jwt = p->jw_type % 1000;
    /* hundreds digit 1 = ALT (pfSolve_Plan_JonesWorland.h:265) */
    const int alt = jwt / 100 == 1;
    if (alt) jwt -= 100;
    if (jwt < 10 || jwt > 14) return -2;
    int jw = jwt % 10;
    const int64_t m = p->mask;
    int rtM = (m & PFO_RUNTIME_OFFSETM) != 0, rtV = (m & PFO_RUNTIME_OFFSETV) != 0;
    /* ALT restated for the parallel layout with baked offsets: each baked position
     * (0, N, 2N) becomes a multiple of the ALT block stride (pfSolve_JonesWorland.h:205-222) */
    if (alt && (natural || rtM || rtV || jw == 4)) return -8;
    rs->alt = alt;
    const int64_t NB = alt ? pfo_alt_stride(N, p->Msplit0) : N;   /* block stride of baked positions */
    int64_t offM = rtM ? p->run_offsetM : p->cfg_offsetM;
    int64_t offV = rtV ? p->run_offsetV : p->cfg_offsetV;
    int64_t offT = (m & PFO_RUNTIME_OFFSETT) ? p->run_offsetT : p->cfg_offsetT;
    rs->offSol = (m & PFO_RUNTIME_OFFSETSOLUTION) ? p->run_offsetSolution : p->cfg_offsetSolution;
    rs->stride = (m & PFO_RUNTIME_OUTPUTBUFFERSTRIDE) ? p->run_outStride : p->cfg_outStride;
    rs->izp_rt = (m & PFO_RUNTIME_INPUTZEROPAD) != 0;
    rs->ozp_rt = (m & PFO_RUNTIME_OUTPUTZEROPAD) != 0;
    rs->izp0 = rs->izp_rt ? p->run_izp[0] : p->cfg_izp[0];
    rs->izp1 = rs->izp_rt ? p->run_izp[1] : p->cfg_izp[1];
    rs->ozp0 = rs->ozp_rt ? p->run_ozp[0] : p->cfg_ozp[0];
    rs->ozp1 = rs->ozp_rt ? p->run_ozp[1] : p->cfg_ozp[1];
    rs->scale_rt = (m & PFO_RUNTIME_SCALEC) != 0;
    rs->scaleC = rs->scale_rt ? p->run_scaleC : p->cfg_scaleC;

    rs->P = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    if (!rs->P) return -100;
    int permuted = !natural && pfo_jw_uses_parallel_thomas(N, 1, p->jw_type);
    for (int64_t r = 0; r < N; r++)
        rs->P[r] = alt ? pfo_reorder_i_alt(r, N, N, p->Msplit0) : (permuted ? pfo_reorder_i0(r, N, N, p->Msplit0) : r);

    const int ub = p->upperBanded;
    if (jw == 4) {
        /* fused step: dgbmv(V) -> JW13 tridiagonal (T) -> JW13 upper bidiagonal (M) */
        rs->band = 1;
        rs->offV = offV;
        rs->ldS = offT; rs->mdS = offT + N; rs->udS = offT + 2 * N;
        rs->sv = SV_TRI;
        rs->mdM = offM; rs->udM = offM + N;
        return 0;
    }
    if (ub == 2) {
        if (jw != 3) { free_resolved(rs); return -3; }   /* only diagonal matvec + tri solve is defined */
        rs->ldS = 0; rs->mdA = NB; rs->udS = 2 * NB;
        rs->mv = MV_DIAG; rs->sv = SV_TRI;
        return 0;
    }
    if (ub != 0 && ub != 1) { free_resolved(rs); return -4; }
    if (jw == 0 || jw == 1) {
        if (ub) { rs->mv = MV_UPPER; rs->udA = rtM ? offM : NB; rs->mdA = rtM ? offM + N : 0; }
        else    { rs->mv = MV_LOWER; rs->mdA = rtM ? offM : NB; rs->ldA = rtM ? offM + N : 0; }
    }
    if (jw == 0 || jw == 2) {
        if (ub) { rs->sv = SV_LOWER; rs->ldS = rtV ? offV : 2 * NB; }
        else    { rs->sv = SV_UPPER; rs->udS = rtV ? offV : 2 * NB; }
    }
    if (jw == 3) {
        /* upperBanded is toggled for the load (pfSolve_JonesWorland.h:124-127,202-205) */
        rs->mv = MV_DIAG;
        if (ub) { rs->sv = SV_LOWER; rs->mdA = rtM ? offM : NB; rs->ldS = rtM ? offM + N : 0; }
        else    { rs->sv = SV_UPPER; rs->udS = rtM ? offM : NB; rs->mdA = rtM ? offM + N : 0; }
    }
    return 0;
}

/* One iteration of a chained step (numConsecutiveJWIterations = K > 1,
 * PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:108-330): iteration i
 * has N_i = N -/+ i rows (upperBanded 1: shrinking, 0: growing, :110-113); its matvec is
 * rectangular, N_i rows by N_{i-1} columns; the matvec diagonal is applied only in the first
 * (upperBanded 1) or the last (upperBanded 0) iteration and is an implicit 1 elsewhere
 * (pfSolve_MatVecMul.h:74,98; loads pfSolve_ReadWrite.h:1650-1660); baked positions change
 * with i (pfSolve_JonesWorland.h:159-168,189-198 for the matvec, :291-296,320-324 for the
 * solve), runtime ones use the current N_i.  Matrices are stored in reorder_i0 order of
 * N_i rows with the register layout of the largest system (Nalign = N, or N + K - 1 when
 * growing; pfSolve_Plan_JonesWorland.h:277-282).  Modes restated: jw 10, 11, 12 with
 * upperBanded 0 or 1. */
typedef struct jw_iter {
    int64_t N, Ncols;
    int mv, sv, unit;
    int64_t mdA, udA, ldA, ldS, udS;
    int64_t* P;
} jw_iter;

static int resolve_iteration(const pfo_jw_problem* p, int K, int i, jw_iter* it)
{
    memset(it, 0, sizeof(*it));
    const int natural = p->jw_type > 1000;
    const int jw = (p->jw_type % 1000) % 10;
    const int ub = p->upperBanded;
    if ((p->jw_type % 1000) < 10 || (p->jw_type % 1000) > 12 || (ub != 0 && ub != 1) || K < 1) return -5;
    const int64_t N0 = p->N;
    it->N = ub ? N0 - i : N0 + i;
    if (it->N < 1) return -6;
    it->Ncols = i == 0 ? N0 : (ub ? it->N + 1 : it->N - 1);
    const int64_t Nalign = ub ? N0 : N0 + K - 1;
    const int64_t m = p->mask;
    const int rtM = (m & PFO_RUNTIME_OFFSETM) != 0, rtV = (m & PFO_RUNTIME_OFFSETV) != 0;
    const int64_t offM = rtM ? p->run_offsetM : p->cfg_offsetM;
    const int64_t offV = rtV ? p->run_offsetV : p->cfg_offsetV;
    const int64_t Ni = it->N;
    it->P = (int64_t*)malloc(sizeof(int64_t) * (size_t)Ni);
    if (!it->P) return -100;
    const int permuted = !natural && pfo_jw_uses_parallel_thomas(N0, K, p->jw_type);
    for (int64_t r = 0; r < Ni; r++) it->P[r] = permuted ? pfo_reorder_i0(r, Ni, Nalign, p->Msplit0) : r;
    if (jw == 0 || jw == 1) {
        if (ub) {
            it->mv = MV_UPPER;
            if (rtM) { it->udA = offM; it->mdA = offM + Ni; it->unit = i != 0; }
            else if (i == 0) { it->mdA = 0; it->udA = Ni; }
            else { it->udA = 0; it->unit = 1; }
        } else {
            it->mv = MV_LOWER;
            if (rtM) { it->mdA = offM; it->ldA = offM + Ni; it->unit = i != K - 1; }
            else if (i == K - 1) { it->ldA = 0; it->mdA = Ni; }
            else { it->ldA = 0; it->unit = 1; }
        }
    }
    if (jw == 0 || jw == 2) {
        if (ub) { it->sv = SV_LOWER; it->ldS = rtV ? offV : (i == 0 ? 2 * Ni : Ni); }
        else    { it->sv = SV_UPPER; it->udS = rtV ? offV : (i == K - 1 ? 2 * Ni : Ni); }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* double-double arithmetic (Dekker 1971; Knuth TAOCP 4.2.2 TwoSum;    */
/* Joldes, Muller, Popescu, ACM TOMS 44(2) 2017, Alg. 6, 12, 17)       */
/* ------------------------------------------------------------------ */
pfo_dd pfo_two_sum(double a, double b)
{
    pfo_dd r;
    double s = a + b;
    double bb = s - a;
    r.hi = s;
    r.lo = (a - (s - bb)) + (b - bb);
    return r;
}

static pfo_dd fast_two_sum(double a, double b)
{
    pfo_dd r;
    double s = a + b;
    r.hi = s;
    r.lo = b - (s - a);
    return r;
}

pfo_dd pfo_two_prod(double a, double b)
{
    pfo_dd r;
    r.hi = a * b;
    r.lo = fma(a, b, -r.hi);
    return r;
}

pfo_dd pfo_dd_add(pfo_dd x, pfo_dd y)
{
    pfo_dd s = pfo_two_sum(x.hi, y.hi);
    pfo_dd t = pfo_two_sum(x.lo, y.lo);
    double c = s.lo + t.hi;
    pfo_dd v = fast_two_sum(s.hi, c);
    double w = t.lo + v.lo;
    return fast_two_sum(v.hi, w);
}

pfo_dd pfo_dd_sub(pfo_dd x, pfo_dd y)
{
    pfo_dd ny = {-y.hi, -y.lo};
    return pfo_dd_add(x, ny);
}

pfo_dd pfo_dd_mul(pfo_dd x, pfo_dd y)
{
    pfo_dd c = pfo_two_prod(x.hi, y.hi);
    double tl0 = x.lo * y.lo;
    double tl1 = fma(x.hi, y.lo, tl0);
    double cl2 = fma(x.lo, y.hi, tl1);
    double cl3 = c.lo + cl2;
    return fast_two_sum(c.hi, cl3);
}

/* DW / DW (Joldes et al. Alg. 17): relative error <= 15u^2 + 56u^3 */
pfo_dd pfo_dd_div(pfo_dd x, pfo_dd y)
{
    double th = x.hi / y.hi;
    /* r = y * th (DW times FP, Alg. 9-style: exact product of hi plus lo term) */
    pfo_dd c = pfo_two_prod(y.hi, th);
    double cl = fma(y.lo, th, c.lo);
    pfo_dd r = fast_two_sum(c.hi, cl);
    double pih = x.hi - r.hi;
    double dl = x.lo - r.lo;
    double d = pih + dl;
    double tl = d / y.hi;
    return fast_two_sum(th, tl);
}

static inline f128 dd_to_q(pfo_dd a) { return (f128)a.hi + (f128)a.lo; }
static inline pfo_dd q_to_dd(f128 v)
{
    pfo_dd r;
    r.hi = (double)v;
    r.lo = (double)(v - (f128)r.hi);
    return r;
}
pfo_dd pfo_q_add(pfo_dd a, pfo_dd b) { return q_to_dd(dd_to_q(a) + dd_to_q(b)); }
pfo_dd pfo_q_mul(pfo_dd a, pfo_dd b) { return q_to_dd(dd_to_q(a) * dd_to_q(b)); }
pfo_dd pfo_q_div(pfo_dd a, pfo_dd b) { return q_to_dd(dd_to_q(a) / dd_to_q(b)); }
double pfo_dd_relerr(pfo_dd a, pfo_dd b)
{
    f128 qb = dd_to_q(b);
    f128 d = dd_to_q(a) - qb;
    if (d < 0) d = -d;
    if (qb < 0) qb = -qb;
    if (qb == 0) return (double)d;
    return (double)(d / qb);
}

static inline pfo_dd dd_from_double(double v) { pfo_dd r = {v, 0.0}; return r; }

/* ------------------------------------------------------------------ */
/* type instantiations of the JW core                                  */
/* ------------------------------------------------------------------ */
/* FP64 */
#define R_T double
#define IN_T double
#define OUT_T double
#define R_LOAD(v) (v)
#define R_STORE(dst, v) ((dst) = (v))
#define R_ADD(a, b) ((a) + (b))
#define R_SUB(a, b) ((a) - (b))
#define R_MUL(a, b) ((a) * (b))
#define R_DIV(a, b) ((a) / (b))
#define R_ZERO 0.0
#define R_ONE 1.0
#define R_FROM_DOUBLE(v) (v)
#define R_FN(name) name##_f64
#include "jw_core.inc"
#undef R_T
#undef IN_T
#undef OUT_T
#undef R_LOAD
#undef R_STORE
#undef R_ADD
#undef R_SUB
#undef R_MUL
#undef R_DIV
#undef R_ZERO
#undef R_ONE
#undef R_FROM_DOUBLE
#undef R_FN

/* __float128 on FP64 storage; the store keeps the result in a side array */
static double* g_q_lo_out;  /* set by pfo_jw_step_f128 (single-threaded use) */
static inline void q_store_f64(double* dst, f128 v, double* base_out)
{
    double hi = (double)v;
    *dst = hi;
    if (g_q_lo_out) g_q_lo_out[dst - base_out] = (double)(v - (f128)hi);
}
static double* g_q_base;
#define R_T f128
#define IN_T double
#define OUT_T double
#define R_LOAD(v) ((f128)(v))
#define R_STORE(dst, v) q_store_f64(&(dst), (v), g_q_base)
#define R_ADD(a, b) ((a) + (b))
#define R_SUB(a, b) ((a) - (b))
#define R_MUL(a, b) ((a) * (b))
#define R_DIV(a, b) ((a) / (b))
#define R_ZERO ((f128)0)
#define R_ONE ((f128)1)
#define R_FROM_DOUBLE(v) ((f128)(v))
#define R_FN(name) name##_f128
#include "jw_core.inc"
#undef R_T
#undef IN_T
#undef OUT_T
#undef R_LOAD
#undef R_STORE
#undef R_ADD
#undef R_SUB
#undef R_MUL
#undef R_DIV
#undef R_ZERO
#undef R_ONE
#undef R_FROM_DOUBLE
#undef R_FN

/* double-double arithmetic on DD storage */
#define R_T pfo_dd
#define IN_T pfo_dd
#define OUT_T pfo_dd
#define R_LOAD(v) (v)
#define R_STORE(dst, v) ((dst) = (v))
#define R_ADD(a, b) pfo_dd_add((a), (b))
#define R_SUB(a, b) pfo_dd_sub((a), (b))
#define R_MUL(a, b) pfo_dd_mul((a), (b))
#define R_DIV(a, b) pfo_dd_div((a), (b))
#define R_ZERO dd_from_double(0.0)
#define R_ONE dd_from_double(1.0)
#define R_FROM_DOUBLE(v) dd_from_double(v)
#define R_FN(name) name##_dd
#include "jw_core.inc"
#undef R_T
#undef IN_T
#undef OUT_T
#undef R_LOAD
#undef R_STORE
#undef R_ADD
#undef R_SUB
#undef R_MUL
#undef R_DIV
#undef R_ZERO
#undef R_ONE
#undef R_FROM_DOUBLE
#undef R_FN

/* __float128 on DD storage, result rounded to DD */
#define R_T f128
#define IN_T pfo_dd
#define OUT_T pfo_dd
#define R_LOAD(v) dd_to_q(v)
#define R_STORE(dst, v) ((dst) = q_to_dd(v))
#define R_ADD(a, b) ((a) + (b))
#define R_SUB(a, b) ((a) - (b))
#define R_MUL(a, b) ((a) * (b))
#define R_DIV(a, b) ((a) / (b))
#define R_ZERO ((f128)0)
#define R_ONE ((f128)1)
#define R_FROM_DOUBLE(v) ((f128)(v))
#define R_FN(name) name##_q_dd
#include "jw_core.inc"
#undef R_T
#undef IN_T
#undef OUT_T
#undef R_LOAD
#undef R_STORE
#undef R_ADD
#undef R_SUB
#undef R_MUL
#undef R_DIV
#undef R_ZERO
#undef R_ONE
#undef R_FROM_DOUBLE
#undef R_FN

int pfo_jw_step_f64(const pfo_jw_problem* p, const double* buffer, double* out)
{
    return jw_step_f64(p, buffer, out, 0, p->B);
}

int pfo_jw_step_f64_range(const pfo_jw_problem* p, const double* buffer, double* out,
                          int64_t b0, int64_t b1, int threads)
{
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
    int err = 0;
    int64_t chunk = (b1 - b0 + nt - 1) / nt;
#pragma omp parallel num_threads(nt) reduction(| : err)
    {
        int tid = omp_get_thread_num();
        int64_t s = b0 + (int64_t)tid * chunk;
        int64_t e = s + chunk < b1 ? s + chunk : b1;
        if (s < e) err |= jw_step_f64(p, buffer, out, s, e);
    }
    return err;
#else
    (void)threads;
    return jw_step_f64(p, buffer, out, b0, b1);
#endif
}

int pfo_jw_step_dd_range(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out, int64_t b0, int64_t b1,
                         int threads)
{
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
    int err = 0;
    int64_t chunk = (b1 - b0 + nt - 1) / nt;
#pragma omp parallel num_threads(nt) reduction(| : err)
    {
        int tid = omp_get_thread_num();
        int64_t s = b0 + (int64_t)tid * chunk;
        int64_t e = s + chunk < b1 ? s + chunk : b1;
        if (s < e) err |= jw_step_dd(p, buffer, out, s, e);
    }
    return err;
#else
    (void)threads;
    return jw_step_dd(p, buffer, out, b0, b1);
#endif
}

int pfo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int pfo_jw_step_f128(const pfo_jw_problem* p, const double* buffer, double* out, double* out_lo)
{
    g_q_lo_out = out_lo;
    g_q_base = out;
    int r = jw_step_f128(p, buffer, out, 0, p->B);
    g_q_lo_out = NULL;
    g_q_base = NULL;
    return r;
}

int pfo_jw_step_dd(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out)
{
    return jw_step_dd(p, buffer, out, 0, p->B);
}

int pfo_jw_chain_f64(const pfo_jw_problem* p, int K, const double* const* buffers, double* out)
{
    return jw_chain_f64(p, K, buffers, out, 0, p->B);
}

int pfo_jw_chain_f128(const pfo_jw_problem* p, int K, const double* const* buffers, double* out)
{
    g_q_lo_out = NULL;
    g_q_base = out;
    int r = jw_chain_f128(p, K, buffers, out, 0, p->B);
    g_q_base = NULL;
    return r;
}

int pfo_jw_step_f128_dd(const pfo_jw_problem* p, const pfo_dd* buffer, pfo_dd* out)
{
    return jw_step_q_dd(p, buffer, out, 0, p->B);
}

int pfo_dgbmv_f64(const pfo_jw_problem* p, int LDA, int KU, int KL, const double* buffer, double* out)
{
    return dgbmv_f64(p, LDA, KU, KL, buffer, out);
}

int pfo_dgbmv_f128(const pfo_jw_problem* p, int LDA, int KU, int KL, const double* buffer, double* out)
{
    g_q_lo_out = NULL;
    g_q_base = out;
    int r = dgbmv_f128(p, LDA, KU, KL, buffer, out);
    g_q_base = NULL;
    return r;
}

/* ------------------------------------------------------------------ */
/* block plan                                                          */
/* ------------------------------------------------------------------ */
/* runtime-scaleC branches of PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_RegistersScaleC.h:40-201 */
double pfo_block_scale_factor(int op, int64_t n, int lshift, double s)
{
    double x = (double)n, l = (double)lshift;
    switch (op) {
    case 1: return s;
    case 2: return 2.0 * (x + l + 1.0) * sqrt((x + 1.0) / (x + s + l + 1.5));
    case 3: return 4.0 * (x + l + 2.0) * (x + l + 3.0) *
                   sqrt((x + 1.0) * (x + 2.0) / ((x + s + l + 2.5) * (x + s + l + 3.5)));
    case 4: return (x + l + 1.0) * (4.0 * l + 6.0) * sqrt((x + 1.0) / (x + s + l + 1.5));
    default: return 1.0;
    }
}

long double pfo_block_scale_factor_ld(int op, int64_t n, int lshift, long double s)
{
    long double x = (long double)n, l = (long double)lshift;
    switch (op) {
    case 1: return s;
    case 2: return 2.0L * (x + l + 1.0L) * sqrtl((x + 1.0L) / (x + s + l + 1.5L));
    case 3: return 4.0L * (x + l + 2.0L) * (x + l + 3.0L) *
                   sqrtl((x + 1.0L) * (x + 2.0L) / ((x + s + l + 2.5L) * (x + s + l + 3.5L)));
    case 4: return (x + l + 1.0L) * (4.0L * l + 6.0L) * sqrtl((x + 1.0L) / (x + s + l + 1.5L));
    default: return 1.0L;
    }
}

/* PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/pfSolve_ReadWrite.h:2275-2371,
 * PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_block.h:82-106.
 * Complex buffers hold (re,im) pairs; complex element k of a complex buffer is doubles [2k, 2k+1]. */
int pfo_block_f64(const pfo_block_problem* p, const double* in, double* out)
{
    const int layout = p->block % 10;
    const int accumulate = ((p->block / 10) % 10) == 2;
    const int op = p->block / 100;
    if (layout < 1 || layout > 7) return -1;
    const int in_complex = (layout == 2 || layout == 3 || layout == 4);
    const int out_complex = (layout == 5 || layout == 6 || layout == 7);
    const int skip_read = (op == 1 && !p->scaleC_runtime && p->scaleC == 0.0);
    for (int64_t y = 0; y < p->B; y++) {
        for (int64_t xr = 0; xr < p->N; xr++) {
            double re = 0.0, im = 0.0;
            if (!skip_read && pfo_zeropad_keep(xr, p->izp[0], p->izp[1], p->izp_runtime)) {
                int64_t idx = y * p->inStride + p->offsetM;
                if (layout == 7) idx *= 2;
                idx += xr;
                if (in_complex) { re = in[2 * idx]; im = in[2 * idx + 1]; }
                else if (layout == 5) re = in[idx];
                else if (layout == 6) im = in[idx];
                else if (layout == 7) { re = in[idx]; im = in[idx + p->inStride]; }
                else re = in[idx];
            }
            if (!skip_read && op >= 1 && op <= 4) {
                double f = pfo_block_scale_factor(op, xr, p->lshift, p->scaleC);
                re *= f;
                im *= f;
            }
            if (!pfo_zeropad_keep(xr, p->ozp[0], p->ozp[1], p->ozp_runtime)) continue;
            int64_t o = y * p->outStride + p->offsetSolution;
            if (layout == 4) o *= 2;
            o += xr;
            if (layout == 2) { if (accumulate) re += out[o]; out[o] = re; }
            else if (layout == 3) { if (accumulate) im += out[o]; out[o] = im; }
            else if (layout == 4) {
                if (accumulate) { re += out[o]; im += out[o + p->outStride]; }
                out[o] = re;
                out[o + p->outStride] = im;
            } else if (out_complex) {
                if (accumulate) { re += out[2 * o]; im += out[2 * o + 1]; }
                out[2 * o] = re;
                out[2 * o + 1] = im;
            } else {
                if (accumulate) re += out[o];
                out[o] = re;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* synthetic workloads                                                 */
/* ------------------------------------------------------------------ */
void pfo_gen_raw_matrices(uint64_t seed, int64_t N, int KL, int KU,
                          double* Vband, double* Ta, double* Tb, double* Tc,
                          double* Md, double* Me)
{
    uint64_t s = seed;
    const int LDA = KL + KU + 1;
    for (int j = 0; j < LDA; j++)
        for (int64_t col = 0; col < N; col++) {
            double v = pfo_uniform(&s, -1.0, 1.0);
            int64_t r = col - KU + j;       /* row of band entry (j, col) */
            Vband[(int64_t)j * N + col] = (r >= 0 && r < N) ? v : 0.0;
        }
    for (int64_t r = 0; r < N; r++) {
        Ta[r] = pfo_uniform(&s, -0.5, 0.5);
        Tb[r] = pfo_uniform(&s, 2.0, 3.0);
        Tc[r] = pfo_uniform(&s, -0.5, 0.5);
    }
    Ta[0] = 0.0;
    Tc[N - 1] = 0.0;
    for (int64_t r = 0; r < N; r++) {
        Md[r] = pfo_uniform(&s, 1.0, 2.0);
        Me[r] = pfo_uniform(&s, -0.5, 0.5);
    }
    Me[N - 1] = 0.0;
}

/* pre-normalisation as the reference harness does it:
 * benchmark_scripts/vkFFT_scripts/src/sample_1_benchmark_VkFFT_double.cpp:2332-2337 */
void pfo_pack_fused(int64_t N, int KL, int KU, const double* Vband, const double* Ta,
                    const double* Tb, const double* Tc, const double* Md, const double* Me,
                    int permuted, int Msplit0, int64_t offV, int64_t offT, int64_t offM,
                    double* buffer)
{
    const int LDA = KL + KU + 1;
    for (int64_t k = 0; k < (int64_t)LDA * N; k++) buffer[offV + k] = Vband[k];
    for (int64_t r = 0; r < N; r++) {
        int64_t P = permuted ? pfo_reorder_i0(r, N, N, Msplit0) : r;
        double ib = 1.0 / Tb[r];
        buffer[offT + N + P] = ib;
        buffer[offT + P] = Ta[r] * ib;
        buffer[offT + 2 * N + P] = Tc[r] * ib;
        double id = 1.0 / Md[r];
        buffer[offM + P] = id;
        buffer[offM + N + P] = Me[r] * id;
    }
}

void pfo_pack_fused_dd(int64_t N, int KL, int KU, const pfo_dd* Vband, const pfo_dd* Ta,
                       const pfo_dd* Tb, const pfo_dd* Tc, const pfo_dd* Md, const pfo_dd* Me,
                       int permuted, int Msplit0, int64_t offV, int64_t offT, int64_t offM,
                       pfo_dd* buffer)
{
    const int LDA = KL + KU + 1;
    for (int64_t k = 0; k < (int64_t)LDA * N; k++) buffer[offV + k] = Vband[k];
    for (int64_t r = 0; r < N; r++) {
        int64_t P = permuted ? pfo_reorder_i0(r, N, N, Msplit0) : r;
        f128 ib = (f128)1 / dd_to_q(Tb[r]);
        buffer[offT + N + P] = q_to_dd(ib);
        buffer[offT + P] = q_to_dd(dd_to_q(Ta[r]) * ib);
        buffer[offT + 2 * N + P] = q_to_dd(dd_to_q(Tc[r]) * ib);
        f128 id = (f128)1 / dd_to_q(Md[r]);
        buffer[offM + P] = q_to_dd(id);
        buffer[offM + N + P] = q_to_dd(dd_to_q(Me[r]) * id);
    }
}

void pfo_fused_raw_f128(int64_t N, int KL, int KU, const double* Vband, const double* Ta,
                        const double* Tb, const double* Tc, const double* Md, const double* Me,
                        const double* x, double* y)
{
    f128* u = (f128*)malloc(sizeof(f128) * (size_t)N);
    f128* cp = (f128*)malloc(sizeof(f128) * (size_t)N);
    const int LDA = KL + KU + 1;
    for (int64_t r = 0; r < N; r++) {
        f128 acc = 0;
        for (int j = 0; j < LDA; j++) {
            int64_t col = r + KU - j;
            if (col < 0 || col >= N) continue;
            acc += (f128)Vband[(int64_t)j * N + col] * (f128)x[col];
        }
        u[r] = acc;
    }
    /* raw Thomas on T (sample_0_benchmark_VkFFT_single.cpp:437-458) */
    cp[0] = (f128)Tc[0] / (f128)Tb[0];
    u[0] = u[0] / (f128)Tb[0];
    for (int64_t j = 1; j < N; j++) {
        f128 den = (f128)Tb[j] - (f128)Ta[j] * cp[j - 1];
        cp[j] = (j + 1 < N ? (f128)Tc[j] : (f128)0) / den;
        u[j] = (u[j] - (f128)Ta[j] * u[j - 1]) / den;
    }
    for (int64_t j = N - 2; j >= 0; j--) u[j] -= cp[j] * u[j + 1];
    /* upper bidiagonal back-substitution on M */
    u[N - 1] = u[N - 1] / (f128)Md[N - 1];
    for (int64_t j = N - 2; j >= 0; j--) u[j] = (u[j] - (f128)Me[j] * u[j + 1]) / (f128)Md[j];
    for (int64_t r = 0; r < N; r++) y[r] = (double)u[r];
    free(u);
    free(cp);
}

/* Gaussian elimination with partial pivoting in __float128 (brute force for tiny N) */
void pfo_dense_solve_f128(int64_t N, const double* A, const double* b, double* x)
{
    f128* M = (f128*)malloc(sizeof(f128) * (size_t)(N * (N + 1)));
    for (int64_t i = 0; i < N; i++) {
        for (int64_t j = 0; j < N; j++) M[i * (N + 1) + j] = (f128)A[i * N + j];
        M[i * (N + 1) + N] = (f128)b[i];
    }
    for (int64_t k = 0; k < N; k++) {
        int64_t piv = k;
        f128 best = M[k * (N + 1) + k] < 0 ? -M[k * (N + 1) + k] : M[k * (N + 1) + k];
        for (int64_t i = k + 1; i < N; i++) {
            f128 v = M[i * (N + 1) + k] < 0 ? -M[i * (N + 1) + k] : M[i * (N + 1) + k];
            if (v > best) { best = v; piv = i; }
        }
        if (piv != k)
            for (int64_t j = 0; j <= N; j++) {
                f128 t = M[k * (N + 1) + j];
                M[k * (N + 1) + j] = M[piv * (N + 1) + j];
                M[piv * (N + 1) + j] = t;
            }
        for (int64_t i = k + 1; i < N; i++) {
            f128 f = M[i * (N + 1) + k] / M[k * (N + 1) + k];
            for (int64_t j = k; j <= N; j++) M[i * (N + 1) + j] -= f * M[k * (N + 1) + j];
        }
    }
    /* back substitution kept entirely in f128 */
    f128* xq = (f128*)malloc(sizeof(f128) * (size_t)N);
    for (int64_t i = N - 1; i >= 0; i--) {
        f128 s = M[i * (N + 1) + N];
        for (int64_t j = i + 1; j < N; j++) s -= M[i * (N + 1) + j] * xq[j];
        xq[i] = s / M[i * (N + 1) + i];
    }
    for (int64_t i = 0; i < N; i++) x[i] = (double)xq[i];
    free(xq);
    free(M);
}
