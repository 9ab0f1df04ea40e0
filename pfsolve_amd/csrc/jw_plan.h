/*
 * jw_plan.h -- plan state and the host-side resolution of reference modes.
 */
#ifndef PFSOLVE_JW_PLAN_H
#define PFSOLVE_JW_PLAN_H

#include "pfsolve_amd.h"
#include "jw_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PF_KIND_NONE = 0, PF_KIND_JW = 1, PF_KIND_BLOCK = 2, PF_KIND_GBMV = 3 };

struct pfsolve_plan {
    int kind;
    int precision;            /* JW_PREC_* */
    int64_t N;
    int L, R, waves;
    int cu_count, occupancy;
    void* table;              /* device: coefficient table, one per stream slot */
    size_t table_bytes;
    size_t table_stride;      /* bytes between the slots' tables */
    int64_t lds_bytes;
    /* wave-per-system FP64 path (table format 2): rows per lane (0 = not available),
     * waves per workgroup and workgroups per CU */
    int R64, waves64, occupancy64;
    /* compute / store split kernel for R64 <= 8 (0 = not used), its workgroups per
     * CU, and the device flag it raises when a bounded wait runs out */
    int split64, occupancy_split;
    int* d_err;
    int wait_polls;           /* jw_run_desc.wait_polls of the split kernel's launches */
    hipEvent_t ev_start, ev_stop;   /* optional timing events around the main kernel */
    /* wave-per-system double-double kernel (table format 2 in double-double): rows per
     * lane (0 = not used), waves per workgroup, workgroups per CU */
    int dd64_R, dd64_waves, dd64_occ;
    /* chained steps: iterations (1 = none), rows per lane of the chained kernel, its
     * workgroups per CU; a slot's table area holds K format-2 tables of chain_tab bytes */
    int chain_K, chain_R, chain_occ;
    size_t chain_tab;
    /* ALT variant: two parity subsystems, one format-1 table of alt_tab bytes each */
    int alt;
    size_t alt_tab;
    /* long systems cut into chunks of 64 long_R rows (long_R = 0: not available; long_only:
     * the only path of this plan).  A slot's area holds, from long_tab_off, long_ntab
     * format-1 tables of long_tab_elems elements, and from long_red_off the reduced data
     * for slices of long_red_sys systems; long_occ chunk workgroups fit one CU.  FP64 plans
     * run 16-byte aligned systems on the pair machinery instead (format-2 tables in the same
     * space, long64_occ of its workgroups per CU; 0: not available) */
    int long_R, long_only, long_occ, long64_occ;
    int64_t long_nc, long_ntab, long_tab_elems, long_red_sys;
    size_t long_tab_off, long_red_off;
    /* pivot hand-over between table builds: ticket and per-build flags (ints), then the
     * per-build last pivots from long_gval_off */
    size_t long_sync_off, long_gval_off;
    /* bytes of a slot's area the paths other than chunks use (the chunk fallback of FP64
     * 1025-2048 rows adds its tables after them), and the area actually provided (0: the
     * full table_stride, as initializePfSolve allocates it) */
    size_t main_bytes, area_cap;
};

/* Applies the reference's configuration defaults (pfSolve_InitializeApp.h:981-1212)
 * and validates the JW / block / dgbmv configuration. */
PfSolveResult jw_configure(const PfSolveConfiguration* in, PfSolveConfiguration* out);
/* kind of plan a configuration asks for */
int jw_plan_kind(const PfSolveConfiguration* cfg);
int jw_precision(const PfSolveConfiguration* cfg);
/* 1 when a plain JW step of N rows is cut into chunks (FP64 N > 2048, double-double
 * N > 512): longer than the one-wavefront kernels hold */
int jw_is_chunked(int64_t N, int precision);
/* geometry of the JW kernels for N rows */
PfSolveResult jw_select_geometry(int64_t N, int precision, int* L, int* R);

/* Per-launch resolution: which diagonals the prologue reads and what the main
 * kernel applies (pfSolve_JonesWorland.h:117-348). */
PfSolveResult jw_resolve_launch(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                                int64_t nsys, jw_source_desc* src, jw_run_desc* run);
PfSolveResult jw_resolve_block(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                               int64_t nsys, jw_block_desc* d);
PfSolveResult jw_resolve_gbmv(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp,
                              int64_t nsys, jw_gbmv_desc* d);
/* block stride of a diagonal in the ALT storage (get_ALT_stride) */
int64_t jw_alt_stride(int64_t N, uint64_t Msplit0);
/* iteration i of a chained step (numConsecutiveJWIterations > 1) */
PfSolveResult jw_resolve_chain_iteration(const PfSolveConfiguration* cfg, const PfSolveLaunchParams* lp, int i,
                                         jw_source_desc* src);
/* largest matrix / RHS element index a JW launch touches, and sanity of offsets */
void jw_index_extent(const jw_source_desc* s, const jw_run_desc* r, int64_t* mat_max, int64_t* rhs_max);
PfSolveResult jw_check_offsets(const jw_source_desc* s, const jw_run_desc* r);
/* keep range [lo, hi) of a zero-pad pair, clipped to [0, N) */
void jw_keep_range(int64_t zp0, int64_t zp1, int runtime, int64_t N, int64_t* lo, int64_t* hi);

#ifdef __cplusplus
}
#endif
#endif
