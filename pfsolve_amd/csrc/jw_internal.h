/*
 * jw_internal.h -- descriptors shared by the C host code (jw_plan.c,
 * pfsolve_api.c, jw_batch.c, jw_serve.c) and the gfx950 device code
 * (jw_device.hip, jw_main64.hip).  Plain C.
 *
 * Data model of the JW step on the device
 * ---------------------------------------
 * Every supported operator is y = B2^{-1} B1^{-1} F^{-1} (V x), row 0 scaled,
 * where V is banded (|offset| <= 2) and F (unit lower bidiagonal), B1, B2
 * (upper bidiagonal) are the factors of the solve matrices.  A table build
 * (a prologue kernel, or each workgroup of the compute / store split kernel)
 * turns the caller's matrices (in the reference storage, any mode) into one
 * coefficient table per matrix set:
 *   per row r:  vm2 vm1 v0 vp1 vp2 | l | g h | D E
 *     u_r = sum_d v_d[r] x_{r+d}
 *     w_r = u_r - l_r w_{r-1}                  (forward, F)
 *     z_r = g_r w_r - h_r z_{r+1}              (backward, B1)
 *     y_r = D_r z_r - E_r y_{r+1}              (backward, B2)
 *   per lane t and scan level k: the products of the recurrence multipliers
 *   over the lanes that level k combines.
 * The main kernel applies the table to every right-hand side.
 */
#ifndef PFSOLVE_JW_INTERNAL_H
#define PFSOLVE_JW_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { JW_BAND_NONE = 0, JW_BAND_DIAG = 1, JW_BAND_UPPER = 2, JW_BAND_LOWER = 3, JW_BAND_GB = 4 };
enum { JW_SOLVE_NONE = 0, JW_SOLVE_LOWER = 1, JW_SOLVE_UPPER = 2, JW_SOLVE_TRI = 3 };
enum { JW_PREC_F64 = 1, JW_PREC_DD = 2, JW_PREC_DD_F64MEM = 3 };

#define JW_NCOEF 10           /* coefficients per row in the table */
#define JW_NPAIR 5            /* ... stored as 5 pairs */
#define JW_MAX_BAND 2         /* |offset| of V */
/* table format 2 (64 lanes, R rows per lane): 8 coefficients per row, then
 * 3 recurrences x 8 scan multipliers per lane */
#define JW_TAB64_ELEMS(R) (8 * (R) * 64 + 24 * 64)

/* Where the prologue finds the raw coefficients of one matrix set.
 * Offsets are element offsets into the matrix buffer. */
typedef struct jw_source_desc {
    int64_t N;
    int32_t band_kind;        /* JW_BAND_* */
    int32_t solve_kind;       /* JW_SOLVE_* */
    int64_t mdA, udA, ldA;    /* JW matvec diagonals (storage permuted when `permuted`) */
    int64_t offV;             /* band-major V, natural order (JW_BAND_GB) */
    int32_t KL, KU;
    int64_t ldS, udS;         /* unit-diagonal solve off-diagonals */
    int64_t mdS;              /* fused step: T pre-scale diagonal, -1 = none */
    int32_t has_m;            /* fused step: upper bidiagonal M present */
    int32_t permuted;         /* 1: diagonals stored in reorder_i0 order */
    int64_t mdM, udM;
    int32_t Msplit0;
    int32_t mem_f64;          /* DD compute with FP64 storage */
    /* chained iterations: columns of the (rectangular) matvec, 0 = N; the row count of
     * the register layout the storage permutation uses, 0 = N; matvec diagonal an
     * implicit 1 (not read) */
    int64_t Ncols, Nalign;
    int32_t unit_diag, pad_chain;
    /* ALT variant: this descriptor is the subsystem of the rows of parity alt_parity of
     * systems of alt_N rows, stored in reorder_i_alt order (alt = 1) */
    int32_t alt, alt_parity;
    int64_t alt_N;
} jw_source_desc;

/* Tables of a chained step (numConsecutiveJWIterations = K <= JW_CHAIN_MAX), built by one
 * launch with one workgroup per iteration; passed by value (no device copy of the jobs). */
#define JW_CHAIN_MAX 16
typedef struct jw_chain_jobs {
    jw_source_desc src[JW_CHAIN_MAX];
    const void* matrices[JW_CHAIN_MAX];
    void* table;              /* iteration k's format-2 table at table + k * tab_stride bytes */
    int64_t tab_stride;
    int32_t K, R;
} jw_chain_jobs;

/* One launch of the main kernel over a batch of right-hand sides. */
typedef struct jw_run_desc {
    void* rhs;                /* systems, solved in place */
    const void* table;        /* coefficient table (prologue output) */
    int64_t nsys;             /* real systems */
    int64_t stride;           /* elements between systems */
    int64_t offSol;           /* element offset of system 0 */
    int64_t in_lo, in_hi;     /* rows read:    [in_lo, in_hi)  (already clipped to [0,N)) */
    int64_t out_lo, out_hi;   /* rows written: [out_lo, out_hi) */
    int32_t N;
    int32_t band, kl, ku;     /* band present, lower / upper band width used */
    int32_t fwd, b1, b2;      /* recurrences present */
    int32_t apply_scale;
    double scale_hi, scale_lo;
    int32_t vec2;             /* FP64: 16-byte aligned pairs may use 128-bit access */
    int32_t mem_f64;          /* DD compute with FP64 storage */
    /* fused block operators (pair = re/im of one packed complex system) */
    int32_t fused_io;         /* 1: read in_c, write out_c; 2: read in_c, write rhs (real systems) */
    int32_t in_op, out_op, out_acc;
    int32_t in_lshift, out_lshift;
    double in_scale, out_scale;
    const void* in_c;         /* packed complex input */
    void* out_c;              /* packed complex output */
    int64_t in_stride_c, in_off_c, out_stride_c, out_off_c;   /* complex elements */
    /* polls before a bounded wait of the compute / store split kernel gives up:
     * 0 the default (2^22), negative none (the time-out path at once) */
    int32_t wait_polls;
    /* elements between consecutive rows of a system (0 or 1: contiguous; 2: one parity
     * subsystem of the ALT variant) -- general kernel only */
    int32_t row_step;
} jw_run_desc;

/* Long systems (FP64 N > 2048, double-double N > 512): every system is cut into chunks of
 * C = 64 R rows (one wavefront per chunk and pair of systems).  Chunk c's format-1 table
 * (64 lanes, R rows per lane) sits at table + c * tab_elems elements; its pivots continue
 * those of the chunk before.  A chunk's result is linear in its right-hand sides and in
 * three carries: w of the row before it, z and y of the row after it.  The reduced data
 * `red` (elements of the arithmetic type) holds, per chunk c and system s of the slice,
 * array a at ((a * red_sys + s) * nchunks + c) (a system's chunks contiguous):
 *   a = 0..2  F, Z, Y: the chunk's w at its last row and z, y at its first row, all
 *             carries zero (REDUCE)
 *   a = 3..6  the right-hand side of rows rb, rb+1, rb+C-2, rb+C-1 as read (REDUCE),
 *             for the neighbours' banded matvec (the step runs in place)
 *   a = 7..9  the carries w_in, z_in, y_in (the reduced solve)
 * then, at 10 * nchunks * red_sys + 9 c + 3 j + o, the response of output o (w last,
 * z first, y first) to a unit carry j (w_in, z_in, y_in) with zero right-hand side (COEF). */
enum { JW_LONG_COEF = 0, JW_LONG_REDUCE = 1, JW_LONG_FINISH = 2 };
#define JW_LONG_ARRAYS 10
typedef struct jw_long_desc {
    jw_run_desc run;          /* rhs (slice), nsys (slice), stride, offSol, keep ranges, N, recurrences, scale */
    const void* table;
    int64_t tab_elems;        /* elements of one chunk's table */
    void* red;
    int64_t red_sys;          /* systems per reduced array row (the slice capacity) */
    int32_t nchunks, mode;
} jw_long_desc;

/* One matrix set for the batched prologue (one workgroup each). */
typedef struct jw_prologue_job {
    jw_source_desc src;
    const void* matrices;
    void* table;
    int32_t L, R, format, pad;
} jw_prologue_job;

/* Uneven-size batch: one item per matrix set, and work units (a range of
 * the set's system pairs) in queue order. */
typedef struct jw_batch_item {
    void* rhs;                /* systems of this set, solved in place */
    const void* table;        /* the set's format-2 table */
    int64_t nsys, stride, offSol;
    int32_t N, kl, ku, apply_scale;
    double scale;             /* row-0 scaleC */
} jw_batch_item;

typedef struct jw_batch_unit {
    int32_t item;             /* index into the item array */
    int32_t pair0, pair1;     /* system pairs [pair0, pair1) of the item */
    int32_t pad;
} jw_batch_unit;

typedef struct jw_block_desc {
    const void* in;
    void* out;
    int64_t N, B;
    int64_t inStride, outStride, offIn, offOut;
    int64_t in_lo, in_hi, out_lo, out_hi;
    int32_t layout, accumulate, op, lshift, skip_read;
    double scale;
} jw_block_desc;

typedef struct jw_gbmv_desc {
    const void* mat;
    void* rhs;
    int64_t N, nsys, stride, offSol, offM;
    int64_t in_lo, in_hi, out_lo, out_hi;
    int32_t LDA, KU, KL;
} jw_gbmv_desc;

/* device entry points (jw_device.hip) */
int jwdev_supported(int L, int R, int precision);
int64_t jwdev_table_elems(int L, int R);
int64_t jwdev_lds_bytes(int L, int R, int precision);
int jwdev_prologue(const jw_source_desc* d, const void* matrices, void* table, int L, int R,
                   int precision, int format, void* stream);
int jwdev_prologue_batch(const jw_prologue_job* d_jobs, int njobs, int max_RL, int precision, int* d_counters,
                         int ncounters, void* stream);
/* wave-per-system FP64 kernel (table format 2, 64 lanes, R in {2,4,8,16}) */
int64_t jwdev_table64_elems(int R);
int jwdev_main64_waves(int R);
int64_t jwdev_main64_lds_bytes(int R, int waves);
int jwdev_main64_occupancy(int R, int waves);
int jwdev_main64(const jw_run_desc* r, int R, int waves, int grid, void* stream);
/* the same with compute / store split (R <= 8): compute wavefronts never store,
 * storer wavefronts write their results; *d_err is raised if a bounded wait ran out */
int jwdev_main64s_shape(int R, int* ncomp, int* nstore);
int64_t jwdev_main64s_lds_bytes(int R, int ncomp);
int jwdev_main64s_occupancy(int R);
/* builds the table of `src` in each workgroup's LDS (no prologue launch needed) */
int jwdev_main64s(const jw_run_desc* r, const jw_source_desc* src, const void* matrices, int R, int grid, int* d_err,
                  void* stream);
/* persistent uneven-batch kernel over a unit queue; *d_counter must be 0 at launch */
int jwdev_batch64_occupancy(int R, int waves);
int jwdev_zero_counters(int* d_counters, int n, void* stream);
/* spw (R = 2 only): systems per wavefront pass, 2 (pairs), 4 (two 32-lane groups for sets
 * with N <= 64) or 8 (four 16-lane groups for N <= 32, two groups for 32 < N <= 64) */
int jwdev_batch64(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                  int waves, int grid, int spw, void* stream);
/* the same queue with direct I/O (R <= 16): LDS holds the table only, every wavefront
 * loads one pass ahead and stores from registers; up to 8 wavefronts per workgroup */
int jwdev_batch64d_occupancy(int R, int waves);
int jwdev_batch64d(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int waves, int grid, void* stream);
/* the same queue with each set's table built in the workgroup's LDS from the matrices
 * of d_jobs[item] when the workgroup's set changes (R <= 16, waves >= 4): no table build
 * launch and no table in global memory */
int jwdev_batch64b_occupancy(int R, int waves);
int jwdev_batch64b(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int waves, int grid, const jw_prologue_job* d_jobs, void* stream);
/* the same queue with the compute / store split (R <= 16): `ncomp` compute
 * wavefronts (0: default); *d_err is raised if a bounded wait ran out (after
 * `wait_polls` polls: 0 the default, negative none) */
int jwdev_batch64s_shape(int R, int ncomp, int* ncomp_out, int* nstore_out);
int64_t jwdev_batch64s_lds_bytes(int R, int ncomp);
int jwdev_batch64s_occupancy(int R, int ncomp);
int jwdev_batch64s(const jw_batch_item* d_items, const jw_batch_unit* d_units, int nunits, int* d_counter, int R,
                   int ncomp, int grid, int* d_err, int wait_polls, void* stream);
int jwdev_main(const jw_run_desc* r, int L, int R, int precision, int waves, int grid, void* stream);
/* wave-per-system double-double kernel (table format 2 in double-double, 64 lanes, R in
 * {2,4,8}: N <= 512); DD or FP64 storage (mem_f64), any alignment; `waves` per workgroup
 * (at most 8) */
int64_t jwdev_dd64_table_elems(int R);
int jwdev_dd64_occupancy(int R, int waves);
int jwdev_dd64(const jw_run_desc* r, int R, int waves, int grid, void* stream);
/* chained steps: the K tables (one launch), then the kernel that keeps each pair of systems
 * in registers through the K iterations (FP64, 16-byte aligned pairs, R in {2,4,8,16}) */
int jwdev_prologue_chain(const jw_chain_jobs* jobs, void* stream);
int jwdev_chain64_occupancy(int R);
int jwdev_chain64(const jw_run_desc* r, const void* tables, int64_t tab_elems, int K, int R, int grid, void* stream);
/* long systems: rows per lane of a chunk, chunks per table build, the table of every
 * chunk (one workgroup per build of `nsub` chunks; pivots handed from build to build),
 * the chunk kernel in one of the JW_LONG_* modes, the reduced solve */
int jwdev_long_rows(int precision);
int jwdev_long_nsub(int precision);
int64_t jwdev_long_lds_bytes(int precision);
int jwdev_long_occupancy(int precision);
/* sync: 1 + builds ints (ticket, per-build flags), gval: one value per build, err: the
 * device error flag (used by tridiagonal solves only) */
int jwdev_long_prologue(const jw_source_desc* src, const void* matrices, void* table, int64_t tab_elems,
                        int precision, int* sync, void* gval, int* err, void* stream);
int jwdev_long_chunks(const jw_long_desc* d, int precision, int grid_x, void* stream);
int jwdev_long_reduced(const jw_long_desc* d, int precision, void* stream);
/* long FP64 systems whose rows are 16-byte aligned: the chunk modes on the pair machinery
 * (format-2 tables of JW_TAB64_ELEMS(8) elements per chunk, built by jwdev_long_prologue
 * when given that table size); a persistent grid of cu_count * occupancy workgroups at most */
int jwdev_long64_nsub(void);   /* chunks per table build */
int64_t jwdev_long64_lds_bytes(void);
int jwdev_long64_occupancy(void);
int jwdev_long64(const jw_long_desc* d, int cu_count, int occupancy, void* stream);
/* the front of such a step in one launch: every chunk's table (format 2), COEF and REDUCE of
 * the slice in d; sync holds 32 + builds ints and gval one 8-byte slot per build (both set
 * by the call), err the device error flag (a bounded wait that ran out) */
int jwdev_long64_front(const jw_source_desc* src, const void* matrices, const jw_long_desc* d, int* sync, void* gval,
                       int* err, int cu_count, int occupancy, void* stream);
int jwdev_block(const jw_block_desc* d, void* stream);
int jwdev_gbmv(const jw_gbmv_desc* d, void* stream);
int jwdev_device_info(int* cu_count, int* lds_per_block, int* arch_major, int* arch_minor);
const char* jwdev_last_error(void);
void jwdev_set_error(const char* msg);
/* this GPU's scratch of the allocation-free entry point: table, error flag, table capacity */
int jwdev_noalloc_scratch(void** table, int** err, size_t* table_bytes);

#ifdef __cplusplus
}
#endif
#endif
