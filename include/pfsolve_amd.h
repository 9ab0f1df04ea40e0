/*
 * pfsolve_amd.h -- extensions of the gfx950 JW library beyond the reference
 * plan surface.  None of these has a counterpart in the reference; they sit
 * beside the drop-in entry points of pfSolve.h.
 */
#ifndef PFSOLVE_AMD_H
#define PFSOLVE_AMD_H

#include "pfSolve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* jw_type % 10 == 4: fused Jones-Worland step y = M^{-1} T^{-1} (V x), the
 * composition of the reference's dgbmv plan (V, band-major, KL = KU <= 2),
 * its tridiagonal JW plan (jw_type 13, upperBanded 2) and its bidiagonal JW
 * plan (jw_type 13, upperBanded 0), in one pass over the right-hand sides.
 *   V at offsetV:  buffer[offsetV + j*N + col], j = 0..KL+KU (natural order)
 *   T at offsetT:  [ld | md | ud], pre-normalised a/b, 1/b, c/b
 *   M at offsetM:  [md | ud],      pre-normalised 1/d, e/d
 * T and M diagonals are stored in reorder_i0 order (jw_type 14) or in natural
 * order (jw_type 1014), as the reference's JW plans read them.  scaleC
 * multiplies solution row 0 once. */
#define JWT_FUSED_BAND_TRI_BIDIAG 4
#define RUNTIME_OFFSETT 512

/* ---- ABI introspection for language bindings ---- */
size_t pfsolve_sizeof_configuration(void);
size_t pfsolve_sizeof_launch_params(void);
size_t pfsolve_sizeof_application(void);
/* byte offset of a named field, -1 if unknown */
int64_t pfsolve_offsetof_configuration(const char* field);
int64_t pfsolve_offsetof_launch_params(const char* field);
/* size of an extension struct of this header by type name, 0 if unknown */
size_t pfsolve_sizeof_ext(const char* type_name);

/* ---- plan introspection (no device needed) ---- */
typedef struct pfsolve_geometry {
    int32_t lanes;            /* lanes per system group (L) */
    int32_t rows_per_lane;    /* R */
    int32_t waves;            /* wavefronts per workgroup */
    int32_t precision;        /* 1 FP64, 2 double-double, 3 double-double with FP64 memory */
    int64_t table_bytes;      /* coefficient table per matrix set */
    int64_t lds_bytes;        /* LDS per workgroup */
    int32_t fast_rows_per_lane;   /* wave-per-system kernel (FP64, or double-double N <= 512): R (0 = not used) */
    int32_t fast_waves;           /* ... and its computing waves per workgroup */
    int32_t fast_store_waves;     /* ... storing waves per workgroup (0: each wave stores its own) */
    int32_t pad;
} pfsolve_geometry;

/* Validates a configuration as initializePfSolve does, without touching the
 * device; fills the geometry the plan would use (g may be NULL). */
PfSolveResult pfsolve_validate_configuration(const PfSolveConfiguration* cfg, pfsolve_geometry* g);
PfSolveResult pfsolve_plan_geometry(const PfSolveApplication* app, pfsolve_geometry* g);
/* storage position of logical row i (the reference's reorder_i0) */
int64_t pfsolve_reorder_i0(int64_t i, int64_t N, int64_t Nalign, int64_t Msplit0);
/* kernel variants compiled into the library */
int pfsolve_kernel_supported(int lanes, int rows_per_lane, int precision);
/* last device-side error text of the calling thread */
const char* pfsolve_last_error(void);

/* One plan's step without allocating: the work of initializePfSolve(configuration) +
 * PfSolveAppend(launchParams) + deletePfSolve, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference, whose plans always allocate.  The
 * coefficient table and the device error flag live in a static device array of the
 * library (pfsolve_noalloc_scratch_bytes(), 144 KiB + 256 B per GPU, held until the
 * process ends); configurations whose table would not fit (long systems, FP64 N > 2048
 * or double-double N > 512, whose per-chunk tables and reduced data are larger) return
 * PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER.  Nothing is allocated on the
 * device or the host (no hipMalloc, hipFree, hipEventCreate, malloc).  Runs on
 * configuration->stream[0] (or the null stream); num_streams > 1 is refused
 * (PFSOLVE_ERROR_INVALID_QUEUE).  All calls on one GPU share its scratch: calls whose
 * launches overlap on that GPU (different streams, or host threads) race; calls in
 * one stream are ordered.  The error flag is cleared by each call, so a bounded wait
 * that ran out is not reported (PfSolveSync has no plan to read it from).  Results are
 * bit-identical to PfSolveAppend of the same configuration and launch parameters. */
PfSolveResult pfsolve_append_noalloc(const PfSolveConfiguration* configuration,
                                     const PfSolveLaunchParams* launchParams);
size_t pfsolve_noalloc_scratch_bytes(void);

/* Bounded waits of the compute / store split kernel: polls before a wait gives up
 * and raises the device error flag that PfSolveSync reports (0: the default, 2^22;
 * negative: none, so every launch takes the time-out path at once and still
 * exits normally).  For tests of failure handling; no counterpart in the
 * reference. */
PfSolveResult pfsolve_set_wait_polls(PfSolveApplication* app, int32_t polls);

/* Timing hook: when set, every PfSolveAppend records `start` right before and
 * `stop` right after the main JW kernel on the plan's stream (NULL clears). */
PfSolveResult pfsolve_set_kernel_events(PfSolveApplication* app, hipEvent_t start, hipEvent_t stop);

/* ---- uneven-size batch: many matrix sets (one per harmonic degree), one step ----
 *
 * Each set holds the fused-step matrices of one size N (the jw_type 14 layout:
 * V band-major at offsetV, T [a/b | 1/b | c/b] at offsetT, M [1/d | e/d] at
 * offsetM, T and M in reorder_i0 order unless naturalOrder) and `nsys` real
 * systems solved in place.  pfsolve_batch_run turns every set's matrices into
 * its coefficient table in one launch, then solves all systems in one
 * persistent launch per row class (N <= 128, 256, 512, 1024, 2048): workgroups take
 * work units (a set and a range of its system pairs) from a queue ordered
 * largest N first.  pfsolve_batch_run_plain does the same work with one
 * plain plan launch pair per set, for comparison.  No counterpart in the
 * reference, which launches one plan per size. */
typedef struct pfsolve_batch_set {
    int64_t N;                /* rows, 2 <= N <= 2048 */
    int64_t nsys;             /* real systems sharing the set's matrices (>= 0) */
    const void* matrices;     /* device pointer, FP64 */
    int64_t offsetV, offsetT, offsetM;
    void* rhs;                /* device pointer, FP64, 16-byte aligned */
    int64_t stride;           /* elements between systems: >= N and even */
    int64_t offsetSolution;   /* element offset of system 0: even */
    int32_t naturalOrder;     /* 1: T and M diagonals in natural order (as jw_type 1014) */
    int32_t KL, KU;           /* band of V, 0..2 each */
    int32_t pad;
    double scaleC;            /* multiplies solution row 0 (1.0: none) */
} pfsolve_batch_set;

#define PFSOLVE_BATCH_CLASSES 5   /* row classes R = 2, 4, 8, 16, 32 rows per lane */

/* What the autotuner chooses, per row class c (R = 2 << c). 0 = default. */
typedef struct pfsolve_batch_tuning {
    int32_t waves[PFSOLVE_BATCH_CLASSES];       /* wavefronts per workgroup (LDS tiling) */
    int32_t unit_pairs[PFSOLVE_BATCH_CLASSES];  /* system pairs per queue unit */
    int32_t grid_per_cu[PFSOLVE_BATCH_CLASSES]; /* persistent workgroups per CU */
    int32_t rows_per_lane_shift[PFSOLVE_BATCH_CLASSES]; /* 1: sets of this class run with 2R rows
                                                           per lane (half the lanes, one scan level
                                                           fewer), when 2R <= 16 */
    /* unit_pairs counts systems, not pairs, for the R = 32 class (one system per pass) */
    int32_t kernel[PFSOLVE_BATCH_CLASSES];      /* 0: default, 1: every wavefront loads, computes and
                                                   stores its pairs, 2: compute / store split
                                                   (R <= 16; `waves` counts the computing
                                                   wavefronts), 3: direct I/O (R <= 16: the
                                                   table alone in LDS, up to 8 wavefronts),
                                                   4: as 1, with each set's table built in the
                                                   workgroup's LDS from its matrices (R <= 16,
                                                   >= 4 wavefronts; no table build launch and
                                                   no table through HBM) */
    int32_t order;            /* 0: largest N first, 1: sets as given */
    int32_t systems_per_wave; /* R = 2 class on queue kernel 1: systems one wavefront solves
                                 per pass, 2 (a pair; 0 = 2), 4 (two groups of 32 lanes, for
                                 sets with N <= 64) or 8 (four groups of 16 lanes, N <= 32;
                                 sets with 32 < N <= 64 then use 4) */
} pfsolve_batch_tuning;

typedef struct pfsolve_batch_info {
    int64_t sets, systems, units;
    int64_t table_bytes;      /* device tables of all sets */
    int64_t algorithmic_bytes;    /* per step: 8 * (N read + N written) per system + 80 N per set */
    int32_t classes_used;     /* bit c set: row class c has work */
    int32_t launches;         /* kernel launches per pfsolve_batch_run */
    pfsolve_batch_tuning tuning;  /* in effect */
} pfsolve_batch_info;

typedef struct pfsolve_batch pfsolve_batch;

/* Validates the sets, allocates the tables and the work queue (device memory
 * owned by the batch) and uploads the queue; synchronous.  tuning: NULL uses
 * the defaults, or the entry of a table loaded with pfsolve_tuning_load.
 * A batch's steps share its work queues and tables: steps of one batch must not
 * overlap on the device (issue them on one stream, or order the streams). */
PfSolveResult pfsolve_batch_create(pfsolve_batch** out, const pfsolve_batch_set* sets, int64_t nsets,
                                   const pfsolve_batch_tuning* tuning);
PfSolveResult pfsolve_batch_run(pfsolve_batch* b, hipStream_t stream);
/* run = prepare + solve.  prepare writes every set's coefficient table from its
 * matrices (except for row classes on queue kernel 4, whose solve builds them in
 * LDS); solve applies the resident tables to the right-hand sides.  A
 * caller whose matrices do not change between steps prepares once and then
 * only solves. */
PfSolveResult pfsolve_batch_prepare(pfsolve_batch* b, hipStream_t stream);
PfSolveResult pfsolve_batch_solve(pfsolve_batch* b, hipStream_t stream);
PfSolveResult pfsolve_batch_run_plain(pfsolve_batch* b, hipStream_t stream);
/* Records pfsolve_batch_run on `stream` (not the null stream) into a hipGraph,
 * then replays it with pfsolve_batch_launch_graph.  Every queue kernel leaves its
 * work queue empty when it ends (no clearing node in the graph), so every replay
 * is a full step. */
PfSolveResult pfsolve_batch_capture_graph(pfsolve_batch* b, hipStream_t stream);
/* solve_only = 1: the graph holds pfsolve_batch_solve (tables prepared beforehand) */
PfSolveResult pfsolve_batch_capture_graph_mode(pfsolve_batch* b, hipStream_t stream, int solve_only);
/* the same with `steps` (1..1024) consecutive steps in the one graph: a replay costs one
 * graph launch (measured about 7 us more than a stream launch on ROCm 7) for all of them */
PfSolveResult pfsolve_batch_capture_graph_steps(pfsolve_batch* b, hipStream_t stream, int solve_only, int steps);
PfSolveResult pfsolve_batch_launch_graph(pfsolve_batch* b, hipStream_t stream);
/* Waits for `stream` and reports a raised device error flag (a bounded wait
 * inside the compute / store split kernel ran out) as
 * PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE, clearing it. */
PfSolveResult pfsolve_batch_sync(pfsolve_batch* b, hipStream_t stream);
/* Bound of the split queue kernel's waits, as pfsolve_set_wait_polls for plans
 * (0: the default; negative: none, for tests of failure handling). */
PfSolveResult pfsolve_batch_set_wait_polls(pfsolve_batch* b, int32_t polls);
/* Autotuner: times every candidate (kernel, waves, unit size, grid, rows per
 * lane, queue order) on scratch right-hand sides of the batch's shapes (the
 * caller's systems are not touched), keeps the fastest per class, applies it
 * to the batch and returns it in *best (may be NULL).  `repeats` timed runs
 * per candidate (>= 1).  Synchronises the stream.  Candidates are timed as
 * whole steps (pfsolve_batch_run); with solve_only = 1 as solves on tables
 * prepared beforehand (pfsolve_batch_solve), for callers whose matrices do not
 * change between steps.  The process-wide table keeps the last result. */
PfSolveResult pfsolve_batch_tune(pfsolve_batch* b, hipStream_t stream, int repeats, pfsolve_batch_tuning* best);
PfSolveResult pfsolve_batch_tune_mode(pfsolve_batch* b, hipStream_t stream, int repeats, int solve_only,
                                      pfsolve_batch_tuning* best);
PfSolveResult pfsolve_batch_set_tuning(pfsolve_batch* b, const pfsolve_batch_tuning* t);
PfSolveResult pfsolve_batch_get_info(const pfsolve_batch* b, pfsolve_batch_info* info);
void pfsolve_batch_destroy(pfsolve_batch* b);

/* ---- serving loop ----
 *
 * pfsolve_serve runs a request loop over `reqs` (sorted by arrival time): each
 * request is a bundle of `nsys` real systems of size N (host input rhs_in and
 * output rhs_out, systems contiguous) whose fused-step matrices (the jw_type
 * 14 / 1014 layout, host memory, unchanged during the run) are identified by
 * matrix_key.  Requests are admitted in arrival order while device memory (one
 * arena of hbm_budget bytes for bundles in flight and resident matrices) and
 * max_inflight allow; matrices and their coefficient table stay resident per key
 * (any number of keys) until the arena needs their room: keys with no request in
 * flight are then evicted least recently used first.  Every admitted request
 * waiting for launch joins the next launch on a free stream slot (up to max_batch
 * requests: a new key's matrices and table, copies in, one persistent queue kernel
 * per row class, copies out), so new requests are launched while earlier launches
 * are still on the device; finished launches are found from their events.
 * realtime = 1 admits a request no earlier than its arrival_s (seconds from the
 * start, wall clock); realtime = 0 treats every request as arrived at the start.
 * Host buffers should be pinned for the copies to overlap.  A request that cannot
 * fit the budget even alone is rejected (status PFSOLVE_ERROR_FAILED_TO_ALLOCATE),
 * as is a key reused with another N (PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH) or with
 * another band (KL, KU), matrix offsets or storage order (PFSOLVE_ERROR_UNSUPPORTED_RADIX).  A launch that fails ends its requests with
 * the failure's status and the loop goes on with the others (fail_launch = k > 0
 * makes the k-th launch fail before its kernels, to test that path).  dry_run = n
 * > 0 runs the same policy without a device: host memory for the arena, copies
 * instead of transfers, no kernels (every result equals its input), each launch
 * complete after n polls, and every completion checks that its keys' matrices were
 * not overwritten while it was in flight (a mismatch: status
 * PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE, counted in integrity_failures).  No
 * counterpart in the reference. */
typedef struct pfsolve_serve_config {
    int64_t hbm_budget;       /* bytes of device memory the loop may hold */
    int32_t max_inflight;     /* requests admitted and not finished, at most */
    int32_t streams;          /* stream slots (launches in flight), 1..16 */
    int32_t max_batch;        /* requests per launch */
    int32_t realtime;
    int32_t fail_launch;      /* 0, or the launch (1-based) made to fail */
    int32_t dry_run;          /* 0: the GPU; n > 0: host dry run, launches done after n polls */
} pfsolve_serve_config;

typedef struct pfsolve_serve_request {
    int64_t N, nsys;
    uint64_t matrix_key;
    const double* matrices;   /* host */
    int64_t offsetV, offsetT, offsetM;
    int32_t naturalOrder, KL, KU, pad;
    double scaleC;
    const double* rhs_in;     /* host, nsys * N */
    double* rhs_out;          /* host, nsys * N */
    double arrival_s;
    double admit_s, done_s;   /* filled in: seconds from the start of the run */
    int32_t status, pad2;     /* filled in: PfSolveResult of the request */
} pfsolve_serve_request;

typedef struct pfsolve_serve_stats {
    int64_t requests, rejected, systems, launches;
    double wall_s, systems_per_s;
    double p50_ms, p99_ms, max_ms;    /* completion - arrival (realtime) or - start */
    double mean_batch;                /* requests per launch */
    int64_t peak_bytes;               /* arena bytes in use at the peak */
    int64_t resident_keys, resident_bytes;   /* at the end of the run */
    int64_t evictions;                /* keys evicted to make room */
    int64_t failed;                   /* requests ended by a failed launch */
    int64_t integrity_failures;       /* dry run: completions whose matrices were overwritten */
    int64_t peak_keys;                /* resident keys at the peak */
} pfsolve_serve_stats;

PfSolveResult pfsolve_serve(const pfsolve_serve_config* cfg, pfsolve_serve_request* reqs, int64_t nreq,
                            pfsolve_serve_stats* stats);

/* Tuning table on disk (text, one line per row class and batch size bucket):
 * pfsolve_batch_tune results are added to the process-wide table, which
 * pfsolve_batch_create consults when given no tuning. */
PfSolveResult pfsolve_tuning_save(const char* path);
PfSolveResult pfsolve_tuning_load(const char* path);
void pfsolve_tuning_clear(void);

#ifdef __cplusplus
}
#endif
#endif
