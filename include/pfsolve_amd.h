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

/* ---- plan introspection (no device needed) ---- */
typedef struct pfsolve_geometry {
    int32_t lanes;            /* lanes per system group (L) */
    int32_t rows_per_lane;    /* R */
    int32_t waves;            /* wavefronts per workgroup */
    int32_t precision;        /* 1 FP64, 2 double-double, 3 double-double with FP64 memory */
    int64_t table_bytes;      /* coefficient table per matrix set */
    int64_t lds_bytes;        /* LDS per workgroup */
    int32_t fast_rows_per_lane;   /* wave-per-system FP64 kernel: R (0 = not used) */
    int32_t fast_waves;           /* ... and its waves per workgroup */
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

/* Timing hook: when set, every PfSolveAppend records `start` right before and
 * `stop` right after the main JW kernel on the plan's stream (NULL clears). */
PfSolveResult pfsolve_set_kernel_events(PfSolveApplication* app, hipEvent_t start, hipEvent_t stop);

#ifdef __cplusplus
}
#endif
#endif
