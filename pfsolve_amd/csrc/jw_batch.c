/*
 * jw_batch.c -- uneven-size batches of fused JW steps (pfsolve_amd.h).
 *
 * A batch is a list of matrix sets of different sizes N (one per harmonic
 * degree in a spectral code) with their right-hand sides.  The reference
 * solves such a batch with one plan per size (one JIT-compiled kernel launch
 * per N, PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:80-159 per plan).
 * Here one step is:
 *   1. jw_prologue_batch_kernel: one workgroup per set writes the set's
 *      coefficient table (format 2) and zeroes the queue counters (only the
 *      counters when every row class builds its tables in LDS);
 *   2. per row class (R rows per lane, N <= 64 R): one persistent launch of a
 *      queue kernel over that class's work units.
 * Work units are (set, range of system pairs), listed largest N first, so
 * the long units start early and the short ones fill the tail.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "jw_plan.h"

#define NCLS PFSOLVE_BATCH_CLASSES
#define NKERNELS 4                    /* queue kernels 1..NKERNELS (0: the default) */

struct pfsolve_batch {
    int64_t nsets;
    pfsolve_batch_set* sets;          /* host copy */
    int* cls;                         /* row class of each set (R = 2 << cls) */
    int cu_count;
    /* device memory */
    void* tables;                     /* one table slot per set, sized for 2R rows per lane */
    int64_t* table_off;               /* host: byte offset of each set's slot */
    int64_t table_bytes;
    jw_prologue_job* d_jobs;          /* every set's table build, indexed by set */
    jw_prologue_job* d_gjobs;         /* the builds of the sets whose tables go to global memory */
    int64_t n_gjobs;
    jw_batch_item* d_items;
    jw_batch_unit* d_units;           /* units of all launch groups, back to back */
    int* d_counters;                  /* per launch group a queue (next unit, workgroups left; the
                                         queue kernels leave it at 0), then the error flag */
    /* launch groups: sets run with R_eff = 2 << g rows per lane */
    int64_t unit_first[NCLS], unit_count[NCLS];
    int group_waves[NCLS], group_grid[NCLS], group_kernel[NCLS];
    int wait_polls;                   /* bound of the split kernel's waits (0: default) */
    int64_t total_units, total_systems;
    pfsolve_batch_tuning tuning;
    int max_RL;
    /* captured step */
    hipGraph_t graph;
    hipGraphExec_t graph_exec;
};

/* ---- tuning table (process-wide) ---- */
typedef struct tune_entry {
    int cls, bucket;
    int waves, unit_pairs, grid_per_cu, shift, order, kernel, spw;
} tune_entry;
static tune_entry* g_tune;
static int g_ntune;

static int pairs_bucket(int64_t pairs)
{
    int b = 0;
    while (b < 40 && ((int64_t)1 << (b + 1)) <= pairs) b++;
    return b;
}

/* systems per wavefront pass of the R = 2 class: 0 (default), 2, 4 or 8 */
static int spw_ok(int spw) { return spw == 0 || spw == 2 || spw == 4 || spw == 8; }

static void tune_put(int cls, int bucket, const pfsolve_batch_tuning* t)
{
    for (int i = 0; i < g_ntune; i++) {
        if (g_tune[i].cls == cls && g_tune[i].bucket == bucket) {
            g_tune[i].waves = t->waves[cls];
            g_tune[i].unit_pairs = t->unit_pairs[cls];
            g_tune[i].grid_per_cu = t->grid_per_cu[cls];
            g_tune[i].shift = t->rows_per_lane_shift[cls];
            g_tune[i].order = t->order;
            g_tune[i].kernel = t->kernel[cls];
            g_tune[i].spw = t->systems_per_wave;
            return;
        }
    }
    tune_entry* n = (tune_entry*)realloc(g_tune, sizeof(tune_entry) * (size_t)(g_ntune + 1));
    if (!n) return;
    g_tune = n;
    tune_entry e = {cls, bucket, t->waves[cls], t->unit_pairs[cls], t->grid_per_cu[cls], t->rows_per_lane_shift[cls],
                    t->order, t->kernel[cls], t->systems_per_wave};
    g_tune[g_ntune++] = e;
}

static const tune_entry* tune_get(int cls, int bucket)
{
    const tune_entry* best = NULL;
    int bd = 1 << 30;
    for (int i = 0; i < g_ntune; i++) {
        if (g_tune[i].cls != cls) continue;
        const int dd = abs(g_tune[i].bucket - bucket);
        if (dd < bd) { bd = dd; best = &g_tune[i]; }
    }
    return best;
}

void pfsolve_tuning_clear(void)
{
    free(g_tune);
    g_tune = NULL;
    g_ntune = 0;
}

PfSolveResult pfsolve_tuning_save(const char* path)
{
    if (!path) return PFSOLVE_ERROR_EMPTY_FILE;
    FILE* f = fopen(path, "w");
    if (!f) return PFSOLVE_ERROR_EMPTY_FILE;
    fprintf(f, "# pfsolve_amd batch tuning: class rows_per_lane pairs_log2 waves unit_pairs grid_per_cu shift order "
               "kernel systems_per_wave\n");
    for (int i = 0; i < g_ntune; i++)
        fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", g_tune[i].cls, 2 << g_tune[i].cls, g_tune[i].bucket,
                g_tune[i].waves, g_tune[i].unit_pairs, g_tune[i].grid_per_cu, g_tune[i].shift, g_tune[i].order,
                g_tune[i].kernel, g_tune[i].spw);
    fclose(f);
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_tuning_load(const char* path)
{
    if (!path) return PFSOLVE_ERROR_EMPTY_FILE;
    FILE* f = fopen(path, "r");
    if (!f) return PFSOLVE_ERROR_EMPTY_FILE;
    char line[256];
    int bad = 0;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        /* the kernel and systems-per-wavefront columns are optional (0: default) */
        int c, r, b, w, u, g, s, o, k = 0, spw = 0;
        const int nf = sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &c, &r, &b, &w, &u, &g, &s, &o, &k, &spw);
        if ((nf < 8 || nf > 10) || c < 0 || c >= NCLS || r != (2 << c) || w < 0 || u < 0 || g < 0 || s < 0 ||
            s > 1 || o < 0 || o > 1 || k < 0 || k > NKERNELS || !spw_ok(spw)) {
            bad = 1;
            continue;
        }
        pfsolve_batch_tuning t;
        memset(&t, 0, sizeof(t));
        t.waves[c] = w;
        t.unit_pairs[c] = u;
        t.grid_per_cu[c] = g;
        t.rows_per_lane_shift[c] = s;
        t.order = o;
        t.kernel[c] = k;
        t.systems_per_wave = spw;
        tune_put(c, b, &t);
    }
    fclose(f);
    return bad ? PFSOLVE_ERROR_EMPTY_FILE : PFSOLVE_SUCCESS;
}

/* ---- helpers ---- */
static int set_class(int64_t N)
{
    int c = 0;
    while (c < NCLS - 1 && (int64_t)(128 << c) < N) c++;
    return c;   /* N <= 64 * (2 << c) */
}

/* the doubled rows-per-lane variant exists for R = 2, 4, 8 (to 4, 8, 16) */
#define NSHIFT 3

static int eff_group(const pfsolve_batch* b, int64_t s)
{
    const int c = b->cls[s];
    return c + (c < NSHIFT && b->tuning.rows_per_lane_shift[c] ? 1 : 0);
}

/* wavefront passes of a set: pairs of systems, or systems for R = 32 */
static int64_t set_passes(const pfsolve_batch* b, int64_t s)
{
    return eff_group(b, s) == NCLS - 1 ? b->sets[s].nsys : (b->sets[s].nsys + 1) / 2;
}

static void source_of(const pfsolve_batch_set* st, jw_source_desc* src)
{
    memset(src, 0, sizeof(*src));
    src->N = st->N;
    src->band_kind = JW_BAND_GB;
    src->offV = st->offsetV;
    src->KL = st->KL;
    src->KU = st->KU;
    src->solve_kind = JW_SOLVE_TRI;
    src->ldS = st->offsetT;
    src->mdS = st->offsetT + st->N;
    src->udS = st->offsetT + 2 * st->N;
    src->has_m = 1;
    src->mdM = st->offsetM;
    src->udM = st->offsetM + st->N;
    src->Msplit0 = 704;
    /* the reference's JW plans read permuted storage except on their N <= 64 path */
    src->permuted = !st->naturalOrder && st->N > 64;
}

static void item_of(const pfsolve_batch_set* st, const void* table, void* rhs, jw_batch_item* it)
{
    memset(it, 0, sizeof(*it));
    it->rhs = rhs;
    it->table = table;
    it->nsys = st->nsys;
    it->stride = st->stride;
    it->offSol = st->offsetSolution;
    it->N = (int32_t)st->N;
    it->kl = st->KL;
    it->ku = st->KU;
    it->apply_scale = st->scaleC != 1.0;
    it->scale = st->scaleC;
}

static int group_waves_default(int g)
{
    return jwdev_main64_waves(2 << g);
}

/* kernel of launch group g: 1 = load, compute and store in every wavefront
 * (input through an LDS staging), 2 = compute / store split (R <= 16), 3 = direct
 * I/O (R <= 16: LDS holds the table only, loads one pass ahead, stores from
 * registers, up to 8 wavefronts on one table), 4 = kernel 1 with each set's table
 * built in the workgroup's LDS when its set changes (R <= 16, >= 4 wavefronts; no
 * table build launch, no table through HBM).  Default: the split where it is
 * measured faster (R <= 8: 5.35 against 4.94 TB/s at N = 512); at R = 16 the
 * staged kernels reach the same rate with the 4-5 wavefronts LDS leaves them, and
 * direct I/O measured slower (harmonic batch, solve only: 256-295 us against
 * 213-227 us: a lane's 16-byte accesses 128 bytes apart reach memory less well than
 * the staged, contiguous ones), so it is a tuner candidate only. */
static int group_kernel_default(int g) { return g == 0 ? 1 : (g <= 2 ? 2 : 1); }

/* systems per wavefront pass of the R = 2 class on queue kernel 1 when not tuned: four
 * 16-lane groups for sets with N <= 32, two 32-lane groups for N <= 64 (measured on small-N
 * batches: 1.5 - 2 x the pair kernel's rate) */
static int spw_effective(const pfsolve_batch* b) { return b->tuning.systems_per_wave ? b->tuning.systems_per_wave : 8; }

static int group_kernel_ok(int g, int k) { return k == 1 || ((k == 2 || k == 3 || k == 4) && g <= 3); }

typedef struct n_index {
    int64_t N, idx;
} n_index;

/* largest N first; equal N keep their given order */
static int cmp_n_desc(const void* a, const void* c)
{
    const n_index* x = (const n_index*)a;
    const n_index* y = (const n_index*)c;
    if (x->N != y->N) return x->N > y->N ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* passes per queue unit of set s (`pairs` passes in all): the tuned size, else `defup` when
 * set (and smaller than the set), else the whole set */
static int64_t unit_passes(const pfsolve_batch* b, int64_t s, int64_t pairs, int64_t defup)
{
    const int64_t tuned = b->tuning.unit_pairs[b->cls[s]];
    if (tuned > 0) return tuned;
    if (defup > 0 && defup < pairs) return defup;
    return pairs > 0 ? pairs : 1;
}

/* fills the units (host, grouped by effective row class, queue order inside a
 * group) and the launch shapes from b->tuning */
static PfSolveResult build_queue(pfsolve_batch* b, jw_batch_unit** out_units)
{
    const int64_t S = b->nsets;
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(S > 0 ? S : 1));
    n_index* keys = (n_index*)malloc(sizeof(n_index) * (size_t)(S > 0 ? S : 1));
    if (!order || !keys) { free(order); free(keys); return PFSOLVE_ERROR_MALLOC_FAILED; }
    for (int64_t i = 0; i < S; i++) { keys[i].N = b->sets[i].N; keys[i].idx = i; }
    if (b->tuning.order == 0 && S > 1) qsort(keys, (size_t)S, sizeof(n_index), cmp_n_desc);
    for (int64_t i = 0; i < S; i++) order[i] = keys[i].idx;
    free(keys);
    /* untuned unit size: whole sets, unless a group has fewer sets than 4 units per CU; then
     * units of at least 16 passes that give each CU about 4 of them */
    int64_t nsets_g[NCLS] = {0}, passes_g[NCLS] = {0}, defup[NCLS] = {0};
    for (int64_t s = 0; s < S; s++) {
        nsets_g[eff_group(b, s)]++;
        passes_g[eff_group(b, s)] += set_passes(b, s);
    }
    for (int g = 0; g < NCLS; g++) {
        const int64_t target = 4 * (int64_t)(b->cu_count > 0 ? b->cu_count : 1);
        if (nsets_g[g] < target) {
            defup[g] = (passes_g[g] + target - 1) / target;
            if (defup[g] < 16) defup[g] = 16;
        }
    }
    /* count units per group */
    int64_t cnt[NCLS] = {0};
    for (int64_t k = 0; k < S; k++) {
        const int64_t s = order[k];
        const int g = eff_group(b, s);
        const int64_t pairs = set_passes(b, s);
        const int64_t up = unit_passes(b, s, pairs, defup[g]);
        cnt[g] += (pairs + up - 1) / up;
    }
    int64_t total = 0;
    for (int g = 0; g < NCLS; g++) {
        b->unit_first[g] = total;
        b->unit_count[g] = cnt[g];
        total += cnt[g];
    }
    if (total > 0x7fffffff) { free(order); return PFSOLVE_ERROR_EMPTY_size; }
    jw_batch_unit* u = (jw_batch_unit*)calloc((size_t)(total > 0 ? total : 1), sizeof(jw_batch_unit));
    if (!u) { free(order); return PFSOLVE_ERROR_MALLOC_FAILED; }
    int64_t fill[NCLS];
    for (int g = 0; g < NCLS; g++) fill[g] = b->unit_first[g];
    for (int64_t k = 0; k < S; k++) {
        const int64_t s = order[k];
        const int g = eff_group(b, s);
        const int64_t pairs = set_passes(b, s);
        const int64_t up = unit_passes(b, s, pairs, defup[g]);
        for (int64_t p = 0; p < pairs; p += up) {
            jw_batch_unit* x = &u[fill[g]++];
            x->item = (int32_t)s;
            x->pair0 = (int32_t)p;
            x->pair1 = (int32_t)(p + up < pairs ? p + up : pairs);
        }
    }
    free(order);
    b->total_units = total;
    *out_units = u;
    /* launch shapes */
    for (int g = 0; g < NCLS; g++) {
        b->group_waves[g] = 0;
        b->group_grid[g] = 0;
        if (!b->unit_count[g]) continue;
        /* the tuning of the native class that feeds this group wins; default otherwise */
        int w = 0, gpc = 0, k = 0;
        for (int c = 0; c < NCLS; c++) {
            const int ge = c + (c < NSHIFT && b->tuning.rows_per_lane_shift[c] ? 1 : 0);
            if (ge != g) continue;
            if (b->tuning.waves[c] > w) w = b->tuning.waves[c];
            if (b->tuning.grid_per_cu[c] > gpc) gpc = b->tuning.grid_per_cu[c];
            if (b->tuning.kernel[c] > k) k = b->tuning.kernel[c];
        }
        if (!group_kernel_ok(g, k)) k = group_kernel_default(g);
        int occ = 0;
        if (k == 4) {
            /* the table build needs at least 4 wavefronts */
            const int wmax = group_waves_default(g);
            if (w <= 0 || w > wmax) w = wmax;
            if (w < 4) w = 4;
            occ = jwdev_batch64b_occupancy(2 << g, w);
        } else if (k == 3) {
            if (w <= 0 || w > 8) w = 8;
            occ = jwdev_batch64d_occupancy(2 << g, w);
        } else if (k == 2) {
            int nc = 0, ns = 0;
            if (!jwdev_batch64s_shape(2 << g, w, &nc, &ns)) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
            w = nc;
            occ = jwdev_batch64s_occupancy(2 << g, w);
        } else {
            const int wmax = group_waves_default(g);
            if (w <= 0 || w > wmax) w = wmax;
            if (w < 1) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
            occ = jwdev_batch64_occupancy(2 << g, w);
        }
        if (occ < 1) return PFSOLVE_ERROR_FAILED_TO_SET_DYNAMIC_SHARED_MEMORY;
        if (gpc <= 0 || gpc > occ) gpc = occ;
        int64_t grid = (int64_t)b->cu_count * gpc;
        if (grid > b->unit_count[g]) grid = b->unit_count[g];
        b->group_waves[g] = w;
        b->group_grid[g] = (int)(grid > 0 ? grid : 1);
        b->group_kernel[g] = k;
    }
    return PFSOLVE_SUCCESS;
}

static PfSolveResult upload_queue(pfsolve_batch* b)
{
    jw_batch_unit* u = NULL;
    PfSolveResult r = build_queue(b, &u);
    if (r != PFSOLVE_SUCCESS) return r;
    if (b->d_units) { hipFree(b->d_units); b->d_units = NULL; }
    const size_t bytes = sizeof(jw_batch_unit) * (size_t)(b->total_units > 0 ? b->total_units : 1);
    if (hipMalloc((void**)&b->d_units, bytes) != hipSuccess) { free(u); return PFSOLVE_ERROR_FAILED_TO_ALLOCATE; }
    hipError_t e = hipMemcpy(b->d_units, u, bytes, hipMemcpyHostToDevice);
    free(u);
    if (e != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_COPY;
    /* the prologue jobs carry the effective rows per lane of each set; the sets of groups
     * that build their tables in LDS are left out of the table build launch */
    const size_t njb = (size_t)(b->nsets > 0 ? b->nsets : 1);
    jw_prologue_job* jobs = (jw_prologue_job*)calloc(njb, sizeof(jw_prologue_job));
    jw_prologue_job* gjobs = (jw_prologue_job*)calloc(njb, sizeof(jw_prologue_job));
    if (!jobs || !gjobs) { free(jobs); free(gjobs); return PFSOLVE_ERROR_MALLOC_FAILED; }
    b->max_RL = 0;
    b->n_gjobs = 0;
    for (int64_t s = 0; s < b->nsets; s++) {
        source_of(&b->sets[s], &jobs[s].src);
        jobs[s].matrices = b->sets[s].matrices;
        jobs[s].table = (char*)b->tables + b->table_off[s];
        jobs[s].L = 64;
        jobs[s].R = 2 << eff_group(b, s);
        jobs[s].format = 2;
        if (b->group_kernel[eff_group(b, s)] == 4) continue;
        gjobs[b->n_gjobs++] = jobs[s];
        if (64 * jobs[s].R > b->max_RL) b->max_RL = 64 * jobs[s].R;
    }
    e = hipMemcpy(b->d_jobs, jobs, sizeof(jw_prologue_job) * (size_t)b->nsets, hipMemcpyHostToDevice);
    if (e == hipSuccess && b->n_gjobs > 0)
        e = hipMemcpy(b->d_gjobs, gjobs, sizeof(jw_prologue_job) * (size_t)b->n_gjobs, hipMemcpyHostToDevice);
    free(jobs);
    free(gjobs);
    if (e != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_COPY;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = NULL; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = NULL; }
    return PFSOLVE_SUCCESS;
}

static PfSolveResult upload_items(pfsolve_batch* b, jw_batch_item* d_items, void* const* rhs_override)
{
    jw_batch_item* it = (jw_batch_item*)calloc((size_t)(b->nsets > 0 ? b->nsets : 1), sizeof(jw_batch_item));
    if (!it) return PFSOLVE_ERROR_MALLOC_FAILED;
    for (int64_t s = 0; s < b->nsets; s++)
        item_of(&b->sets[s], (char*)b->tables + b->table_off[s], rhs_override ? rhs_override[s] : b->sets[s].rhs,
                &it[s]);
    hipError_t e = hipMemcpy(d_items, it, sizeof(jw_batch_item) * (size_t)b->nsets, hipMemcpyHostToDevice);
    free(it);
    return e == hipSuccess ? PFSOLVE_SUCCESS : PFSOLVE_ERROR_FAILED_TO_COPY;
}

void pfsolve_batch_destroy(pfsolve_batch* b)
{
    if (!b) return;
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->tables) hipFree(b->tables);
    if (b->d_jobs) hipFree(b->d_jobs);
    if (b->d_gjobs) hipFree(b->d_gjobs);
    if (b->d_items) hipFree(b->d_items);
    if (b->d_units) hipFree(b->d_units);
    if (b->d_counters) hipFree(b->d_counters);
    free(b->sets);
    free(b->cls);
    free(b->table_off);
    free(b);
}

static PfSolveResult check_set(const pfsolve_batch_set* s)
{
    if (s->N < 2 || s->N > 2048) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
    if (s->nsys < 0) return PFSOLVE_ERROR_EMPTY_size;
    if (s->KL < 0 || s->KL > 2 || s->KU < 0 || s->KU > 2) return PFSOLVE_ERROR_UNSUPPORTED_RADIX;
    if (!s->matrices) return PFSOLVE_ERROR_EMPTY_buffer;
    if (s->offsetV < 0 || s->offsetT < 0 || s->offsetM < 0) return PFSOLVE_ERROR_EMPTY_bufferSize;
    if (s->nsys == 0) return PFSOLVE_SUCCESS;
    if (!s->rhs) return PFSOLVE_ERROR_EMPTY_outputBuffer;
    if (((uintptr_t)s->rhs % 16) != 0 || s->stride < s->N || (s->stride % 2) != 0 || s->offsetSolution < 0 ||
        (s->offsetSolution % 2) != 0)
        return PFSOLVE_ERROR_EMPTY_outputBufferSize;
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_create(pfsolve_batch** out, const pfsolve_batch_set* sets, int64_t nsets,
                                   const pfsolve_batch_tuning* tuning)
{
    if (!out) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    *out = NULL;
    if (nsets < 1 || !sets) return PFSOLVE_ERROR_EMPTY_size;
    if (nsets > 0x7fffffff) return PFSOLVE_ERROR_EMPTY_size;
    for (int64_t s = 0; s < nsets; s++) {
        PfSolveResult r = check_set(&sets[s]);
        if (r != PFSOLVE_SUCCESS) return r;
    }
    pfsolve_batch* b = (pfsolve_batch*)calloc(1, sizeof(*b));
    if (!b) return PFSOLVE_ERROR_MALLOC_FAILED;
    b->nsets = nsets;
    b->sets = (pfsolve_batch_set*)malloc(sizeof(pfsolve_batch_set) * (size_t)nsets);
    b->cls = (int*)malloc(sizeof(int) * (size_t)nsets);
    b->table_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)nsets);
    if (!b->sets || !b->cls || !b->table_off) { pfsolve_batch_destroy(b); return PFSOLVE_ERROR_MALLOC_FAILED; }
    memcpy(b->sets, sets, sizeof(pfsolve_batch_set) * (size_t)nsets);
    int lds = 0, maj = 0, min = 0;
    if (jwdev_device_info(&b->cu_count, &lds, &maj, &min) != 0) { pfsolve_batch_destroy(b); return PFSOLVE_ERROR_INVALID_DEVICE; }
    int64_t off = 0, pairs_cls[NCLS] = {0};
    for (int64_t s = 0; s < nsets; s++) {
        b->cls[s] = set_class(sets[s].N);
        b->table_off[s] = off;
        int rmax = 2 << b->cls[s];
        if (b->cls[s] < NSHIFT) rmax *= 2;   /* room for the doubled rows-per-lane variant */
        off += ((int64_t)jwdev_table64_elems(rmax) * 8 + 255) / 256 * 256;
        b->total_systems += sets[s].nsys;
        pairs_cls[b->cls[s]] += b->cls[s] == NCLS - 1 ? sets[s].nsys : (sets[s].nsys + 1) / 2;
    }
    b->table_bytes = off;
    if (hipMalloc(&b->tables, (size_t)off) != hipSuccess ||
        hipMalloc((void**)&b->d_jobs, sizeof(jw_prologue_job) * (size_t)nsets) != hipSuccess ||
        hipMalloc((void**)&b->d_gjobs, sizeof(jw_prologue_job) * (size_t)nsets) != hipSuccess ||
        hipMalloc((void**)&b->d_items, sizeof(jw_batch_item) * (size_t)nsets) != hipSuccess ||
        hipMalloc((void**)&b->d_counters, sizeof(int) * (2 * NCLS + 1)) != hipSuccess) {
        pfsolve_batch_destroy(b);
        return PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
    }
    if (hipMemset(b->d_counters, 0, sizeof(int) * (2 * NCLS + 1)) != hipSuccess) {
        pfsolve_batch_destroy(b);
        return PFSOLVE_ERROR_FAILED_TO_COPY;
    }
    if (tuning) {
        b->tuning = *tuning;
    } else {
        memset(&b->tuning, 0, sizeof(b->tuning));
        for (int c = 0; c < NCLS; c++) {
            if (!pairs_cls[c]) continue;
            const tune_entry* e = tune_get(c, pairs_bucket(pairs_cls[c]));
            if (!e) continue;
            b->tuning.waves[c] = e->waves;
            b->tuning.unit_pairs[c] = e->unit_pairs;
            b->tuning.grid_per_cu[c] = e->grid_per_cu;
            b->tuning.rows_per_lane_shift[c] = e->shift;
            b->tuning.kernel[c] = e->kernel;
            b->tuning.order = e->order;
            if (c == 0) b->tuning.systems_per_wave = e->spw;
        }
    }
    for (int c = NSHIFT; c < NCLS; c++) b->tuning.rows_per_lane_shift[c] = 0;
    PfSolveResult r = upload_queue(b);
    if (r == PFSOLVE_SUCCESS) r = upload_items(b, b->d_items, NULL);
    if (r != PFSOLVE_SUCCESS) { pfsolve_batch_destroy(b); return r; }
    *out = b;
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_set_tuning(pfsolve_batch* b, const pfsolve_batch_tuning* t)
{
    if (!b || !t) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    for (int c = 0; c < NCLS; c++)
        if (t->waves[c] < 0 || t->unit_pairs[c] < 0 || t->grid_per_cu[c] < 0 || t->rows_per_lane_shift[c] < 0 ||
            t->rows_per_lane_shift[c] > 1 || t->kernel[c] < 0 || t->kernel[c] > NKERNELS)
            return PFSOLVE_ERROR_EMPTY_size;
    if (t->order < 0 || t->order > 1 || !spw_ok(t->systems_per_wave)) return PFSOLVE_ERROR_EMPTY_size;
    b->tuning = *t;
    for (int c = NSHIFT; c < NCLS; c++) b->tuning.rows_per_lane_shift[c] = 0;
    return upload_queue(b);
}

/* the tables that live in global memory (that launch also clears the queues).  With none,
 * nothing is launched: every queue kernel leaves its queue empty when it ends, so a step
 * whose row classes all build their tables in LDS is one launch per class. */
static int run_prologue(pfsolve_batch* b, hipStream_t stream)
{
    if (b->n_gjobs > 0)
        return jwdev_prologue_batch(b->d_gjobs, (int)b->n_gjobs, b->max_RL, JW_PREC_F64, b->d_counters, 2 * NCLS,
                                    stream);
    return 0;
}

static PfSolveResult launch_step(pfsolve_batch* b, const jw_batch_item* d_items, hipStream_t stream, int only_group,
                                 int with_prologue)
{
    if (with_prologue && run_prologue(b, stream) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    for (int g = 0; g < NCLS; g++) {
        if (!b->unit_count[g] || (only_group >= 0 && g != only_group)) continue;
        const int k = b->group_kernel[g];
        const int rc = k == 4 ? jwdev_batch64b(d_items, b->d_units + b->unit_first[g], (int)b->unit_count[g],
                                               b->d_counters + 2 * g, 2 << g, b->group_waves[g], b->group_grid[g],
                                               b->d_jobs, stream)
                       : k == 2 ? jwdev_batch64s(d_items, b->d_units + b->unit_first[g], (int)b->unit_count[g],
                                               b->d_counters + 2 * g, 2 << g, b->group_waves[g], b->group_grid[g],
                                               b->d_counters + 2 * NCLS, b->wait_polls, stream)
                       : k == 3 ? jwdev_batch64d(d_items, b->d_units + b->unit_first[g], (int)b->unit_count[g],
                                                 b->d_counters + 2 * g, 2 << g, b->group_waves[g], b->group_grid[g], stream)
                                : jwdev_batch64(d_items, b->d_units + b->unit_first[g], (int)b->unit_count[g],
                                                b->d_counters + 2 * g, 2 << g, b->group_waves[g], b->group_grid[g],
                                                spw_effective(b), stream);
        if (rc != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_run(pfsolve_batch* b, hipStream_t stream)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    return launch_step(b, b->d_items, stream, -1, 1);
}

PfSolveResult pfsolve_batch_prepare(pfsolve_batch* b, hipStream_t stream)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    return run_prologue(b, stream) == 0 ? PFSOLVE_SUCCESS : PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
}

PfSolveResult pfsolve_batch_solve(pfsolve_batch* b, hipStream_t stream)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    return launch_step(b, b->d_items, stream, -1, 0);
}

PfSolveResult pfsolve_batch_run_plain(pfsolve_batch* b, hipStream_t stream)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    for (int64_t s = 0; s < b->nsets; s++) {
        const pfsolve_batch_set* st = &b->sets[s];
        const int R = 2 << b->cls[s];
        void* table = (char*)b->tables + b->table_off[s];
        jw_source_desc src;
        source_of(st, &src);
        if (jwdev_prologue(&src, st->matrices, table, 64, R, JW_PREC_F64, 2, stream) != 0)
            return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
        if (st->nsys <= 0) continue;
        jw_run_desc run;
        memset(&run, 0, sizeof(run));
        run.rhs = st->rhs;
        run.table = table;
        run.nsys = st->nsys;
        run.stride = st->stride;
        run.offSol = st->offsetSolution;
        run.in_lo = 0;
        run.in_hi = st->N;
        run.out_lo = 0;
        run.out_hi = st->N;
        run.N = (int32_t)st->N;
        run.band = 1;
        run.kl = st->KL;
        run.ku = st->KU;
        run.fwd = run.b1 = run.b2 = 1;
        run.apply_scale = st->scaleC != 1.0;
        run.scale_hi = st->scaleC;
        run.vec2 = 1;
        const int waves = jwdev_main64_waves(R);
        const int occ = jwdev_main64_occupancy(R, waves);
        if (waves < 1 || occ < 1) return PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH;
        const int64_t npass = R == 32 ? st->nsys : (st->nsys + 1) / 2;
        int64_t grid = (npass + waves - 1) / waves;
        if (grid > (int64_t)b->cu_count * occ) grid = (int64_t)b->cu_count * occ;
        if (jwdev_main64(&run, R, waves, (int)grid, stream) != 0) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_capture_graph(pfsolve_batch* b, hipStream_t stream)
{
    return pfsolve_batch_capture_graph_mode(b, stream, 0);
}

PfSolveResult pfsolve_batch_capture_graph_mode(pfsolve_batch* b, hipStream_t stream, int solve_only)
{
    return pfsolve_batch_capture_graph_steps(b, stream, solve_only, 1);
}

PfSolveResult pfsolve_batch_capture_graph_steps(pfsolve_batch* b, hipStream_t stream, int solve_only, int steps)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    if (!stream) return PFSOLVE_ERROR_INVALID_QUEUE;
    if (steps < 1 || steps > 1024) return PFSOLVE_ERROR_EMPTY_size;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = NULL; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = NULL; }
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed) != hipSuccess)
        return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    PfSolveResult r = PFSOLVE_SUCCESS;
    for (int i = 0; i < steps && r == PFSOLVE_SUCCESS; i++)
        r = solve_only ? pfsolve_batch_solve(b, stream) : launch_step(b, b->d_items, stream, -1, 1);
    hipGraph_t g = NULL;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (r != PFSOLVE_SUCCESS) { if (g) hipGraphDestroy(g); return r; }
    if (e != hipSuccess || !g) return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    b->graph = g;
    if (hipGraphInstantiate(&b->graph_exec, g, NULL, NULL, 0) != hipSuccess) {
        hipGraphDestroy(g);
        b->graph = NULL;
        b->graph_exec = NULL;
        return PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
    }
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_launch_graph(pfsolve_batch* b, hipStream_t stream)
{
    if (!b || !b->graph_exec) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    return hipGraphLaunch(b->graph_exec, stream) == hipSuccess ? PFSOLVE_SUCCESS : PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
}

PfSolveResult pfsolve_batch_set_wait_polls(pfsolve_batch* b, int32_t polls)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    b->wait_polls = polls;
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_sync(pfsolve_batch* b, hipStream_t stream)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    if (hipStreamSynchronize(stream) != hipSuccess) return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
    int e = 0;
    if (hipMemcpy(&e, b->d_counters + 2 * NCLS, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
    if (e) {
        (void)hipMemset(b->d_counters + 2 * NCLS, 0, sizeof(int));
        return PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE;
    }
    return PFSOLVE_SUCCESS;
}

PfSolveResult pfsolve_batch_get_info(const pfsolve_batch* b, pfsolve_batch_info* info)
{
    if (!b || !info) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    memset(info, 0, sizeof(*info));
    info->sets = b->nsets;
    info->systems = b->total_systems;
    info->units = b->total_units;
    info->table_bytes = b->table_bytes;
    int64_t alg = 0;
    for (int64_t s = 0; s < b->nsets; s++) alg += 16 * b->sets[s].N * b->sets[s].nsys + 80 * b->sets[s].N;
    info->algorithmic_bytes = alg;
    /* per step: the table build launch (when some class reads tables from global memory)
     * and one queue launch per class */
    int launches = b->n_gjobs > 0 ? 1 : 0;
    for (int g = 0; g < NCLS; g++)
        if (b->unit_count[g]) { info->classes_used |= 1 << g; launches++; }
    info->launches = launches;
    info->tuning = b->tuning;
    return PFSOLVE_SUCCESS;
}

/* ---- autotuner ---- */
static int time_group(pfsolve_batch* b, const jw_batch_item* d_items, hipStream_t stream, int g, int repeats,
                      int solve_only, float* ms_out)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return -1; }
    float best = 1e30f;
    int rc = 0;
    /* warm-up (also writes the tables) */
    if (launch_step(b, d_items, stream, g, 1) != PFSOLVE_SUCCESS) rc = -1;
    for (int k = 0; rc == 0 && k < repeats; k++) {
        /* the whole step (the global tables' build, then the queue launches), or with
         * solve_only the queue launches alone on tables built beforehand */
        if (solve_only && run_prologue(b, stream) != 0) { rc = -1; break; }
        hipEventRecord(e0, stream);
        if (launch_step(b, d_items, stream, g, !solve_only) != PFSOLVE_SUCCESS) { rc = -1; break; }
        hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = -1; break; }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    /* a candidate whose bounded waits ran out is an error, not a timing */
    if (rc == 0 && pfsolve_batch_sync(b, stream) != PFSOLVE_SUCCESS) rc = -1;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms_out = best;
    return rc;
}

PfSolveResult pfsolve_batch_tune(pfsolve_batch* b, hipStream_t stream, int repeats, pfsolve_batch_tuning* best_out)
{
    return pfsolve_batch_tune_mode(b, stream, repeats, 0, best_out);
}

PfSolveResult pfsolve_batch_tune_mode(pfsolve_batch* b, hipStream_t stream, int repeats, int solve_only,
                                      pfsolve_batch_tuning* best_out)
{
    if (!b) return PFSOLVE_ERROR_PLAN_NOT_INITIALIZED;
    if (repeats < 1) repeats = 1;
    /* scratch right-hand sides with the batch's shapes */
    int64_t total = 0;
    int64_t* roff = (int64_t*)malloc(sizeof(int64_t) * (size_t)b->nsets);
    void** rptr = (void**)malloc(sizeof(void*) * (size_t)b->nsets);
    if (!roff || !rptr) { free(roff); free(rptr); return PFSOLVE_ERROR_MALLOC_FAILED; }
    for (int64_t s = 0; s < b->nsets; s++) {
        const pfsolve_batch_set* st = &b->sets[s];
        roff[s] = total;
        const int64_t n = st->nsys > 0 ? (st->nsys - 1) * st->stride + st->offsetSolution + st->N : 0;
        total += (n + 1) / 2 * 2;
    }
    double* scratch = NULL;
    jw_batch_item* d_items = NULL;
    PfSolveResult res = PFSOLVE_SUCCESS;
    if (hipMalloc((void**)&scratch, (size_t)(total > 0 ? total : 2) * 8) != hipSuccess ||
        hipMalloc((void**)&d_items, sizeof(jw_batch_item) * (size_t)b->nsets) != hipSuccess) {
        res = PFSOLVE_ERROR_FAILED_TO_ALLOCATE;
        goto done;
    }
    if (hipMemsetAsync(scratch, 0, (size_t)(total > 0 ? total : 2) * 8, stream) != hipSuccess) {
        res = PFSOLVE_ERROR_FAILED_TO_COPY;
        goto done;
    }
    for (int64_t s = 0; s < b->nsets; s++) rptr[s] = scratch + roff[s];
    res = upload_items(b, d_items, rptr);
    if (res != PFSOLVE_SUCCESS) goto done;
    {
        const pfsolve_batch_tuning start = b->tuning;
        pfsolve_batch_tuning cur = start;
        int64_t pairs_cls[NCLS] = {0};
        for (int64_t s = 0; s < b->nsets; s++)
            pairs_cls[b->cls[s]] += b->cls[s] == NCLS - 1 ? b->sets[s].nsys : (b->sets[s].nsys + 1) / 2;
        static const int units_try[] = {0, 1, 2, 4, 8, 16, 32, 64};
        for (int c = 0; c < NCLS && res == PFSOLVE_SUCCESS; c++) {
            if (!pairs_cls[c]) continue;
            float best_ms = 1e30f;
            pfsolve_batch_tuning best = cur;
            const int shifts = c < NSHIFT ? 2 : 1;
            for (int sh = 0; sh < shifts && res == PFSOLVE_SUCCESS; sh++) {
                const int g = c + sh;
                for (int k = 1; k <= NKERNELS && res == PFSOLVE_SUCCESS; k++) {
                    if (!group_kernel_ok(g, k)) continue;
                    /* a few wave counts: for the one-wavefront kernel the largest, its half and
                     * quarter; for the split kernel (computing wavefronts) the default, its half
                     * and the most LDS holds */
                    int wl[3] = {0, 0, 0};
                    if (k == 1) {
                        const int wmax = group_waves_default(g);
                        wl[0] = wmax;
                        wl[1] = (wmax + 1) / 2;
                        wl[2] = (wmax + 3) / 4;
                    } else if (k == 3) {
                        wl[0] = 8;
                        wl[1] = 4;
                        wl[2] = 2;
                    } else if (k == 4) {
                        /* the table build needs >= 4 wavefronts */
                        const int wmax = group_waves_default(g);
                        wl[0] = wmax;
                        wl[1] = (wmax + 1) / 2 >= 4 ? (wmax + 1) / 2 : 0;
                        wl[2] = 4;
                    } else {
                        int nc = 0, ns = 0;
                        if (jwdev_batch64s_shape(2 << g, 0, &nc, &ns)) wl[0] = nc;
                        if (wl[0] > 1 && jwdev_batch64s_shape(2 << g, wl[0] / 2, &nc, &ns)) wl[1] = nc;
                        if (jwdev_batch64s_shape(2 << g, 64, &nc, &ns) && nc != wl[0]) wl[2] = nc;
                    }
                    for (int wi = 0; wi < 3 && res == PFSOLVE_SUCCESS; wi++) {
                        const int w = wl[wi];
                        if (w < 1 || (wi > 0 && w == wl[0]) || (wi == 2 && w == wl[1])) continue;
                        const int occ = k == 2   ? jwdev_batch64s_occupancy(2 << g, w)
                                        : k == 3 ? jwdev_batch64d_occupancy(2 << g, w)
                                        : k == 4 ? jwdev_batch64b_occupancy(2 << g, w)
                                                 : jwdev_batch64_occupancy(2 << g, w);
                        for (int gpc = 1; gpc <= occ && gpc <= 2; gpc++) {
                            for (size_t ui = 0; ui < sizeof(units_try) / sizeof(units_try[0]); ui++) {
                                pfsolve_batch_tuning t = cur;
                                t.rows_per_lane_shift[c] = sh;
                                t.kernel[c] = k;
                                t.waves[c] = w;
                                t.grid_per_cu[c] = gpc;
                                t.unit_pairs[c] = units_try[ui];
                                /* the whole solve is timed: a class moved to the next group shares
                                 * that launch with the next class's sets */
                                b->tuning = t;
                                res = upload_queue(b);
                                if (res != PFSOLVE_SUCCESS) break;
                                float ms = 0.f;
                                if (time_group(b, d_items, stream, -1, repeats, solve_only, &ms) != 0) {
                                    res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                                    break;
                                }
                                if (ms < best_ms) { best_ms = ms; best = t; }
                            }
                            if (res != PFSOLVE_SUCCESS) break;
                        }
                    }
                }
            }
            cur = best;
        }
        if (res == PFSOLVE_SUCCESS) {
            /* systems per wavefront pass of the R = 2 class on queue kernel 1 (sub-wavefront
             * groups for its sets with N <= 64) */
            b->tuning = cur;
            res = upload_queue(b);
            int small = 0;
            for (int64_t s = 0; s < b->nsets && res == PFSOLVE_SUCCESS; s++)
                if (b->sets[s].N <= 64 && b->sets[s].nsys > 0 && eff_group(b, s) == 0) small = 1;
            if (res == PFSOLVE_SUCCESS && small && b->group_kernel[0] == 1) {
                static const int spw_try[] = {2, 4, 8};
                float best_ms = 1e30f;
                int best_spw = cur.systems_per_wave;
                for (int i = 0; i < 3 && res == PFSOLVE_SUCCESS; i++) {
                    pfsolve_batch_tuning t = cur;
                    t.systems_per_wave = spw_try[i];
                    b->tuning = t;
                    float ms = 0.f;
                    if (time_group(b, d_items, stream, -1, repeats, solve_only, &ms) != 0) {
                        res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
                        break;
                    }
                    if (ms < best_ms) { best_ms = ms; best_spw = spw_try[i]; }
                }
                cur.systems_per_wave = best_spw;
            }
        }
        if (res == PFSOLVE_SUCCESS) {
            /* queue order: largest N first vs as given, on the whole step */
            float ms_order[2] = {1e30f, 1e30f};
            for (int o = 0; o < 2 && res == PFSOLVE_SUCCESS; o++) {
                pfsolve_batch_tuning t = cur;
                t.order = o;
                b->tuning = t;
                res = upload_queue(b);
                if (res != PFSOLVE_SUCCESS) break;
                if (time_group(b, d_items, stream, -1, repeats, solve_only, &ms_order[o]) != 0)
                    res = PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL;
            }
            cur.order = ms_order[1] < ms_order[0] ? 1 : 0;
        }
        b->tuning = res == PFSOLVE_SUCCESS ? cur : start;
        PfSolveResult r2 = upload_queue(b);
        if (res == PFSOLVE_SUCCESS) res = r2;
        if (res == PFSOLVE_SUCCESS) {
            for (int c = 0; c < NCLS; c++)
                if (pairs_cls[c]) tune_put(c, pairs_bucket(pairs_cls[c]), &cur);
            if (best_out) *best_out = cur;
        }
    }
done:
    if (stream) hipStreamSynchronize(stream);
    else hipDeviceSynchronize();
    if (d_items) hipFree(d_items);
    if (scratch) hipFree(scratch);
    free(roff);
    free(rptr);
    return res;
}
