/*
 * fishfry.h -- C-ABI of the MI355X (gfx950) periodic Poisson solver.
 *
 * The drop-in boundary for fishfry's solver plugin.  The reference binds its
 * solver as the template parameter P of FishFry<P> (FishFry.hpp:103-104,
 * main.cpp:48 instantiates FishFry<ParisPencil>); P must provide a
 * constructor, bytes(), solve() and a destructor (ParisPencil.hpp:23-44).
 * Each of those maps onto one entry point here; the C++ facade
 * fishfry_amd/host/ParisPencil.hpp wraps them back into that class.
 *
 * Conventions: plain pointers and sizes only.  Device buffers are HIP device
 * pointers; a stream is a hipStream_t passed as void* (NULL = the null
 * stream).  Every entry point returns an ff_status and never exits the
 * process; ff_last_error() returns the calling thread's last message.
 */
#ifndef FISHFRY_H
#define FISHFRY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FF_ABI_VERSION 1

typedef enum ff_status {
  FF_OK = 0,
  FF_ERR_INVALID = 1,     /* bad argument (sizes, null pointers, options)          */
  FF_ERR_UNSUPPORTED = 2, /* valid for the reference, not supported here (m != 1)  */
  FF_ERR_BYTES = 3,       /* work arrays smaller than ff_plan_bytes()              */
  FF_ERR_HIP = 4,         /* a HIP runtime call failed                             */
  FF_ERR_NOMEM = 5,       /* device or host allocation failed                      */
  FF_ERR_REJECTED = 6,    /* serving: request does not fit the memory budget      */
  FF_ERR_TIMEOUT = 7,     /* serving: wait timed out                                */
  FF_ERR_NOT_FOUND = 8    /* serving: unknown request id                            */
} ff_status;

typedef enum ff_precision {
  FF_FP64 = 0,        /* FP64 storage and arithmetic in every pass                */
  FF_FP32_REFINE = 1  /* FP32 passes, FP64 residual + mean removal, refine_steps
                         corrections (0 = plain FP32 solve)                         */
} ff_precision;

/* ff_options.flags */
#define FF_PRESERVE_INPUT 1 /* FP64: the plan owns a third work array (ff_plan_bytes()
                               more device memory) and solve leaves density intact;
                               FP32 modes always preserve density                   */
#define FF_BLOCKED_LAYOUT 2 /* FP64: between the y passes store the half spectrum in
                               kh blocks of KT = 8 (y-blocked [kb][i][j][kt] after the y
                               forward and y inverse, x-blocked [kb][j][i][kt] after the
                               x pass) so every column pass writes contiguous tiles;
                               default: the reference's [i][j][kh] throughout
                               (ff_plan_describe "kt": 0 or KT)                       */

#define FF_SLAB_PIPELINE 4  /* FP64: run the z forward / y forward pair, and later the
                               y inverse / z inverse pair, one slab of x-planes at a time
                               so the second pass of each pair reads what the first
                               just wrote while it is still in the 256 MB Infinity
                               Cache (plain spectrum layout only)                      */

#define FF_SYMBOL_3PT 8     /* finite-difference symbol of the 3-point Laplacian on every
                               axis in place of the spectral filter: for mode k of an
                               n-point axis with spacing h = (hi - lo)/(n - 1) and
                               s = sin(pi k/n) the axis term is (4/h^2) s^2; the solve
                               inverts the 7-point stencil exactly.  Named, not
                               implemented, by the reference (make-frontier.sh:9-10,
                               make-summit.sh:7-8).  FP32 refinement computes its FP64
                               residual with the stencil in real space (one pass over
                               f and phi instead of an FP64 transform round trip)     */
#define FF_SYMBOL_5PT 16    /* the same with the 5-point (fourth-order) Laplacian: axis
                               term (4/h^2) s^2 (1 + s^2/3), 13-point stencil residual   */
#define FF_PENCIL_LAYOUT 32 /* every column pass transposes as it transforms, so it
                               reads strided columns and writes its lines contiguously:
                               [kh][i][j] after the y forward pass, [kh][j][i] after the
                               x pass, [i][kh][j] after the y inverse pass (which the z
                               inverse reads); the reference's pencil transposes
                               (ParisPencil.cpp:167-377) folded into the passes.
                               FP32 plans (on request): their FP32 passes, between the
                               two halves of the work area (FP64 residual passes stay
                               in place).  Default for FP64 plans of nx*ny*(nz/2+1) >=
                               2^25 points and nz <= 1024 (ff_plan_describe "pencil")  */
#define FF_PLAIN_LAYOUT 64  /* the reference's [i][j][kh] between all passes even where
                               FF_PENCIL_LAYOUT would be the default                    */

typedef struct ff_options {
  int precision;    /* ff_precision                                             */
  int refine_steps; /* FF_FP32_REFINE only: 0..8                                */
  int flags;        /* 0 or FF_PRESERVE_INPUT | FF_BLOCKED_LAYOUT | FF_SLAB_PIPELINE,
                       plus at most one of FF_SYMBOL_3PT / FF_SYMBOL_5PT          */
  int slab_planes;  /* FF_SLAB_PIPELINE: x-planes per slab (0: about 64 MB of
                       half spectrum per slab)                                  */
} ff_options;

/* One phase time stamp: the reference's TimeStamp (TimeStamp.hpp:7), a
 * (name, seconds) pair; seconds are on a steady host clock. */
typedef struct ff_stamp {
  char name[64];
  double seconds;
} ff_stamp;

#define FF_SOLVE_STAMPS 14 /* stamps written by ff_solve, ParisPencil.cpp:98-424 */

typedef struct ff_plan ff_plan;

/* ParisPencil::ParisPencil(n, lo, hi, m, id) -- ParisPencil.hpp:23,
 * ParisPencil.cpp:7-82.  n: global cells per axis; lo/hi: physical bounds
 * (hi is one cell below the periodic end); m: tasks per axis, must be {1,1,1}
 * (FF_ERR_UNSUPPORTED otherwise); id must be {0,0,0}.  opts may be NULL (FP64).
 * Builds radix plans and twiddle tables on the current HIP device. */
ff_status ff_plan_create(const int n[3], const double lo[3], const double hi[3], const int m[3],
                         const int id[3], const ff_options* opts, ff_plan** out);

/* ParisPencil::bytes() -- ParisPencil.hpp:30, formula ParisPencil.cpp:49-75.
 * Bytes each of density and potential must hold.  0 for a NULL plan. */
size_t ff_plan_bytes(const ff_plan* plan);

/* ParisPencil::solve(bytes, density, potential, stamps) -- ParisPencil.hpp:44,
 * ParisPencil.cpp:95-425.  density: [nx][ny][nz] FP64 input, clobbered
 * (FP64 mode) or preserved (FP32 modes); potential: receives phi in the same
 * layout.  bytes < ff_plan_bytes() -> FF_ERR_BYTES.  Work is enqueued on
 * stream and the call returns after it completes (the reference ends with a
 * device synchronize, ParisPencil.cpp:423).  stamps may be NULL; otherwise
 * it receives up to *n_stamps entries (FF_SOLVE_STAMPS with the reference's
 * phase names) and *n_stamps is set to the count written. */
ff_status ff_solve(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream,
                   ff_stamp* stamps, int* n_stamps);

/* Same as ff_solve but only enqueues the work on stream (no host wait, no
 * stamps): for callers that overlap solves or capture them in a HIP graph. */
ff_status ff_solve_async(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream);

/* No counterpart in the reference.  ff_solve_async for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the work area the solve's kernels need
 * beyond density and potential -- FP32 + refinement plans: spectrum, residual
 * and two sums (ff_plan_bytes() + 4*nx*ny*nz + 512 bytes or so); FP64 plans
 * with FF_PRESERVE_INPUT: a third ff_plan_bytes() array -- is the library's
 * static device array of 64 MiB (one per GPU, held until the process ends)
 * instead of memory the plan allocated.  The call allocates nothing and
 * calls no allocating routine; plans whose work area exceeds 64 MiB are
 * rejected with FF_ERR_INVALID.  The plan's device must be the current
 * device.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (other streams, other threads) race.  FP64 plans without
 * FF_PRESERVE_INPUT need no work area and behave as ff_solve_async. */
ff_status ff_solve_noalloc(const ff_plan* plan, size_t bytes, double* density, double* potential, void* stream);

/* ParisPencil::~ParisPencil -- ParisPencil.cpp:84-93.  NULL is a no-op. */
ff_status ff_plan_destroy(ff_plan* plan);

/* Thread-local text of the last error ("" if none). */
const char* ff_last_error(void);

/* The reference's per-array byte formula for any task grid m
 * (ParisPencil.cpp:24-26,49-75), without creating a plan. */
size_t ff_reference_bytes(const int n[3], const int m[3]);

/* Human-readable JSON description of a plan (radix plans, tiles, bytes). */
ff_status ff_plan_describe(const ff_plan* plan, char* buf, size_t len);

/* One pass of the solve, for per-pass parity checks and timing:
 * 0 z forward (real src -> complex dst), 1 y forward, 2 x forward*filter*
 * inverse, 3 y inverse, 4 z inverse + scale (complex src -> real dst).
 * Real arrays are [i][j][k].  The half spectrum is [i][j][kh] (the reference's
 * D2Z output layout); with FF_BLOCKED_LAYOUT pass 1 writes y-blocked, pass 2
 * reads y-blocked and writes x-blocked, pass 3 reads x-blocked and writes
 * y-blocked, pass 4 reads y-blocked.  Blocked layouts, with kh = kb*KT + kt,
 * nb = ceil(nh/KT), W = KT (nh - (nb-1)*KT for the last block):
 *   y-blocked (i, j, kh) at kb*nx*ny*KT + (i*ny + j)*W + kt
 *   x-blocked (i, j, kh) at kb*nx*ny*KT + (j*nx + i)*W + kt            */

/* Tuning overrides consulted when a plan is created (process-wide, thread
 * safe).  kind: 0 z forward rows, 1 y columns, 2 x columns, 3 z inverse
 * rows; n: transform length
 * (nz/2 for even-nz rows); precision: ff_precision of the pass (0 FP64,
 * 1 FP32); E: elements per thread; T: lines per workgroup (power of two for
 * columns).  E = 0 removes the override.  Invalid combinations are rejected
 * when the override is set. */
typedef struct ff_tune {
  int kind;
  int n;
  int precision;
  int E;
  int T;
  int flags; /* FF_TUNE_SPLIT_LDS | FF_TUNE_TWO_GROUPS (column passes) | FF_TUNE_LDS_TWIDDLES | FF_TUNE_GENERIC */
} ff_tune;

/* ff_tune.flags (column passes):
 * FF_TUNE_SPLIT_LDS   exchange real and imaginary parts separately through
 *                     half the LDS (more workgroups per CU, twice the barriers)
 * FF_TUNE_TWO_GROUPS  a workgroup transforms two groups of T columns one after
 *                     the other, so it reads and writes 2*T-column rows whole
 *                     (E = 8 or 16)                                          */
#define FF_TUNE_SPLIT_LDS 1
#define FF_TUNE_TWO_GROUPS 2
/* FF_TUNE_GENERIC     the kernel that reads its line plan at run time even
 *                     where one with the plan of (n, E, T) compiled in exists
 *                     (ff_tune_query reports FF_TUNE_GENERIC when no such
 *                     kernel exists for the configuration)                 */
#define FF_TUNE_GENERIC 4
/* FF_TUNE_LDS_TWIDDLES column or row passes with a compiled-in plan: the
 *                     workgroup copies half the twiddle table to LDS and the
 *                     transforms read it there instead of through the vector
 *                     memory path (more LDS per workgroup; results unchanged) */
#define FF_TUNE_LDS_TWIDDLES 8

ff_status ff_tune_set(const ff_tune* t);
void ff_tune_reset(void);
/* The configuration a plan created now would use for (kind, n, precision);
 * ncols is the column count of the pass (nh), used to cap T. */
ff_status ff_tune_query(int kind, int n, int precision, int ncols, ff_tune* out);
ff_status ff_solve_pass(const ff_plan* plan, int pass, const double* src, double* dst, void* stream);

/* No counterpart in the reference.  out = L_h phi in FP64 with the plan's
 * finite-difference stencil (FF_SYMBOL_3PT / FF_SYMBOL_5PT plans; periodic,
 * [nx][ny][nz] arrays of nx*ny*nz doubles): the operator the solve inverts,
 * so L_h(solve(f)) = f - mean(f).  FF_ERR_INVALID for spectral plans.
 * Enqueued on stream. */
ff_status ff_stencil_apply(const ff_plan* plan, const double* phi, double* out, void* stream);

/* ----------------------------------------------------------------------- */
/* Batched solves: many independent grids (uneven shapes, FP64 or FP32 +    */
/* refinement each) launched together.  No reference counterpart (the       */
/* reference solves one grid per process); each instance follows            */
/* ParisPencil::solve's contract and numerics.                              */

typedef struct ff_batch_item {
  int n[3];         /* cells per axis (each >= 2)                                */
  double lo[3];     /* physical bounds as in ff_plan_create                       */
  double hi[3];
  int precision;    /* ff_precision                                               */
  int refine_steps; /* FF_FP32_REFINE only                                        */
} ff_batch_item;

#define FF_BATCH_GRAPH 1      /* capture the fork/join launch sequence in a HIP graph on
                                 first solve and replay it while buffers are unchanged */
#define FF_BATCH_PERSISTENT 4 /* one persistent kernel over every instance: removed (its
                                 kernel did not complete on MI355X), rejected with
                                 FF_ERR_UNSUPPORTED; FF_BATCH_STEPS replaces it         */
#define FF_BATCH_GROUPED 8    /* instances of one shape, precision and refinement count
                                 run each pass as one launch (one grid row per instance,
                                 per-instance buffers and domains from a device table);
                                 the table is rebuilt, after a device synchronization,
                                 when the buffers change                                 */
#define FF_BATCH_STEPS 16     /* pass step s of every instance of one precision and
                                 refinement count is ONE launch of a multi-shape kernel:
                                 a device table holds each instance's pass arguments and
                                 workgroup range, and each workgroup runs the tile of the
                                 instance it falls in (no workgroup waits on another).
                                 Lines without a compiled-in multi-shape variant (lengths
                                 other than 32, 48, 64, 96, 128, 160, 256 and, for the z
                                 passes, nz other than 32 ... 256; Bluestein and odd nz)
                                 run that step as extra grouped launches on the same
                                 stream.  Exclusive with FF_BATCH_GROUPED; the table is
                                 rebuilt, after a device synchronization, when the
                                 buffers change                                          */

typedef struct ff_batch ff_batch;

/* Plans every instance; the instances run on a few internal streams forked
 * from, and joined back onto, the caller's stream. */
ff_status ff_batch_create(int count, const ff_batch_item* items, int flags, ff_batch** out);
/* Bytes each of instance index's density / potential must hold. */
size_t ff_batch_bytes(const ff_batch* b, int index);
/* Enqueue one solve of every instance on stream: density[i] (clobbered for
 * FP64 instances, preserved for FP32 ones) -> potential[i]; returns after
 * enqueueing.  FF_BATCH_GRAPH needs a non-default stream. */
ff_status ff_batch_solve(ff_batch* b, double* const* density, double* const* potential, void* stream);
ff_status ff_batch_wait(ff_batch* b, void* stream);
/* Autotuner: times every valid (elements per thread, lines per workgroup)
 * configuration of each distinct pass of the batch on the given buffers
 * (reps launches each; the buffers' contents are overwritten), records the
 * fastest with ff_tune_set and re-plans.  tuned: passes tuned; ms_before /
 * ms_after: batch solve time before and after. */
ff_status ff_batch_autotune(ff_batch* b, double* const* density, double* const* potential, void* stream, int reps,
                            int* tuned, double* ms_before, double* ms_after);
ff_status ff_batch_describe(const ff_batch* b, char* buf, size_t len);
ff_status ff_batch_destroy(ff_batch* b);

/* ----------------------------------------------------------------------- */
/* Serving: a request loop with admission by HBM footprint and continuous    */
/* batching (no reference counterpart).                                      */

typedef struct ff_server ff_server;

typedef struct ff_server_config {
  size_t hbm_budget;      /* device bytes reserved at start for request buffers      */
  int streams;            /* device streams groups are spread over (1..16)            */
  int max_batch;          /* requests dispatched together per size class              */
  double batch_window_us; /* longest wait for same-class companions while busy        */
  int flags;              /* FF_SERVE_LARGE_ALONE or 0                                */
} ff_server_config;

/* ff_server_config.flags: a group whose estimated work is large (about 3e8
 * point-passes: FP32 + 2 steps from about 260^3, FP64 from about 390^3)
 * runs with the GPU nearly to itself -- it is dispatched once nothing is in
 * flight, and while it runs only small groups (below 6e7 point-passes: FP64
 * up to about 230^3, FP32 + 2 steps up to about 150^3) are dispatched --
 * instead of sharing the GPU with the other streams' groups, which stretches
 * its latency. */
#define FF_SERVE_LARGE_ALONE 1

struct ff_field;

typedef struct ff_request {
  int n[3];                     /* grid extents (benchmark grid: lo = -pi, 2pi period) */
  int precision;                /* ff_precision                                        */
  int refine_steps;             /* FF_FP32_REFINE only                                 */
  const struct ff_field* field; /* f generated on the device from this descriptor, or  */
  const double* f_host;         /* f copied from host memory [n0][n1][n2]              */
  double* phi_host;             /* optional: receives phi [n0][n1][n2]                 */
  int want_norms;               /* field requests: relative L1/L2/Linf vs the exact phi */
} ff_request;

typedef struct ff_result {
  int status;        /* ff_status of the request                       */
  double latency_s;  /* submit -> completion seen by the scheduler     */
  double queue_s;    /* submit -> dispatch                             */
  double norms[3];   /* want_norms: relative L1, L2, Linf              */
  long long points;
} ff_result;

typedef struct ff_serving_stats {
  long long completed, rejected, groups;
  double p50_s, p99_s;      /* latency percentiles of completed requests  */
  double points_per_s;      /* completed points / (last done - first submit) */
  size_t arena_bytes, arena_peak_bytes;
  long long grouped_solves; /* requests solved as grouped launches: requests of
                               one shape, precision and step count in one
                               dispatch group run each pass step as one launch */
} ff_serving_stats;

/* Starts the scheduler thread on the current device (cfg NULL: 8 GiB arena,
 * 4 streams, groups of 16, 200 us window). */
ff_status ff_server_create(const ff_server_config* cfg, ff_server** out);
/* Device bytes a request occupies while admitted (its admission footprint):
 * density and potential, the FP32 refinement work area or (FP64 grids whose
 * z rows run in four steps) the z work array, and for want_norms with a field
 * descriptor the work area of the two norm reductions. */
size_t ff_server_footprint(const ff_request* r);
/* Queues a request; FF_ERR_REJECTED if its footprint exceeds the budget.
 * The field descriptor is copied; f_host / phi_host must stay valid until the
 * request completes. */
ff_status ff_server_submit(ff_server* s, const ff_request* r, uint64_t* id);
/* Waits for a request and returns its result; the result is handed out once
 * (a later wait on the same id, or one that loses the race with another
 * thread's, returns FF_ERR_NOT_FOUND). */
ff_status ff_server_wait(ff_server* s, uint64_t id, double timeout_s, ff_result* out);
ff_status ff_server_stats(ff_server* s, ff_serving_stats* out);
/* Drains the queue, then stops the scheduler and frees everything. */
ff_status ff_server_destroy(ff_server* s);

/* ----------------------------------------------------------------------- */
/* Driver kernels of FishFry (FishFry.hpp:14-101).                          */

#define FF_MAX_MODES 32

typedef struct ff_field {
  int n[3];       /* local extents ni, nj, nk (FishFry.hpp:112-114)                  */
  double lo[3];   /* xLo, yLo, zLo (FishFry.hpp:137-139)                             */
  double d[3];    /* dx, dy, dz (FishFry.hpp:109-111)                                */
  double shift[3];/* phase shift added to x, y, z (0 for the reference benchmark)    */
  double scale;   /* amplitude (1 for the reference benchmark)                       */
  double dd[3];   /* filter constants for the exact potential of extra modes         */
  int nmodes;     /* extra Fourier modes a*cos(k.x) + b*sin(k.x), k integer          */
  int mode_k[FF_MAX_MODES][3];
  double mode_a[FF_MAX_MODES];
  double mode_b[FF_MAX_MODES];
} ff_field;

/* Fill ff_field with the reference benchmark's grid for n at 1x1x1 tasks:
 * lo = -pi, d = 2pi/n, shift 0, scale 1, no extra modes (FishFry.hpp:108-139). */
void ff_field_default(const int n[3], ff_field* out);

/* The init kernel (FishFry.hpp:14-45): f = scale * lap(phi)(x + shift) plus
 * the extra modes, written to f ([ni][nj][nk], FP64).  Enqueued on stream. */
ff_status ff_field_init(const ff_field* field, double* f, void* stream);

/* The diff kernel (FishFry.hpp:57-101) with a deterministic two-stage
 * reduction: out[0] = sum|e|, out[1] = sum e^2, out[2] = max|e|, where e is
 * the exact potential minus phi.  phi == NULL evaluates the exact potential
 * against zero (the norm scales of FishFry.hpp:146-159).  Waits on stream. */
ff_status ff_field_diff(const ff_field* field, const double* phi, double out[3], void* stream);

#ifdef __cplusplus
}
#endif
#endif
