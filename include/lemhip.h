/*
 * lemhip.h — C-ABI of the MI355X (gfx950) landscape-evolution timestep.
 *
 * Drop-in boundary for the per-timestep hot path of PARALLEM's lem.cpp driver
 * (lem.cpp:83-97).  The reference binds this path as C++ free functions
 * taking (Data* data, Data* device[, int iter]); a context replaces the
 * "device" Data mirror and keeps all state resident in HBM across phases and
 * timesteps, so the reference's per-phase allocate/copy/free calls
 * (memory_dev.cu:33-240) become no-ops.  Host buffers passed here are
 * caller-owned, in the reference's host layout: row-major W*H cells, double
 * fields, int fd/SFD/mask, and Slopes as 8 doubles per cell
 * (Slopes[8*i + d], d = N, NE, E, SE, S, SW, W, NW).
 *
 * Conventions: every call returns 0 on success or a negative LEMHIP_E_* code
 * (never exits the process); lemhip_last_error() describes the last failure.
 * One context per host thread; work runs on the context's own HIP stream and
 * every call that fills a host buffer or returns results is synchronous at
 * the boundary (lemhip_slot_set_climate only enqueues its kernel: every later
 * call on the context sees its fields).
 */
#ifndef LEMHIP_H
#define LEMHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LEMHIP_OK 0
#define LEMHIP_E_ARG (-1)
#define LEMHIP_E_HIP (-2)
#define LEMHIP_E_NOMEM (-3)
#define LEMHIP_E_NOT_CONVERGED (-4)
#define LEMHIP_E_STATE (-5)
#define LEMHIP_E_SIZE (-6)   /* input larger than a fixed capacity */

/* Precision forms. */
#define LEMHIP_FP64 0        /* reference dtypes, bit-exact with the FP64 restatement */
#define LEMHIP_FP32_KAHAN 1  /* FP32 bulk, FP64 dem/soilT, float-float compensated fa */

/* Model parameters: the positional par file (default.par:3-35, read by
 * io.cpp:376-580) plus the grid's MapInfo (MapInfo.h:4-12). */
typedef struct {
    int width, height;
    double cellsize;
    int flowdirectiontype;       /* 0 SFD, 1 MFD */
    double stoneval, fineval, soilTval, soilMval, soilBval, nutval;
    int soil_struct, profile_permeability;
    double Qthreshold, Gslope, kappa, Gflowlim, Flow_A, Flow_B, max_incision;
    double ddk, dck, dgk, temp_inflex, temp_sens;
    double rain, temp;           /* runin rain (mm/yr) and temperature (C) */
    int runiniter, start_iter;   /* correctmfdflow's first-iteration outlet rule (mfd_simple.cu:336) */
} lem_params;

/* Per-step totals the reference writes to its budgets CSV (io.cpp:233-263). */
typedef struct {
    double FA_max, totE, totI, totG, totD, totweatherC, totweatherP, totbio;
    int outletcellidx;
    int sinkcount;
    int sinks_pass1;
    int nlabels, nedges;
    int plateau_rounds1, plateau_rounds2;
    int fa_uncomputed, sed_uncomputed;
    /* calcwater (runoffweight.cu:356-377): sum of the runoff weights over the
     * mask cells, the "Total amount of water in grid for runoff" log line */
    double totwater;
    /* Input diagnostics fixed at initialisation (no reference counterpart):
     * phantom_cells counts valid cells in rows 1..H-1 of the first
     * P = (W/16 + 1)*16 - W columns, which the reference's label kernels
     * (newflow.cu:358-360, 434-436) also visit from threads past the end of
     * the previous row; when it is 0 every legal schedule of those kernels
     * gives this library's bounds-checked labels.  soilT_negative counts valid
     * cells whose createSoilTfromformula thickness (memory.cu:207-208) is
     * negative (0 < dem < 400), whatever soilT_mode then does with them. */
    int phantom_cells;
    int soilT_negative;
    /* FP32 form with lem_tuning.fp32_tree_sediment: the cells of the grid's
     * last sediment pass left open by the engine's first rounds and routed by
     * the contraction (0 when the engine ran to the end) */
    int sed_open;
} lem_summary;

/* Fields for upload/download (host layout as described above). */
enum lem_field {
    LEM_F_DEM = 0, LEM_F_MASK = 1, LEM_F_SLOPES = 2, LEM_F_SLOPEPTR = 3,
    LEM_F_SOILT = 4, LEM_F_SOILM = 5, LEM_F_STONE = 6, LEM_F_FINES = 7,
    LEM_F_NUT = 8, LEM_F_SOILB = 9, LEM_F_TOTB = 10,
    LEM_F_RAIN = 11, LEM_F_TEMP = 12, LEM_F_LAST_RAIN = 13, LEM_F_LAST_TEMP = 14,
    LEM_F_FD = 15, LEM_F_SFD = 16, LEM_F_RUNOFF = 17, LEM_F_FA = 18,
    LEM_F_ERO = 19, LEM_F_INCI = 20, LEM_F_GELI = 21, LEM_F_DEPO = 22,
    LEM_F_DZ = 23, LEM_F_WC = 24, LEM_F_WP = 25,
    LEM_F_COUNT = 26
};

typedef struct lemhip_ctx lemhip_ctx;

/* Replaces createDeviceSpace (memory_dev.cu:293-321) and the per-phase
 * setdevicespace_* / cleardevicespace_* calls (memory_dev.cu:33-240): allocates
 * every device array once, on device `device_id`. */
int lemhip_create(const lem_params *p, int precision, int device_id, lemhip_ctx **out);
/* (lemhip_create and lemhip_create_atlas: when the call fails after the
 * context was made, *out is still set so that lemhip_last_error can say why;
 * a non-null *out must then be passed to lemhip_destroy, which frees what was
 * allocated.) */
/* Replaces clearDeviceSpace (memory_dev.cu:324-347). */
int lemhip_destroy(lemhip_ctx *ctx);
const char *lemhip_last_error(const lemhip_ctx *ctx);

/* Host state initialisation with the reference's init semantics, done on the
 * device: createmask (memory.cu:217-242), init_setoutletcell (Data.cpp:38-66),
 * createSoilTfromformula (memory.cu:197-214; soilT_mode 1 clamps it at 0),
 * SetClimateMatrices(-10000) (Data.cpp:211-303), setProcessMatrices
 * (Data.cpp:115-185), copyMask / copylastclimate (memory_dev.cu:242-289).
 * rain may be NULL (rain gradient of SetClimateMatrices) or a per-cell mm/yr
 * field that replaces it. */
int lemhip_init_state(lemhip_ctx *ctx, const double *dem, const double *rain, int soilT_mode);

/* Upload/download one field (host layout). */
int lemhip_upload(lemhip_ctx *ctx, int field, const void *host);
int lemhip_download(lemhip_ctx *ctx, int field, void *host);
int lemhip_get_outlet(lemhip_ctx *ctx, int *outlet);
int lemhip_set_outlet(lemhip_ctx *ctx, int outlet);

/* Phases of one timestep (lem.cpp:83-97). */
int lemhip_flow_direction(lemhip_ctx *ctx, int iter); /* cuFlowDirection, routeflow.cu:686-975 */
int lemhip_runoff(lemhip_ctx *ctx);                   /* computeRunOffWeights + calcpropsSFD, runoffweight.cu:61-138 */
int lemhip_flow_accum(lemhip_ctx *ctx, int iter);     /* correctmfdflow, mfd_simple.cu:238-384 */
int lemhip_erosion(lemhip_ctx *ctx, int iter);        /* erosionGPU, erosion.cu:69-186 */
int lemhip_step(lemhip_ctx *ctx, int iter);           /* all four, lem.cpp:83-97 */
int lemhip_run(lemhip_ctx *ctx, int iter0, int nsteps);

/* Climate forcing: SetClimateMatrices(iter) (Data.cpp:231-271) with the given
 * scalar rain/temp, and copytolastclimate (memory_dev.cu:363-373). */
int lemhip_set_climate(lemhip_ctx *ctx, double rain, double temp);
int lemhip_copy_to_last_climate(lemhip_ctx *ctx);

/* Totals of the last step (the Thrust reductions of erosion.cu:96-115,
 * mfd_accum.cu:459-461, eroincidep.cu:446-455, updates.cu:218-220,
 * mfd_simple.cu:354-355). */
int lemhip_summary(lemhip_ctx *ctx, lem_summary *out);

/* Atlas contexts: many DEMs of different sizes stepped together in one grid
 * of atlas_w x atlas_h cells (no reference counterpart; the reference runs
 * one DEM per process, lem.cpp:54-103).  Each DEM ("slot") occupies its own
 * rectangle, padded by one nodata row and column, so its results are exactly
 * those of a single-DEM context on the same inputs.  The per-slot calls mirror
 * the single-DEM calls above, with host buffers in the DEM's own w*h layout
 * and the outlet in the DEM's own row-major numbering.  Up to 255 slots.
 * slot_open fails with LEMHIP_E_NOMEM when no rectangle is free; slot_close
 * returns the rectangle.  step_slots takes one iteration number per slot
 * (entries of unused slots are ignored).  The single-DEM calls
 * init_state/upload/download/get_outlet/set_outlet/summary return
 * LEMHIP_E_STATE on an atlas context. */
int lemhip_create_atlas(const lem_params *p, int atlas_w, int atlas_h, int precision, int device_id,
                        lemhip_ctx **out);
int lemhip_slot_open(lemhip_ctx *ctx, int width, int height, int *slot);
int lemhip_slot_init(lemhip_ctx *ctx, int slot, const double *dem, const double *rain, int soilT_mode);
int lemhip_slot_close(lemhip_ctx *ctx, int slot);
int lemhip_slot_rect(lemhip_ctx *ctx, int slot, int *x0, int *y0, int *width, int *height);
int lemhip_slot_upload(lemhip_ctx *ctx, int slot, int field, const void *host);
int lemhip_slot_download(lemhip_ctx *ctx, int slot, int field, void *host);
int lemhip_slot_get_outlet(lemhip_ctx *ctx, int slot, int *outlet);
int lemhip_slot_set_outlet(lemhip_ctx *ctx, int slot, int outlet);
int lemhip_slot_set_climate(lemhip_ctx *ctx, int slot, double rain, double temp);
int lemhip_slot_summary(lemhip_ctx *ctx, int slot, lem_summary *out);
int lemhip_step_slots(lemhip_ctx *ctx, const int *iters);

/* Checkpoint / restart of one DEM (slot 0 of a single-DEM context, or an
 * atlas slot): replaces retriveProcessMatrices (ioGDAL.cpp:662-804), which
 * reads the process matrices of a restart back from grids written by
 * savegrids (ioGDAL.cpp:208-324).  The snapshot is a self-describing byte
 * buffer: dem, soilT, mask, the sticky Slopes (which the reference's restart
 * leaves out), soilM, stone, fines, nut, soilB, TotB, the four climate
 * fields and the outlet, in the context's own element types.  Restoring it
 * into a context of the same precision form and DEM size (an open slot, or a
 * single-DEM context straight after lemhip_create) reproduces the state
 * exactly, so the resumed run is bit-identical to an unbroken one.
 * snapshot_size needs an open slot; snapshot needs an initialised one. */
int lemhip_snapshot_size(lemhip_ctx *ctx, int slot, size_t *bytes);
int lemhip_snapshot(lemhip_ctx *ctx, int slot, void *buf, size_t bytes);
int lemhip_restore(lemhip_ctx *ctx, int slot, const void *buf, size_t bytes);

/* Steps several contexts at once (no reference counterpart): each from its
 * own host thread on its own stream, so their kernels overlap on the device.
 * iters[k] is the iteration of single-DEM context k; slot_iters[k] the
 * per-slot iterations of atlas context k (as lemhip_step_slots takes them).
 * Contexts must be distinct and on the same device or on different ones;
 * returns the first failure's code (each context's lemhip_last_error says
 * more). */
int lemhip_step_group(lemhip_ctx *const *ctxs, int n, const int *iters, const int *const *slot_iters);

/* Launch tuning (no reference counterpart).  The dependency engine of the
 * two accumulation passes runs in tiles of dag_tile_w x dag_tile_h cells
 * (64x16, 32x32, 16x64, 128x8, 64x32 or 32x64) and checks convergence on the
 * host every dag_batch round launches; plateau routing advances
 * plateau_rounds_per_launch (4, 8 or 16) synchronous rounds per launch.
 * Results never depend on these choices.  lemhip_set_tuning cancels a
 * pending autotune.  lemhip_autotune makes the next step
 * time every candidate on the context's own data (the timed passes are
 * idempotent), keep the fastest (for the plateau rounds: each count timed
 * twice, keeping its faster time, and 8 per launch unless another count is
 * at least 10% faster), and record the times in tuned_ms:
 * [0..5] flow accumulation per tile shape, [8..10] plateau routing per rounds
 * per launch, [11..13] flow accumulation with batches 4, 8, 32.  In the FP32
 * form under single flow directions with fp32_integer_fa (below), flow
 * accumulation does not use the engine; [0..5] and [11..13] then time
 * sediment routing, the pass the tile shape serves (its in-place fields are
 * saved and restored around the timed runs), unless fp32_tree_sediment is
 * on (its engine rounds have a fixed tile shape; nothing is timed). */
typedef struct {
    int dag_tile_w, dag_tile_h;
    int dag_batch;
    int plateau_rounds_per_launch;
    double tuned_ms[16];
    /* FP32 form with single flow directions: 1 (default) accumulates flow as
     * exact integer subtree sums, 0 runs the dependency engine as the other
     * forms do; both stay within the form's stated bounds */
    int fp32_integer_fa;
    /* 1: the straight-line kernel sequences of the erosion phase (part 1 with
     * its totals; the totals, part 2 and the outlet bookkeeping) are captured
     * once into HIP graphs and replayed each step (recaptured when the band
     * of rows or the slot count changes); 0 (default): launched directly.
     * Results are identical. */
    int use_graphs;
    /* FP32 form with single flow directions: n = 1..4 (default 3) routes
     * sediment with n rounds of the dependency engine on a fixed 32 x 32
     * tiling (the first visits every tile and completes the cells whose
     * upstream lies inside their tile), then contracts the rest of the donor
     * forest (unary chains composed as clamped affine functions of the stone
     * fraction, in synchronous rounds); 0 runs the engine's rounds to the end
     * as the FP64 form does.  The same cells compute, every
     * field but stone is bit-identical between the two, and stone values
     * differ by a few float ulps where a composed chain reached them.  The
     * tile shape of those rounds is fixed (atlas rectangles are placed at
     * multiples of 32, so a DEM's results do not depend on its rectangle) */
    int fp32_tree_sediment;
} lem_tuning;
int lemhip_get_tuning(lemhip_ctx *ctx, lem_tuning *out);
int lemhip_set_tuning(lemhip_ctx *ctx, const lem_tuning *t);
int lemhip_autotune(lemhip_ctx *ctx);

/* Live timing with HIP events on the context stream (no reference
 * counterpart; the reference logs phase times from cudaEvents at
 * routeflow.cu:955-971, mfd_simple.cu:376-380, mfd_accum.cu:474-480,
 * erosion.cu:171-181).  Slots: 0 FD, 1 FlowDirs kernel, 2 plateau rounds,
 * 3 depression routing, 4 runoff kernel, 5 flow accumulation, 6 MOD part 1
 * kernel, 7 sediment routing, 8 MOD part 2 kernel, 9 reductions, 10 step,
 * 11 flow-accumulation round kernels only, 12 sediment round kernels only,
 * 13 plateau round kernels only, 14 spill-height kernels only (the Borůvka
 * phases, without the host reads between them), 15 sediment contraction
 * round kernels only (slot 12 then holds the engine's first rounds).
 * enable resets the accumulators; read returns milliseconds and launch
 * counts per slot. */
int lemhip_profile_enable(lemhip_ctx *ctx, int on);
int lemhip_profile_read(lemhip_ctx *ctx, int n, double *ms, long long *launches);

/* Test hook with no counterpart in the reference: evaluates the device FP64
 * math routines (fn 0 exp, 1 log, 2 log10, 3 pow, 4 tanh, 5 expm1) on n inputs. */
int lemhip_math_eval(int fn, int n, const double *x, const double *y, double *out);

/* The same evaluation without any allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Inputs and output
 * are host arrays as above; the device staging is a fixed array of
 * 3 * LEMHIP_NOALLOC_MAX doubles per device, so n above LEMHIP_NOALLOC_MAX
 * returns LEMHIP_E_SIZE and nothing is run.  Every call on a device shares
 * that device's staging: calls that overlap on one device (from several host
 * threads) race. */
#define LEMHIP_NOALLOC_MAX 65536
int lemhip_math_eval_noalloc(int fn, int n, const double *x, const double *y, double *out);

/* Device properties for measurement code (no reference counterpart). */
int lemhip_device_info(int device_id, char *name, int name_len, int *cu_count, double *mem_gb);
/* Free and total device memory in bytes (no reference counterpart). */
int lemhip_mem_info(int device_id, double *free_bytes, double *total_bytes);
/* Device bytes a context allocates per grid cell in the given precision form
 * (atlas != 0: an atlas context, which adds a slot id per cell), for
 * admission by footprint (no reference counterpart).  The edge and label
 * arrays of depression routing grow with the number of sinks and are not
 * included; *grow_per_cell bounds them per cell for DEMs with at most one
 * sink label per 16 cells.  The FP32 form's fixed figure includes the 64
 * bytes per cell of sediment records (lem_tuning.fp32_tree_sediment),
 * allocated on first use. */
int lemhip_bytes_per_cell(int precision, int atlas, double *fixed_per_cell, double *grow_per_cell);

#ifdef __cplusplus
}
#endif
#endif
