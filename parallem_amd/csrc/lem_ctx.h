// lem_ctx.h — the context that owns one DEM's device-resident state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "lem_device.h"
#include "../../include/lemhip.h"

// live timing of selected kernels/phases with HIP events on the context stream
enum {
    PROF_FD = 0, PROF_FLOWDIRS, PROF_PLATEAU, PROF_FLOOD, PROF_RUNOFF, PROF_FA, PROF_MOD1, PROF_SED,
    PROF_MOD2, PROF_REDUCE, PROF_STEP, PROF_FA_K, PROF_SED_K, PROF_PLAT_K, PROF_MM_K, PROF_CS_K, PROF_COUNT
};
#define PROF_MAXP 256
struct ProfSlot {
    hipEvent_t a[PROF_MAXP], b[PROF_MAXP];
    int n = 0;
    double ms = 0.0;
    long launches = 0;
    long events = 0;
};

// plateau tiles: PTW x PTH cells advanced plat_k synchronous rounds per launch.
// At 8 rounds per launch a 48 x 24 tile has a 64 x 40 region: each wave holds
// one region row and the 512 threads hold exactly 5 cells each.
#define PTW 48
#define PTH 24

// per-cell status of the dependency engine
enum : uint8_t { ST_PENDING = 0, ST_DONE = 1, ST_DEAD = 2, ST_RING_WAIT = 3 };

// Shelf packing of DEM rectangles into the atlas grid.  Each request is
// padded by one nodata column on the right and one nodata row below, so no
// two DEMs are 8-neighbours (an off-grid neighbour and a nodata neighbour
// behave the same on the whole path).
struct ShelfPack {
    struct Shelf {
        int y, h;
        std::vector<std::pair<int, int>> free;  // [x0, x1) intervals, sorted
    };
    int W = 0, H = 0, top = 0;
    std::vector<Shelf> shelves;
    bool alloc(int w, int h, int *x, int *y);
    void release(int x, int y, int w, int h);
};

// A straight-line sequence of launches on the context stream, captured once
// into a HIP graph and replayed while its launch shapes stay the same (key).
struct GraphSeg {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long key = -1;
};

struct lemhip_ctx {
    lem_params p;
    int prec = LEMHIP_FP64;
    int dev = 0;
    int ncu = 256;        // compute units of the device
    hipStream_t s = nullptr;
    int W = 0, H = 0;     // grid rows/cells a step covers (an atlas steps only its used rows)
    long N = 0;
    int Hfull = 0;        // allocated rows
    long NP = 0;          // allocated cells = stride of the Slopes planes
    int nslot = 1;        // slots (DEMs) of the atlas
    int slot_kbmax = 1;   // max slot_kb over active slots
    uint8_t *demid = nullptr;     // per-cell slot id (nullptr: one slot covering the grid)
    DemDesc *ddesc = nullptr;     // device copy of hdesc
    DemDesc hdesc[LEM_MAX_SLOTS + 1] = {};   // active rectangles (w == 0: inactive); entry 255 stays empty
    DemDesc hrect[LEM_MAX_SLOTS] = {};       // reserved rectangles
    uint8_t hstate[LEM_MAX_SLOTS] = {};      // 0 free, 1 reserved, 2 active
    // per-slot input diagnostic, from the host DEM at initialisation:
    //  phantom: valid cells in rows 1..h-1 of columns < P = (w/16 + 1)*16 - w,
    //           which the reference's label kernels also visit from threads
    //           past the row end (newflow.cu:358-360, 434-436); 0 means every
    //           legal schedule of those kernels equals the bounds-checked one
    int hphantom[LEM_MAX_SLOTS] = {};
    bool atlas = false;
    ShelfPack pack;
    DemScal *dsc = nullptr;       // device per-slot scalars (LEM_MAX_SLOTS + 1 entries)
    DemScal *hdsc = nullptr;      // pinned host mirror
    int *slot_flags = nullptr;    // device per-slot first-iteration flags
    int *hflags = nullptr;        // pinned host mirror
    uint8_t *labdem = nullptr;    // slot of each sink label
    void *dstage = nullptr;       // device staging for slot initialisation
    size_t dstage_bytes = 0;
    size_t tsz = 8;            // bytes of one bulk element (8 FP64, 4 FP32 form)
    std::string err;

    // ---- persistent state (HBM-resident across steps) ----
    double *dem = nullptr;     // f64 in every form
    double *soilT = nullptr;   // f64 in every form (sediment mass)
    uint8_t *mask = nullptr;
    void *slopes = nullptr;    // 8 planes of N (d-major), T
    void *soilM = nullptr, *stone = nullptr, *fines = nullptr, *nut = nullptr, *soilB = nullptr, *TotB = nullptr;
    void *rain = nullptr, *temp = nullptr, *last_rain = nullptr, *last_temp = nullptr;

    // ---- per-step products ----
    uint8_t *fd = nullptr, *SFD = nullptr;
    void *w = nullptr;         // runoff weight, T
    void *fa = nullptr;        // f64 (FP64) or float2 hi/lo (FP32 form)
    void *slopeptr = nullptr, *ero = nullptr, *inci = nullptr, *geli = nullptr, *depo = nullptr;
    void *dz = nullptr, *wC = nullptr, *wP = nullptr;

    // ---- scratch ----
    float *sp0 = nullptr, *sp1 = nullptr;   // plateau distances, two buffers per tile
    int *pl_tiles = nullptr;                 // per plateau tile: has-candidates, stamp, 3 rotating lists, buffer version
    void *pl_ctl = nullptr;
    int pl_launch = 0;
    int *cand = nullptr;       // plateau candidates (also holds pass-2 list)
    int *ws = nullptr;         // watershed labels
    int *scan_tmp = nullptr;   // sink flags / prefix
    double *z2 = nullptr;      // raised dem (transient)
    void *cub_tmp = nullptr;
    size_t cub_bytes = 0;
    unsigned long long *hkeys = nullptr, *hvals = nullptr;  // edge hash table
    long hcap = 0;
    uint32_t *elo = nullptr, *ehi = nullptr;
    unsigned long long *ew = nullptr;                        // compacted edges (weight encoded)
    uint32_t *elo2[2] = {nullptr, nullptr}, *ehi2[2] = {nullptr, nullptr};   // edges still able to lower a value
    unsigned long long *ew2[2] = {nullptr, nullptr};
    long ecap = 0;
    unsigned long long *rmm = nullptr;                       // minimax per label
    uint32_t *comp = nullptr;                                // component of labels cut off from label 0
    unsigned long long *cmax = nullptr;
    uint8_t *seen = nullptr;
    long lcap = 0;
    uint8_t *status = nullptr;  // DAG engine status
    uint8_t *dmask = nullptr;   // DAG donor bits (pull order)
    int *dag_list[3] = {nullptr, nullptr, nullptr};   // rotating lists of woken tiles
    int *dag_wake = nullptr;                          // per tile: last round it was queued for, DAG_TILE_DONE when finished
    void *dag_ctl = nullptr;
    int dag_round = 0;
    int dag_rounds_last = 0;
    int plat_k = 8;             // plateau rounds per launch (4, 8 or 16)
    // lemhip_autotune sets both; each tuner clears its own flag once it has run
    bool tune_dag_pending = false;    // the next pass on the engine times tile shapes and batches
    bool tune_plat_pending = false;   // the next plateau pass on a raised surface times rounds per launch
    void *tcache = nullptr;     // MOD part 2's calcite terms per cell (2 x N, T), from the temperatures
    bool tcache_ok = false;     // false after anything that may change a temperature
    double tune_ms[16] = {};    // last tuning: per-candidate times
    int dag_shape = 0;          // tile shape id of the dependency engine
    int dag_batch = 16;         // round launches between host checks
    long dag_tiles_max = 0;     // tile arrays are sized for the shape with the most tiles
    double *partials = nullptr; // deterministic reductions
    // FP32 single-direction flow accumulation in integers (lem_fa32.hip)
    int *fx_shift = nullptr;          // device: fixed-point shift per slot
    int hfx_shift[LEM_MAX_SLOTS + 1] = {};
    int *fx_tag_cell = nullptr;       // cells below the outlet's west neighbour
    unsigned long long *fx_tag_sum = nullptr;
    long fx_tag_cap = 0;
    bool fx_enabled = true;           // false: the dependency engine in every form
    // FP32 sediment routing: the engine's first round, then a contraction of
    // the donor graph (lem_sed32.hip)
    void *cs_rec = nullptr;           // 2 x N unary records, allocated on first use
    int cs_rounds = 3;                // engine rounds before the contraction (0: the engine to the end)
    // MOD's straight-line segments (part 1 with its totals; the totals,
    // part 2 and the outlet bookkeeping) replayed from HIP graphs
    bool use_graphs = false;
    bool capturing = false;
    GraphSeg gseg[2];
    LemScalars *scal = nullptr;   // device
    LemScalars *hscal = nullptr;  // pinned host mirror
    void *stage = nullptr;        // pinned host staging for upload/download
    size_t stage_bytes = 0;
    int *hint = nullptr;          // pinned host ints for small readbacks

    // counters known on the host after each phase
    int h_rounds1 = 0, h_rounds2 = 0, h_nlab = 0, h_nedges = 0, h_sinkcount = 0, h_sinks1 = 0;
    int h_fa_unc = 0, h_sed_unc = 0;
    int h_sed_open = 0;   // sediment cells the contraction took over (last pass)

    // profiling
    bool prof_on = false;
    bool prof_ready = false;
    ProfSlot prof[PROF_COUNT];

    // tiling of the dependency engine
    int tw = 64, th = 16;
    int ntx = 0, nty = 0;
};

#define LEM_CHECK(ctx, call)                                                              \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return LEMHIP_E_HIP;                                                          \
        }                                                                                 \
    } while (0)

#define LEM_LAUNCHED(ctx)                                                                 \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            (ctx)->err = std::string("kernel launch: ") + hipGetErrorString(e_);          \
            return LEMHIP_E_HIP;                                                          \
        }                                                                                 \
    } while (0)

#define LEM_TRY(x)                  \
    do {                            \
        int r_ = (x);               \
        if (r_ != LEMHIP_OK) return r_; \
    } while (0)

static inline void prof_begin(lemhip_ctx *c, int id)
{
    if (!c->prof_on || c->capturing) return;
    ProfSlot &p = c->prof[id];
    if (p.n < PROF_MAXP) (void)hipEventRecord(p.a[p.n], c->s);
}
static inline void prof_end(lemhip_ctx *c, int id, long launches = 1)
{
    if (!c->prof_on || c->capturing) return;
    ProfSlot &p = c->prof[id];
    if (p.n < PROF_MAXP) {
        (void)hipEventRecord(p.b[p.n], c->s);
        p.n++;
        p.launches += launches;
    }
}

// phase entry points implemented in the .hip files
int lem_fd_phase(lemhip_ctx *c, int iter);
int lem_runoff_phase(lemhip_ctx *c);
int lem_calcwater(lemhip_ctx *c);
int lem_fa32_sfd(lemhip_ctx *c, int *fallback);
int lem_sed32(lemhip_ctx *c, int *fallback);
int lem_sed_first(lemhip_ctx *c, int rounds);
int lem_fa_phase(lemhip_ctx *c, int iter);
int lem_mod_phase(lemhip_ctx *c, int iter);
int lem_sync_scalars(lemhip_ctx *c);
// per-slot reductions (lem_mod.hip)
int lem_slot_lowest(lemhip_ctx *c, const double *z, int s0, int ns);
int lem_slot_first_lowest(lemhip_ctx *c, const double *z, int s0, int ns);
int lem_read_slots(lemhip_ctx *c);
int lem_dag_shape_count();
int lem_tune_dag(lemhip_ctx *c);
int lem_tune_sed(lemhip_ctx *c);
void lem_dag_shape_dims(int id, int *tw, int *th);
static inline int lem_grid1d(long n, int block = 256, int cap = 8192)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}
