// lem_device.h — device-side constants and helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/lem_math.h"

// Direction index d = 0..7 runs N, NE, E, SE, S, SW, W, NW (routeflow.cu:701-702);
// ESRI codes per direction (routeflow.h:32).
__device__ __constant__ static const int8_t kDX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ __constant__ static const int8_t kDY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
__device__ __constant__ static const uint8_t kCODE[8] = {64, 128, 1, 2, 4, 8, 16, 32};
// pull order of CalcMFD / MFDsedbud: E, SE, S, SW, W, NW, N, NE (mfd_simple.cu:95-165)
__device__ __constant__ static const int8_t kPULL[8] = {2, 3, 4, 5, 6, 7, 0, 1};

#define LEM_NODATA (-9999.0)

// code a neighbour in direction d must carry to flow into the centre cell
__device__ __forceinline__ uint8_t donor_code(int d) { return kCODE[(d + 4) & 7]; }
// the same tables as compile-time functions, for fully unrolled loops
__host__ __device__ constexpr int cdx(int d) { return (d >= 1 && d <= 3) ? 1 : ((d >= 5) ? -1 : 0); }
__host__ __device__ constexpr int cdy(int d) { return (d == 7 || d <= 1) ? -1 : ((d >= 3 && d <= 5) ? 1 : 0); }
__host__ __device__ constexpr int cpull(int k) { return (k + 2) & 7; }   // pull order E, SE, S, SW, W, NW, N, NE

// direction index of an ESRI code (0 if not a single code)
__device__ __forceinline__ int code_dir(int code)
{
    switch (code) {
    case 64: return 0; case 128: return 1; case 1: return 2; case 2: return 3;
    case 4: return 4; case 8: return 5; case 16: return 6; default: return 7;
    }
}

// order-preserving map double -> uint64 (for atomic min/max on doubles)
__device__ __host__ __forceinline__ uint64_t ord_enc(double x)
{
    uint64_t u = lm_to_bits(x);
    return (u & 0x8000000000000000ULL) ? ~u : (u | 0x8000000000000000ULL);
}
__device__ __host__ __forceinline__ double ord_dec(uint64_t u)
{
    return lm_from_bits((u & 0x8000000000000000ULL) ? (u & 0x7FFFFFFFFFFFFFFFULL) : ~u);
}

// One DEM ("slot") of an atlas context: its rectangle in the atlas grid
// (w == 0: slot unused).  Cells outside every rectangle are nodata.
struct DemDesc {
    int x0, y0, w, h;
};
#define DEM_NONE 255
// sediment prep states: pending (DAG_PENDING, or DAG_PENDING - 1 for a cell
// whose stone mixes, where the FP32 contraction asks for it), -1 dead
#define DAG_PENDING 0x7FFFFFFF
#define LEM_MAX_SLOTS 255

// per-slot scalars (the reference's per-DEM host bookkeeping)
struct DemScal {
    double lowest;        // min over mask cells of dem, from 1000 (routeflow.cu:171-184)
    double base_level;    // dem[outlet] - 0.00012 (updates.cu:151)
    double FA_max;
    double totE, totI, totG, totD, totwC, totwP, totbio;
    double totwater;      // sum of runoff weights over mask cells (runoffweight.cu:356-377)
    unsigned long long lowest_enc, famax_enc;
    int outlet;           // atlas index of the outlet cell
    int argidx;
    int sinks_pass1;
    int sinkcount;
    int zero_seen;
    int first_iter;       // this step is the slot's first iteration (mfd_simple.cu:336)
    unsigned long long fx_outlet_sum;   // the outlet's integer flow sum (FP32 single-direction accumulation)
    int soilT_negative;   // valid cells with a negative formula soil thickness at initialisation
};

// Per-slot reductions split a slot's rows over slot_kb(h) blocks (row k goes
// to block k mod slot_kb(h)); the split depends only on the slot's own height,
// so a DEM's totals do not depend on where it sits in an atlas.
#define SLOT_KB 1024
__host__ __device__ inline int slot_kb(int h)
{
    int kb = (h + 3) / 4;
    return kb < 1 ? 1 : (kb > SLOT_KB ? SLOT_KB : kb);
}

__device__ __forceinline__ int slot_of(const uint8_t *demid, long i) { return demid ? (int)demid[i] : 0; }

// per-step scalars kept on the device
struct LemScalars {
    double lowest;        // min over mask cells of dem, starting at 1000 (routeflow.cu:171-184)
    double base_level;    // dem[outlet] - 0.00012 (updates.cu:151)
    double FA_max;
    double totE, totI, totG, totD, totwC, totwP, totbio;
    int outlet;
    int sinks_pass1;      // mask cells with SFD == 0 after pass-1 routing
    int sinkcount;        // distinct labels in the edge set (newflow.cu:1045) or sinks_pass1
    int nlabels;
    int nedges;
    int plateau_rounds1, plateau_rounds2;
    int fa_uncomputed, sed_uncomputed;
    int ncand;            // plateau candidates (mask cells without a downhill neighbour)
    int flag;             // generic change flag for host-driven loops
    int hash_overflow;
    int w_bad;            // a mask cell's runoff weight is negative or not finite
    unsigned long long lowest_enc;
};
