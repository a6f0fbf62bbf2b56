// lem_sed32.hip — sediment routing of the FP32 form under single flow
// directions (MFDsedbud, mfd_accum.cu:63-252, loop mfd_accum.cu:345-483):
// the dependency engine's first rounds (the first over every tile), then a
// contraction of what is left of the donor forest in synchronous rounds,
// instead of the engine's further tile rounds.  (Under multiple flow
// directions most cells have several open donors at once, so few chains
// form and the rounds approach the depth of the graph; those contexts keep
// the engine.)
//
// Of a computed cell's results only the stone fraction travels downstream:
// deposition comes from the donors' erosion, incision and gelifluction
// (inputs of the pass, summed by the prep kernel), and the new soil
// thickness from the cell's own terms.  The new stone value is
//     y = clamp(a + b * S, 0.01, 99.9),   S = sum of the donors' new values,
// with a = old * stone and b = input (mfd_accum.cu:222-235); a cell with
// zero sediment thickness keeps its stone (y = stone, a constant).  Which
// cells compute at all is fixed by the graph alone: a cell computes once all
// its donors have, never when a donor is dead (row 0 / column 0 of its DEM
// or a nodata neighbour in a blocking direction, mfd_accum.cu:42,390-394),
// and never on or below a cycle.
//
// The engine's first round visits every 32 x 32 tile once and completes the
// cells whose whole upstream lies inside their tile (about 80% of the cells
// on the fBm workloads), each later round the cells whose upstream enters
// their tile only from cells completed before (three rounds by default);
// further tile rounds would revisit most tiles for the rest, and a river
// crossing hundreds of tiles needs as many rounds.  The rest contracts
// instead:
// functions x -> clamp(A + B x, L, U) are closed under composition, so the
// graph contracts in rounds.  Each cell's state after round r is a function
// of the states after round r-1 only (status words carry the round of their
// last change and the kind before it; unary records are double-buffered by
// round parity), so results do not depend on scheduling or list order.  In
// round r a cell
//   * whose donors have all computed computes, with the engine's arithmetic
//     (donor values summed in pull order in float): rake;
//   * with exactly one donor u still open becomes unary: y = F(y_u), F
//     folding in the computed donors' sum;
//   * that is unary with a target t composes t's function into its own and
//     points at t's target when t is unary too (pointer jumping: chains
//     halve every round), computes when t has computed, dies when t has died;
//   * with a dead donor dies.
// A pointer that has passed more cells than its DEM holds went round a
// cycle: the cell dies (never computes).  Rounds end when no cell is open.
//
// Values of the engine's rounds and of rake are bit-identical with the
// engine's.  A composed function is evaluated in double from the float
// parameters, so values reached through one differ from the engine's
// link-by-link float rounding by a few float ulps.  Which cells the first
// round completes depends on their position modulo the tile (32 cells), so
// a DEM gives identical results wherever it sits in an atlas only at
// positions that are multiples of 32.
#include <algorithm>
#include "lem_ctx.h"

// status word: kind (2 bits) | kind before the last change (2) | update
// flag (1) | round of the last change (27)
enum : uint32_t { CS_PEND = 0, CS_DONE = 1, CS_DEAD = 2, CS_UN = 3 };
#define CS_UPD 16u

__device__ __forceinline__ uint32_t cs_kind_at(uint32_t w, uint32_t round)
{
    return (w >> 5) <= round ? (w & 3u) : ((w >> 2) & 3u);
}

__device__ __forceinline__ uint32_t cs_word(uint32_t kind, uint32_t prev, uint32_t upd, uint32_t round)
{
    return kind | (prev << 2) | upd | (round << 5);
}

// y = clamp(A + B x, L, U); a constant c is A = c, B = 0, L = U = c
struct CsRec {
    double A, B;
    float L, U;
    int ptr, hop;
};
static_assert(sizeof(CsRec) == 32, "record layout");

struct CsCtl {
    int cnt[3];
    int pad0[29];
    int bad;
    int pad1[31];
    int unc;
    int pad2[31];
};

__device__ __forceinline__ float cs_eval(const CsRec &f, float x)
{
    double v = f.A + f.B * (double)x;
    if (v < (double)f.L) v = f.L;
    else if (v > (double)f.U) v = f.U;
    return (float)v;
}

// f after g: clamp(f.A + f.B * clamp(g.A + g.B x, g.L, g.U), f.L, f.U)
__device__ __forceinline__ CsRec cs_compose(const CsRec &f, const CsRec &g)
{
    CsRec h;
    h.A = f.A + f.B * g.A;
    h.B = f.B * g.B;
    // the image of g's range under f's affine part (either order of the ends)
    double l = f.A + f.B * (double)g.L, u = f.A + f.B * (double)g.U;
    if (l > u) { const double t = l; l = u; u = t; }
    if (u < (double)f.L) { h.A = f.L; h.B = 0.0; h.L = f.L; h.U = f.L; }
    else if (l > (double)f.U) { h.A = f.U; h.B = 0.0; h.L = f.U; h.U = f.U; }
    else {
        h.L = (float)(l > (double)f.L ? l : (double)f.L);
        h.U = (float)(u < (double)f.U ? u : (double)f.U);
    }
    h.ptr = g.ptr;
    const long hp = (long)f.hop + g.hop;
    h.hop = hp > 0x7FFFFFFF ? 0x7FFFFFFF : (int)hp;
    return h;
}

// Open cells are kept in lists built in chunks of 256 * IT entries per
// workgroup: one global atomic reserves a chunk's output range (workgroups
// adding to one counter serialise, so one atomic per wave would dominate).
// The pass over every cell takes 16 rows per workgroup; a round takes 2, so
// that more workgroups share the gathers of a short list (a thread works
// through its rows one after another).
#define CS_ITEMS 16
#ifndef CS_ROUND_ITEMS
#define CS_ROUND_ITEMS 2
#endif

template <int IT>
struct CsChunk {
    int cell[IT * 256];
    unsigned long long bal[IT][4];
    int off[IT][4];
    int base;
};

// row j of the chunk: record this thread's entry and the wave's keep mask
template <int IT>
__device__ __forceinline__ void cs_chunk_put(CsChunk<IT> &C, int j, bool keep, int cell)
{
    const unsigned long long b = __ballot(keep);
    if ((threadIdx.x & 63) == 0) C.bal[j][threadIdx.x >> 6] = b;
    C.cell[j * 256 + threadIdx.x] = cell;
}

// reserve the chunk's range with one atomic, then write the kept entries in
// row order (to a second list as well when copy is given)
template <int IT>
__device__ __forceinline__ void cs_chunk_flush(CsChunk<IT> &C, int *list, int *count, int *copy = nullptr)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int j = 0; j < IT; j++)
            for (int w = 0; w < 4; w++) {
                C.off[j][w] = tot;
                tot += __popcll(C.bal[j][w]);
            }
        C.base = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < IT; j++) {
        const unsigned long long b = C.bal[j][wave];
        if ((b >> lane) & 1ull) {
            const int at = C.base + C.off[j][wave] + __popcll(b & ((1ull << lane) - 1ull));
            list[at] = C.cell[j * 256 + threadIdx.x];
            if (copy) copy[at] = C.cell[j * 256 + threadIdx.x];
        }
    }
    __syncthreads();
}

// Status words after the engine's first rounds (lem_sed_first): a cell they
// completed is done and one it killed (or the prep kernel marked dead) is
// dead, both with stamp 0; a pending cell gets its rake parameters, as the
// engine's cell body forms them from its unchanged inputs, and is open.
// The open cells form round 2's list (and its copy for k_cs_finish); the
// dead cells are counted as uncomputed.  Each thread rewrites only its own
// cells' words.
__global__ void __launch_bounds__(256) k_cs_from_engine(long N, const uint8_t *__restrict__ mask, int *st,
                                                        const double *__restrict__ soilT,
                                                        const float *__restrict__ dps,
                                                        const float *__restrict__ stone, float *ca, float *cb,
                                                        int *list, int *list0, CsCtl *ctl)
{
    __shared__ CsChunk<CS_ITEMS> C;
    const long CH = 256L * CS_ITEMS;
    int unc = 0;
    for (long c0 = (long)blockIdx.x * CH; c0 < N; c0 += (long)gridDim.x * CH) {
#pragma unroll 1
        for (int j = 0; j < CS_ITEMS; j++) {
            const long i = c0 + j * 256 + threadIdx.x;
            bool keep = false;
            if (i < N) {
                const uint8_t m = mask[i];
                const int s = st[i];
                uint32_t w = cs_word(CS_DONE, CS_DONE, 0, 0);   // done, or outside the mask (never a donor)
                if (m && s == DAG_PENDING) {
                    const double sT = soilT[i];
                    const float sv = stone[i];
                    const double sedthick = sT + (double)dps[i];
                    float a = sv, b = 0.f;
                    uint32_t upd = 0;
                    if (sedthick != 0.0) {
                        const double old = sT / sedthick;
                        a = (float)old * sv;
                        b = (float)(1 - old);
                        upd = CS_UPD;
                    }
                    ca[i] = a;
                    cb[i] = b;
                    w = cs_word(CS_PEND, CS_PEND, upd, 0);
                    keep = true;
                } else if (m && s < 0) {
                    w = cs_word(CS_DEAD, CS_DEAD, 0, 0);
                    unc++;
                }
                st[i] = (int)w;
            }
            cs_chunk_put(C, j, keep, (int)i);
        }
        cs_chunk_flush(C, list, &ctl->cnt[2], list0);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) unc += __shfl_down(unc, o);
    if ((threadIdx.x & 63) == 0 && unc) atomicAdd(&ctl->unc, unc);
}

// One synchronous round r >= 2 over the open cells (list of round r:
// lists[r & 1], count cnt[r % 3]; the cells still open go to the other list
// and cnt[(r + 1) % 3]; cnt[(r + 2) % 3], last read by round r - 1, is
// cleared).
__global__ void __launch_bounds__(256) k_cs_round(int W, uint32_t r, const int *__restrict__ lin, int *lout,
                                                  CsCtl *ctl, uint32_t *stw, const uint8_t *__restrict__ dm,
                                                  float *stone, const float *__restrict__ ca,
                                                  const float *__restrict__ cb, const CsRec *rin, CsRec *rout,
                                                  const uint8_t *__restrict__ demid, const DemDesc *__restrict__ dd)
{
    __shared__ CsChunk<CS_ROUND_ITEMS> C;
    const int n = ctl->cnt[r % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->cnt[(r + 2) % 3] = 0;
    const long CH = 256L * CS_ROUND_ITEMS;
    const uint32_t prev_round = r - 1;
    for (long c0 = (long)blockIdx.x * CH; c0 < n; c0 += (long)gridDim.x * CH) {
#pragma unroll 1
        for (int j = 0; j < CS_ROUND_ITEMS; j++) {
            const long k = c0 + j * 256 + threadIdx.x;
            bool keep = false;
            int i = 0;
            if (k < n) {
                i = lin[k];
                const uint32_t w = stw[i];
                const uint32_t kind = w & 3u, upd = w & CS_UPD;
                uint32_t nk = kind;
                if (kind == CS_PEND) {
                    const unsigned bits = dm[i];
                    float S = 0.f;
                    int cnt = 0, u = -1;
                    bool dead = false;
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        if (!((bits >> q) & 1u)) continue;
                        const int d = cpull(q);
                        const int nb = i + cdy(d) * W + cdx(d);
                        const uint32_t kn = cs_kind_at(stw[nb], prev_round);
                        if (kn == CS_DEAD) dead = true;
                        else if (kn == CS_DONE) S = S + 1.0f * stone[nb];
                        else { cnt++; u = nb; }
                    }
                    const float a = ca[i], b = cb[i];
                    if (dead) nk = CS_DEAD;
                    else if (cnt == 0) {
                        float y = a;
                        if (upd) {
                            y = a + b * S;
                            if (y < 0.01f) y = 0.01f;
                            else if (y > 99.9f) y = 99.9f;
                        }
                        stone[i] = y;
                        nk = CS_DONE;
                    } else if (cnt == 1) {
                        CsRec f;
                        if (upd) { f.A = (double)a + (double)b * (double)S; f.B = b; f.L = 0.01f; f.U = 99.9f; }
                        else { f.A = a; f.B = 0.0; f.L = a; f.U = a; }
                        f.ptr = u;
                        f.hop = 1;
                        rout[i] = f;
                        nk = CS_UN;
                        keep = true;
                    } else {
                        keep = true;
                    }
                } else {   // unary
                    CsRec f = rin[i];
                    const uint32_t kt = cs_kind_at(stw[f.ptr], prev_round);
                    if (kt == CS_DONE) {
                        stone[i] = cs_eval(f, stone[f.ptr]);
                        nk = CS_DONE;
                    } else if (kt == CS_DEAD) {
                        nk = CS_DEAD;
                    } else {
                        if (kt == CS_UN) {
                            f = cs_compose(f, rin[f.ptr]);
                            const DemDesc D = dd[slot_of(demid, i)];   // single-DEM contexts: no slot ids
                            if ((long)f.hop > (long)D.w * D.h) nk = CS_DEAD;   // went round a cycle
                        }
                        if (nk == CS_UN) {
                            rout[i] = f;
                            keep = true;
                        }
                    }
                }
                if (nk != kind) stw[i] = cs_word(nk, kind, upd, r);
            }
            cs_chunk_put(C, j, keep, i);
        }
        cs_chunk_flush(C, lout, &ctl->cnt[(r + 1) % 3]);
    }
}

// Results of the cells the rounds computed (the open cells of round 2's
// list), as the engine's cell body writes them (deposition, fines, soil
// thickness; stone is already in place); the engine's rounds have written
// their own cells; an open cell that never computed is left as it was and
// counted.
__global__ void __launch_bounds__(256) k_cs_finish(int n, const int *__restrict__ list0,
                                                   const uint32_t *__restrict__ stw, const float *__restrict__ dps,
                                                   const float *__restrict__ ero, const float *__restrict__ inci,
                                                   const float *__restrict__ geli, const float *__restrict__ stone,
                                                   float *depo, float *fines, double *soilT, CsCtl *ctl)
{
    int unc = 0;
    const int stride = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const long i = list0[k];
        const uint32_t w = stw[i];
        if ((w & 3u) != CS_DONE) { unc++; continue; }
        const float dp = dps[i];
        const double sT = soilT[i];
        const double sedthick = sT + (double)dp;
        const double toterosion = (double)(ero[i] + inci[i] + geli[i]);
        depo[i] = dp;
        if (w & CS_UPD) fines[i] = 100.f - stone[i];
        double sTn = sT;
        if (toterosion >= sedthick) sTn = 0.0;
        if (sedthick > toterosion) sTn = sedthick - toterosion;
        soilT[i] = sTn;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) unc += __shfl_down(unc, o);
    if ((threadIdx.x & 63) == 0 && unc) atomicAdd(&ctl->unc, unc);
}

template <typename T> int lem_sed_prep(lemhip_ctx *c, int *bad);

int lem_sed32(lemhip_ctx *c, int *fallback)
{
    *fallback = 1;
    // an atlas steps only the band of rows its DEMs use (N <= NP, the band
    // can grow between steps): the records are sized for the whole grid
    const long N = c->N, NP = c->NP;
    if (!c->cs_rec) {
        hipError_t e = hipMalloc((void **)&c->cs_rec, 2 * (size_t)NP * sizeof(CsRec));
        if (e != hipSuccess) {
            c->cs_rec = nullptr;
            c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
            return LEMHIP_E_NOMEM;
        }
    }
    CsRec *rec[2] = {reinterpret_cast<CsRec *>(c->cs_rec), reinterpret_cast<CsRec *>(c->cs_rec) + NP};
    int *lists[2] = {c->cand, c->ws};
    float *ca = c->sp0, *cb = c->sp1;
    uint32_t *stw = reinterpret_cast<uint32_t *>(c->scan_tmp);
    const float *dps = reinterpret_cast<const float *>(c->z2);
    CsCtl *ctl = reinterpret_cast<CsCtl *>(c->dag_ctl);
    static_assert(sizeof(CsCtl) <= 512, "control block");

    LEM_CHECK(c, hipMemsetAsync(ctl, 0, sizeof(CsCtl), c->s));
    LEM_TRY(lem_sed_prep<float>(c, &ctl->bad));
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[25], &ctl->bad, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    // non-finite parameters: the engine's path handles them (nothing of the
    // state has changed yet)
    if (c->hint[25]) return LEMHIP_OK;
    *fallback = 0;
    // the engine's first rounds, the first over every tile (their control block
    // shares the memory of this one)
    LEM_TRY(lem_sed_first(c, c->cs_rounds));
    const long CH = 256L * CS_ITEMS;
    auto chunks = [&](long n) { return (int)std::max(1L, std::min((n + CH - 1) / CH, 4096L)); };
    const long CHR = 256L * CS_ROUND_ITEMS;
    auto chunks_r = [&](long n) { return (int)std::max(1L, std::min((n + CHR - 1) / CHR, 16384L)); };
    LEM_CHECK(c, hipMemsetAsync(ctl, 0, sizeof(CsCtl), c->s));
    // status words and round 2's list (lists[0], count in cnt[2]); its copy
    // for k_cs_finish in the upper half of the raised-surface buffer (dps
    // holds the lower half as floats)
    int *list0 = reinterpret_cast<int *>(c->z2) + NP;
    k_cs_from_engine<<<chunks(N), 256, 0, c->s>>>(N, c->mask, c->scan_tmp, c->soilT, dps, (const float *)c->stone,
                                                  ca, cb, lists[0], list0, ctl);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[24], &ctl->cnt[2], sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    long open = c->hint[24];
    const long open0 = open;
    c->h_sed_open = (int)open;
    uint32_t r = 1;
    long launches = 1;
    const int BATCH = 8;   // rounds between host reads of the open count
    while (open > 0) {
        prof_begin(c, PROF_CS_K);
        for (int q = 0; q < BATCH; q++) {
            ++r;
            k_cs_round<<<chunks_r(open), 256, 0, c->s>>>(c->W, r, lists[r & 1], lists[(r + 1) & 1], ctl, stw, c->dmask,
                                                       (float *)c->stone, ca, cb, rec[(r - 1) & 1], rec[r & 1],
                                                       c->demid, c->ddesc);
        }
        prof_end(c, PROF_CS_K, BATCH);
        LEM_LAUNCHED(c);
        launches += BATCH;
        LEM_CHECK(c, hipMemcpyAsync(&c->hint[24], &ctl->cnt[(r + 1) % 3], sizeof(int), hipMemcpyDeviceToHost,
                                    c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        open = c->hint[24];
        if (r > (1u << 20)) { c->err = "sediment contraction did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    k_cs_finish<<<lem_grid1d(std::max(open0, 1L)), 256, 0, c->s>>>((int)open0, list0, stw, dps, (const float *)c->ero,
                                                  (const float *)c->inci, (const float *)c->geli,
                                                  (const float *)c->stone, (float *)c->depo, (float *)c->fines,
                                                  c->soilT, ctl);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[26], &ctl->unc, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    c->h_sed_unc = c->hint[26];
    c->dag_rounds_last = launches;
    return LEMHIP_OK;
}
