// lem_fd.hip — flow direction phase (cuFlowDirection, routeflow.cu:686-975):
// D8/MFD directions, plateau routing, depression (sink) routing by watershed
// labels and minimax spill heights, the second direction pass on the raised
// surface, and the fd/SFD merge.
#include "lem_ctx.h"
#include <hipcub/hipcub.hpp>
#include <stddef.h>

// ------------------------------------------------------------------ helpers

// one atomic per wavefront: returns this lane's slot, or -1 if !pred
__device__ __forceinline__ int wave_append(int *counter, bool pred)
{
    unsigned long long b = __ballot(pred);
    if (b == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)b) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(b));
    base = __shfl(base, leader);
    return pred ? base + __popcll(b & ((1ULL << lane) - 1ULL)) : -1;
}

// one global atomic per workgroup (block size a multiple of 64, at most 1024);
// every thread of the block must call it
__device__ __forceinline__ int block_append(int *counter, bool pred, int *s_wave, int *s_base)
{
    unsigned long long b = __ballot(pred);
    int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int lane = tid & 63;
    int wave = tid >> 6;
    int nwaves = (blockDim.x * blockDim.y) >> 6;
    if (lane == 0) s_wave[wave] = __popcll(b);
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < nwaves; w++) {
            int cnt = s_wave[w];
            s_wave[w] = tot;
            tot += cnt;
        }
        *s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int pos = *s_base + s_wave[wave] + __popcll(b & ((1ULL << lane) - 1ULL));
    __syncthreads();
    return pred ? pos : -1;
}

// n_mine items per thread appended to a global list with one atomic per block;
// returns this thread's first index
__device__ __forceinline__ int block_append_n(int *counter, int n_mine, int *s_wave, int *s_base)
{
    int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int lane = tid & 63;
    int wave = tid >> 6;
    int nwaves = (blockDim.x * blockDim.y) >> 6;
    int x = n_mine;   // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < nwaves; w++) {
            int cnt = s_wave[w];
            s_wave[w] = tot;
            tot += cnt;
        }
        *s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int pos = *s_base + s_wave[wave] + x - n_mine;
    __syncthreads();
    return pos;
}

// Stream compaction in chunks of 256 * ITEMS elements (1-D blocks of 256):
// one global atomic per chunk instead of one per 256 elements, since many
// workgroups adding to one counter serialise.  Output order is a permutation
// within each chunk.
#define COMPACT_ITEMS 16
template <int ITEMS, class Pred, class Emit>
__device__ __forceinline__ void compact_chunks(long n, int *count, int *s_wave, int *s_base, Pred pred, Emit emit)
{
    const long CH = 256L * ITEMS;
    for (long c0 = blockIdx.x * CH; c0 < n; c0 += (long)gridDim.x * CH) {
        unsigned keep = 0;
        int mine = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            long i = c0 + j * 256 + threadIdx.x;
            bool k = i < n && pred(i);
            keep |= (unsigned)k << j;
            mine += k;
        }
        int pos = block_append_n(count, mine, s_wave, s_base);
#pragma unroll
        for (int j = 0; j < ITEMS; j++)
            if ((keep >> j) & 1u) emit(c0 + j * 256 + threadIdx.x, pos++);
    }
}

// The same keeping the input order: the chunk is read row by row (256
// consecutive elements, coalesced); one atomic reserves the chunk's output
// range, then each row's kept elements take their ranks from a block-wide
// scan of that row.
template <int ITEMS, class Pred, class Emit>
__device__ __forceinline__ void compact_chunks_ordered(long n, int *count, int *s_wave, int *s_base, Pred pred,
                                                       Emit emit)
{
    __shared__ int s_row[ITEMS][4];
    const long CH = 256L * ITEMS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long c0 = blockIdx.x * CH; c0 < n; c0 += (long)gridDim.x * CH) {
        unsigned keep = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            long i = c0 + j * 256 + threadIdx.x;
            bool k = i < n && pred(i);
            keep |= (unsigned)k << j;
            unsigned long long b = __ballot(k);
            if (lane == 0) s_row[j][wave] = __popcll(b);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int j = 0; j < ITEMS; j++)
                for (int w = 0; w < 4; w++) {
                    int c = s_row[j][w];
                    s_row[j][w] = tot;
                    tot += c;
                }
            *s_base = tot ? atomicAdd(count, tot) : 0;
        }
        __syncthreads();
        const int base = *s_base;
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            bool k = (keep >> j) & 1u;
            unsigned long long b = __ballot(k);
            if (k) emit(c0 + j * 256 + threadIdx.x, base + s_row[j][wave] + __popcll(b & ((1ULL << lane) - 1ULL)));
        }
        __syncthreads();
    }
}

static inline int grid_chunks(long n, int cap = 8192)
{
    long g = (n + 256L * COMPACT_ITEMS - 1) / (256L * COMPACT_ITEMS);
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T *sh)
{
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// ----------------------------------------------------------------- lowest

// routeflow.cu:739-741: the outlet is lowered below the lowest cell after the
// directions were computed on the unadjusted surface (one thread per slot)
__global__ void k_slots_outlet_adjust(double *dem, const DemDesc *dd, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    double lowest = ord_dec(ds[j].lowest_enc);
    ds[j].lowest = lowest;
    if (lowest < dem[ds[j].outlet]) dem[ds[j].outlet] = lowest - 0.0001;
}

// routeflow.cu:953: fd[outlet] = WEST
__global__ void k_slots_fd_outlet(uint8_t *fd, const DemDesc *dd, int ns, const DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    fd[ds[j].outlet] = 16;
}

// per-slot counter adds from a block: LDS histogram over slots, then one atomic per slot touched
__device__ __forceinline__ void slot_hist_flush(int *hist, DemScal *ds, int off, int ns)
{
    __syncthreads();
    for (int j = threadIdx.x; j < ns; j += blockDim.x)
        if (hist[j]) atomicAdd(reinterpret_cast<int *>(reinterpret_cast<char *>(&ds[j]) + off), hist[j]);
}

// ---------------------------------------------------------------- FlowDirs

// newflow.cu:49-174.  One thread per cell; a wave covers 64 consecutive cells
// of a row so the centre row and both neighbour rows are read coalesced.
// Slopes take new values only for downhill neighbours (sticky: Slopes keep the first pass's values); pass 2
// passes slopes == nullptr.  Cells without a downhill neighbour are appended
// to the plateau candidate list (the cells whose initial distance is 2e7).
// A 64 x 4 block covers 64 x 64 cells in four 64 x 16 tiles: each tile's
// surface plus a one-cell ring is staged in LDS with every load in flight at
// once, then each thread evaluates four rows from LDS.  The block appends its
// candidates with one atomic at the end (appends from many blocks to one
// counter serialise).  Off-grid ring cells hold NaN: every comparison with it
// is false, so such a neighbour is neither downhill, up nor flat, exactly as
// a skipped one.
template <typename TS>
__global__ void __launch_bounds__(256) k_flowdirs(int W, int H, int cell_size, const double *__restrict__ z,
                                                  const uint8_t *__restrict__ mask, TS *__restrict__ slopes,
                                                  long plane, uint8_t *__restrict__ fd, uint8_t *__restrict__ SFD,
                                                  uint8_t *__restrict__ eqb, int *__restrict__ cand,
                                                  int *__restrict__ ncand)
{
    constexpr int RW = 66, RH = 18, RN = RW * RH, PER = (RN + 255) / 256;
    __shared__ double zs[RN];
    __shared__ int s_wave[4];
    __shared__ int s_base;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int x0 = blockIdx.x * 64;
    const int icol = x0 + threadIdx.x;
    unsigned cand_rows = 0;   // bit 4 * sub + rg: the cell of row group rg of tile sub is a plateau candidate
    for (int sub = 0; sub < 4; sub++) {
        const int y0 = blockIdx.y * 64 + sub * 16;
        if (y0 >= H) break;
        double v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int k = tid + q * 256;
            int gy = y0 - 1 + k / RW, gx = x0 - 1 + k % RW;
            v[q] = lm_nan();
            if (k < RN && gy >= 0 && gy < H && gx >= 0 && gx < W) v[q] = z[(long)gy * W + gx];
        }
        uint32_t mk[4];   // 32-bit: byte registers would be packed, with a wait per load
#pragma unroll
        for (int rg = 0; rg < 4; rg++) {
            int irow = y0 + rg * 4 + threadIdx.y, icol = x0 + threadIdx.x;
            mk[rg] = (icol < W && irow < H) ? mask[(long)irow * W + icol] : 0;
        }
        __syncthreads();   // the previous tile's reads are done
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int k = tid + q * 256;
            if (k < RN) zs[k] = v[q];
        }
        __syncthreads();
#pragma unroll 1
        for (int rg = 0; rg < 4; rg++) {
            int lr = rg * 4 + threadIdx.y;
            int irow = y0 + lr;
            bool inside = icol < W && irow < H;
            long self = (long)irow * W + icol;
            bool active = inside && mk[rg];
            int aspect = 0, aspectsfd = 0;
            unsigned eq = 0;
            if (active) {
                // Every Slopes plane of the cell is rewritten, keeping the old value
                // where the neighbour is not downhill: whole cache lines are written
                // instead of lines with holes in them.  The old values are loaded
                // first and merged after the scan.
                TS sv[8];
                if (slopes) {
#pragma unroll
                    for (int d = 0; d < 8; d++) sv[d] = slopes[d * plane + self];
                }
                const int c0 = (lr + 1) * RW + (threadIdx.x + 1);
                double zc = zs[c0];
                double smax = 0.0;
                int up = 0, flat = 0;
                unsigned down = 0;
                TS nv[8];
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    float dc = (d & 1) ? 1.41f : 1.0f;
                    float den = (float)cell_size * dc;
                    double zn = zs[c0 + kDY[d] * RW + kDX[d]];
                    double stemp = (zc - zn) / (double)den;
                    nv[d] = (TS)stemp;
                    if ((zc > zn) && (zn != -9999)) {
                        down |= 1u << d;
                        if (stemp > smax) { smax = stemp; aspectsfd = kCODE[d]; }
                        aspect += kCODE[d];
                    }
                    if (zc < zn) up++;
                    if (zc == zn) { flat++; eq |= 1u << d; }
                }
                if (slopes) {
#pragma unroll
                    for (int d = 0; d < 8; d++) slopes[d * plane + self] = ((down >> d) & 1u) ? nv[d] : sv[d];
                }
                if (up > 7) { aspect = 0; aspectsfd = 0; }
                if ((flat + up) > 7 && up <= 7) { aspect = 0; aspectsfd = 0; }
                if ((flat < 8) && (flat + up > 7)) { aspect = 0; aspectsfd = 0; }
            }
            if (inside) { fd[self] = (uint8_t)aspect; SFD[self] = (uint8_t)aspectsfd; eqb[self] = (uint8_t)eq; }
            if (active && aspect == 0 && aspectsfd == 0) cand_rows |= 1u << (sub * 4 + rg);
        }
    }
    // one append per block for its candidates (their order is not needed)
    int mine = __popc(cand_rows);
    int base = block_append_n(ncand, mine, s_wave, &s_base);
    while (cand_rows) {
        const int b = __ffs(cand_rows) - 1;
        cand_rows &= cand_rows - 1;
        cand[base++] = (int)((long)(blockIdx.y * 64 + (b >> 2) * 16 + (b & 3) * 4 + threadIdx.y) * W + icol);
    }
}

// ------------------------------------------------------------ plateaus
//
// route_plateausMFD (newflow.cu:264-350) in synchronous rounds: in round
// r a candidate reads its equal-height neighbours' round r-1 distances; each
// strict improvement in the scan (order N..NW) marks that slope and takes its
// direction.  The rounds are temporally blocked: one launch advances K rounds
// for a tile of PTW x PTH cells, loading the tile plus a K-cell halo into LDS
// and recomputing the halo redundantly; after round k only cells at least k
// cells from the halo's outer edge are current, so after K rounds the tile
// itself is exact.  A cell can change in round r only if an equal-height
// neighbour changed in round r-1, so a tile needs the next launch only when a
// cell within K cells of it changed in the last round of this launch (K <=
// tile width and height: in it or in one of its 8 neighbour tiles); tiles
// without candidates never change.

__global__ void k_sp_init(long N, float *sp)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) sp[i] = 0.0f;
}

// candidates start at 2e7 (newflow.cu:252); non-candidates keep 0 and never change
__global__ void k_sp_seed(int W, const int *cand, int n, float *sp, int ntx, int *tile_has)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = cand[k];
    sp[i] = 20000000.0f;
    int t = ((i / W) / PTH) * ntx + (i % W) / PTW;
    if (tile_has[t] == 0) tile_has[t] = 1;
}

// A candidate with an equal-height neighbour at distance 0 (not a
// candidate, or off the grid, where the plateau rounds read 0) takes its
// distance in the first round; every other cell changes in round r only if
// an equal-height neighbour changed in round r-1, i.e. it lies within r-1
// cells of such a seed.  The first launch's K rounds (K <= tile height) can
// therefore change only tiles in the 3x3 tile neighbourhood of a seed's
// tile: those tiles (with candidates) are marked 3, the first list.
__global__ void k_sp_seedmark(int W, int H, const int *cand, int n, const float *sp, const uint8_t *eqb, int ntx,
                              int nty, int *tile_has)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = cand[k];
    const int y = i / W, x = i % W;
    const unsigned e = eqb[i];
    bool seed = false;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const int yy = y + cdy(d), xx = x + cdx(d);
        if ((e >> d) & 1u) seed |= yy < 0 || yy >= H || xx < 0 || xx >= W || sp[(long)yy * W + xx] == 0.0f;
    }
    if (!seed) return;
    const int ty = y / PTH, tx = x / PTW;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int t2y = ty + dy, t2x = tx + dx;
            if (t2y < 0 || t2y >= nty || t2x < 0 || t2x >= ntx) continue;
            const int t2 = t2y * ntx + t2x;
            if (tile_has[t2] != 0) tile_has[t2] = 3;
        }
}

// Launch bookkeeping: launch L visits list L % 3 (count cnt[L % 3]); a tile
// still moving at its end queues itself and its neighbours with candidates in
// list (L+1) % 3, each once (stamp[tile] holds the last launch it was queued
// for), and launch L clears cnt[(L+2) % 3], last read by the publish step of
// launch L-1.
struct PlatCtl {
    int cnt[3];
    int pad0[29];
    int last_change;   // last round (absolute number) in which any cell changed
    int pad1[31];
};

__global__ void k_plat_first_list(int ntiles, const int *tile_has, int *list, PlatCtl *ctl, int slot)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    for (int t0 = blockIdx.x * blockDim.x; t0 < ntiles; t0 += gridDim.x * blockDim.x) {
        int t = t0 + threadIdx.x;
        bool have = t < ntiles && tile_has[t] == 3;
        int at = block_append(&ctl->cnt[slot], have, s_wave, &s_base);
        if (have) list[at] = t;
    }
}

// threads per plateau workgroup: a launch near the end of the pass visits a
// few tiles, each in K synchronised rounds over its region, so the rounds
// are spread over many threads
#define PLAT_NT 512

// Distances live in two buffers per tile: tver[tile] = 2 * (launch that last
// wrote the tile) + (buffer it wrote).  A launch reads every tile from the
// buffer holding its state as of the previous launch and a visited tile
// writes the other one, so no launch reads what another visit of the same
// launch writes, and no copy-back is needed.
__device__ __forceinline__ int plat_buf_as_of(int tv, int launch)
{
    int b = tv & 1;
    return (tv >> 1) == launch ? 1 - b : b;
}

__global__ void k_plat_tver_init(int ntiles, int *tver)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles) tver[t] = -2;   // buffer 0, written before any launch
}

// Appends a workgroup's queued tiles to the next launch's list with one
// atomic (called by every thread of the workgroup, after a barrier).
#define PLAT_QCAP 1024
__device__ __forceinline__ void plat_flush(int *qbuf, int *qn, int *next, int *count)
{
    __shared__ int base;
    const int cnt = *qn;
    if (threadIdx.x == 0) base = atomicAdd(count, cnt);
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) next[base + k] = qbuf[k];
    __syncthreads();
    if (threadIdx.x == 0) *qn = 0;
    __syncthreads();
}

// MARKS: the pass records slope marks (the first pass, on the dem; the pass
// on the raised surface keeps only directions and distances)
template <int K, typename TS, bool MARKS>
__global__ void __launch_bounds__(PLAT_NT) k_plat_block(int W, int H, int launch, int r0, const int *__restrict__ act,
                                                    int *__restrict__ next, float *__restrict__ sp0,
                                                    float *__restrict__ sp1, int *__restrict__ tver,
                                                    const uint8_t *__restrict__ eqb,
                                                    uint8_t *__restrict__ fd, uint8_t *__restrict__ SFD,
                                                    TS *__restrict__ slopes, long plane,
                                                    const int *__restrict__ tile_has, int *__restrict__ tstamp,
                                                    PlatCtl *ctl)
{
    constexpr int RW = PTW + 2 * K, RH = PTH + 2 * K, RN = RW * RH, NI = PTW * PTH;
    constexpr int PER = (RN + PLAT_NT - 1) / PLAT_NT, PERI = (NI + PLAT_NT - 1) / PLAT_NT;
    static_assert(PER <= 32, "region too large for the per-thread cell mask");
    __shared__ float A[RN], B[RN];
    __shared__ uint8_t E[RN];
    __shared__ uint8_t code[NI], marks[NI];
    __shared__ int last;
    __shared__ int s_qmask;   // 3x3 tiles within K cells of a cell that changed in round K
    __shared__ int s_buf[9];
    __shared__ int qbuf[PLAT_QCAP];   // tiles queued for the next launch, not yet appended
    __shared__ int qn;
    int wg_last = 0;                  // thread 0: the last round with a change in this workgroup's visits
    if (threadIdx.x == 0) qn = 0;
    const int n = ctl->cnt[launch % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->cnt[(launch + 2) % 3] = 0;
    const int ntx = (W + PTW - 1) / PTW, nty = (H + PTH - 1) / PTH;
    // the next visit's tile is read during this one (the first together with
    // the count: the grid never exceeds the list's capacity, and an entry
    // past the count is never used)
    int tile_next = act[blockIdx.x];
    for (int q0 = blockIdx.x; q0 < n; q0 += gridDim.x) {
        const int tile = tile_next;
        tile_next = q0 + (int)gridDim.x < n ? act[q0 + gridDim.x] : 0;
        const int tx0 = (tile % ntx) * PTW, ty0 = (tile / ntx) * PTH;
        __syncthreads();
        if (threadIdx.x == 0) { last = 0; s_qmask = 0; }
        // load the region from both distance buffers, and in the same wait
        // which buffer holds each of the 3x3 tiles as of the last launch: the
        // region's loads do not wait for that choice (a visit of the first
        // launches spends most of its time waiting on its loads)
        float v0[PER], v1[PER];
        uint32_t e[PER];   // 32-bit: byte registers would be packed, with a wait per load
        int part[PER];   // 3x3 tile of the region cell, -1 off the grid
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int idx = threadIdx.x + q * PLAT_NT;
            int gy = ty0 - K + idx / RW, gx = tx0 - K + idx % RW;
            v0[q] = 0.f;
            v1[q] = 0.f;
            e[q] = 0;
            part[q] = -1;
            if (idx < RN && gy >= 0 && gy < H && gx >= 0 && gx < W) {
                long i = (long)gy * W + gx;
                int ny = gy < ty0 ? 0 : (gy >= ty0 + PTH ? 2 : 1), nx = gx < tx0 ? 0 : (gx >= tx0 + PTW ? 2 : 1);
                part[q] = ny * 3 + nx;
                v0[q] = sp0[i];
                v1[q] = sp1[i];
                e[q] = eqb[i];
            }
        }
        if (threadIdx.x < 9) {
            int ty = tile / ntx + (int)threadIdx.x / 3 - 1, tx = tile % ntx + (int)threadIdx.x % 3 - 1;
            s_buf[threadIdx.x] = (ty >= 0 && ty < nty && tx >= 0 && tx < ntx) ? plat_buf_as_of(tver[ty * ntx + tx], launch) : 0;
        }
        __syncthreads();
        // both LDS buffers hold every cell, so cells that cannot change (no
        // candidate, or no equal-height neighbour) are never rewritten
        unsigned own = 0;
        float v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) v[q] = part[q] < 0 ? 0.f : (s_buf[part[q]] ? v1[q] : v0[q]);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int idx = threadIdx.x + q * PLAT_NT;
            if (idx < RN) {
                A[idx] = v[q];
                B[idx] = v[q];
                E[idx] = e[q];
                if (v[q] != 0.f && e[q] != 0) own |= 1u << q;
            }
        }
#pragma unroll
        for (int q = 0; q < PERI; q++) {
            int li = threadIdx.x + q * PLAT_NT;
            if (li < NI) { code[li] = 0; marks[li] = 0; }
        }
        __syncthreads();
        float *cur = A, *nxt = B;
        int moved = 0;   // last round in which one of this thread's tile cells changed
        unsigned qm = 0;  // 3x3 tiles within K cells of this thread's cells that changed in round K
        // A round in which no cell of its window changes leaves every later
        // round unchanged too (round k+1 reads only round k values of a window
        // that contains its own), so the launch stops there.
        for (int k = 1; k <= K; k++) {
            bool changed = false;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int idx = threadIdx.x + q * PLAT_NT;
                const int rr = idx / RW, cc = idx % RW;
                const bool act = ((own >> q) & 1) && rr >= k && rr < RH - k && cc >= k && cc < RW - k;
                if (!act) continue;
                // the scan N..NW without branches
                const int b = idx;
                float md = cur[b];
                const unsigned ee = E[b];   // read per round: keeps 8 lane masks per cell out of the registers
                unsigned cd = 0, mk = 0;
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    const float dc = (d & 1) ? 1.41f : 1.0f;
                    const float dis = cur[b + cdy(d) * RW + cdx(d)] + dc;
                    const bool take = ((ee >> d) & 1u) && (md > dis);
                    md = take ? dis : md;
                    cd = take ? (1u << ((d + 6) & 7)) : cd;   // fd code of direction d
                    if (MARKS) mk |= take ? (1u << d) : 0u;
                }
                nxt[idx] = md;
                changed |= cd != 0;
                if (cd && rr >= K && rr < K + PTH && cc >= K && cc < K + PTW) {
                    int li = (rr - K) * PTW + (cc - K);
                    code[li] = (uint8_t)cd;
                    if (MARKS) marks[li] |= (uint8_t)mk;
                    moved = k;
                    if (k == K) {
                        const int ly = rr - K, lx = cc - K;
                        const unsigned rows = 2u | (ly < K ? 1u : 0u) | (ly >= PTH - K ? 4u : 0u);
                        const unsigned cols = 2u | (lx < K ? 1u : 0u) | (lx >= PTW - K ? 4u : 0u);
#pragma unroll
                        for (int ny = 0; ny < 3; ny++)
                            if ((rows >> ny) & 1u) qm |= cols << (3 * ny);
                    }
                }
            }
            const int any = __syncthreads_or(changed);
            float *t = cur; cur = nxt; nxt = t;
            if (!any) break;
        }
        if (moved) atomicMax(&last, moved);
        if (qm) atomicOr(&s_qmask, (int)qm);
        __syncthreads();
        // still moving at the end of the launch: this tile and the neighbours
        // within K cells of a cell that changed in the last round go on (a
        // cell changing in round K + j lies within j cells of one that changed
        // in round K), gathered in LDS and appended to the next list with one
        // atomic per full buffer (workgroups adding to one counter serialise).
        // Before the write-back: the wait for the atomic's returned value
        // would otherwise also wait for those stores
        if (threadIdx.x < 9 && ((s_qmask >> threadIdx.x) & 1)) {
            int ty = tile / ntx + (int)threadIdx.x / 3 - 1, tx = tile % ntx + (int)threadIdx.x % 3 - 1;
            if (ty >= 0 && ty < nty && tx >= 0 && tx < ntx) {
                const int t2 = ty * ntx + tx;
                if (tile_has[t2] && atomicExch(&tstamp[t2], launch + 1) != launch + 1) qbuf[atomicAdd(&qn, 1)] = t2;
            }
        }
        // write back the tile: distances (into its other buffer), directions of
        // changed cells, slope marks (pass 1 only)
        float *sp_out = s_buf[4] ? sp0 : sp1;
#pragma unroll
        for (int q = 0; q < PERI; q++) {
            int li = threadIdx.x + q * PLAT_NT;
            int gy = ty0 + li / PTW, gx = tx0 + li % PTW;
            if (li >= NI || gy >= H || gx >= W) continue;
            long i = (long)gy * W + gx;
            sp_out[i] = cur[(li / PTW + K) * RW + (li % PTW + K)];
            uint8_t cd = code[li];
            if (cd) { fd[i] = cd; SFD[i] = cd; }
            if (MARKS && marks[li]) {
                unsigned mk = marks[li];
#pragma unroll
                for (int d = 0; d < 8; d++)
                    if ((mk >> d) & 1) slopes[d * plane + i] = (TS)0.0001;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            tver[tile] = (launch << 1) | (s_buf[4] ? 0 : 1);
            if (last > 0 && r0 + last > wg_last) wg_last = r0 + last;
        }
        __syncthreads();
        if (qn > PLAT_QCAP - 9) plat_flush(qbuf, &qn, next, &ctl->cnt[(launch + 1) % 3]);
    }
    __syncthreads();
    if (qn > 0) plat_flush(qbuf, &qn, next, &ctl->cnt[(launch + 1) % 3]);
    if (threadIdx.x == 0 && wg_last > 0) atomicMax(&ctl->last_change, wg_last);
}

__global__ void k_slots_sinks_init(int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) ds[j].sinks_pass1 = 0;
}

__global__ void __launch_bounds__(256) k_count_sinks(const int *cand, int n, const uint8_t *SFD,
                                                     const uint8_t *demid, int ns, DemScal *ds)
{
    __shared__ int hist[LEM_MAX_SLOTS + 1];
    for (int j = threadIdx.x; j < ns; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        long i = cand[k];
        if (SFD[i] == 0) atomicAdd(&hist[slot_of(demid, i)], 1);
    }
    slot_hist_flush(hist, ds, (int)offsetof(DemScal, sinks_pass1), ns);
}

// ------------------------------------------------ watershed labels

// DEM-local border of init_watershedMFD (newflow.cu:371): last column, first
// and last row of the cell's own DEM
// A cell of a DEM lies in its last column, first row or last row exactly when
// its east, north or south neighbour is outside the DEM's rectangle: off the
// grid, or in another DEM's rectangle or none (rectangles are convex).  The
// three slot ids are read at once instead of the DEM's rectangle after the
// cell's own id.
__device__ __forceinline__ bool dem_border(int W, int H, int r, int c, long i, const uint8_t *demid)
{
    if (demid) {
        const uint8_t me = demid[i];
        const uint8_t e = c + 1 < W ? demid[i + 1] : DEM_NONE;
        const uint8_t n = r > 0 ? demid[i - W] : DEM_NONE;
        const uint8_t s = r + 1 < H ? demid[i + W] : DEM_NONE;
        return e != me || n != me || s != me;
    }
    return (c == W - 1) || (r == 0) || (r == H - 1);
}

// sink flag: mask, not on the DEM border, SFD == 0 (newflow.cu:363-378)
// Row/column grids (grid2d): cell coordinates without a division per cell.
__global__ void k_sink_flags(int W, int H, const uint8_t *mask, const uint8_t *SFD, const uint8_t *demid,
                             const DemDesc *dd, int *flag)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W) return;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const long i = (long)r * W + c;
        const uint8_t m = mask[i], s = SFD[i];   // both loads in flight
        flag[i] = (m && s == 0 && !dem_border(W, H, r, c, i, demid)) ? 1 : 0;
    }
}

// init_watershedMFD with labels = 1 + rank of the sink cell in atlas row-major order
__global__ void k_init_ws(int W, int H, const uint8_t *mask, const uint8_t *SFD, const int *prefix,
                          const uint8_t *demid, const DemDesc *dd, int *ws, uint8_t *labdem)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W) return;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const long i = (long)r * W + c;
        const uint8_t m = mask[i], s = SFD[i];   // loads in flight together
        const int sl = slot_of(demid, i);
        int v;
        if (!m) v = 0;
        else if (dem_border(W, H, r, c, i, demid)) v = 0;
        else if (s == 0) {
            v = prefix[i] + 1;
            labdem[v] = (uint8_t)sl;
        } else {
            int d = code_dir(s);
            v = -((r + kDY[d]) * W + (c + kDX[d]) + 1);
        }
        ws[i] = v;
    }
}

// init_watershed_sinkMFD (newflow.cu:430-460): SFD==0 cells adopt the largest
// label of an equal-height neighbour; iterated to the fixed point (the
// component maximum, independent of the update order).  SFD==0 cells are a
// subset of the pass-1 plateau candidates.
__global__ void k_sink_merge(int W, int H, const int *cand, int n, const uint8_t *SFD, const double *z,
                             int *ws, int *flag)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    long self = cand[k];
    if (SFD[self] != 0) return;
    int irow = (int)(self / W), icol = (int)(self % W);
    double zc = z[self];
    int cur = ws[self], best = cur;
    for (int d = 0; d < 8; d++) {
        int cx = icol + kDX[d], cy = irow + kDY[d];
        if (cx >= 0 && cx < W && cy >= 0 && cy < H) {
            long nb = (long)cy * W + cx;
            if (z[nb] == zc) {
                int v = ws[nb];
                if (v > best) best = v;
            }
        }
    }
    if (best > cur) { ws[self] = best; *flag = 1; }
}

// Only sinks with an equal-height neighbour can change in the merge above;
// they are collected once (the equal-height bits come from the FlowDirs pass on
// the same surface) and the merge iterates over them alone.
__global__ void __launch_bounds__(256) k_merge_list(const int *cand, int n, const uint8_t *SFD, const uint8_t *eqb,
                                                    int *list, int *count)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += gridDim.x * blockDim.x) {
        int k = k0 + threadIdx.x;
        int i = k < n ? cand[k] : 0;
        bool keep = k < n && SFD[i] == 0 && eqb[i] != 0;
        int at = block_append(count, keep, s_wave, &s_base);
        if (keep) list[at] = i;
    }
}

// cells still pointing at another cell after the first jumps
__global__ void __launch_bounds__(256) k_jump_list(long N, const uint8_t *mask, const int *ws, int *list, int *count)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    compact_chunks<COMPACT_ITEMS>(N, count, s_wave, &s_base, [&](long i) { return mask[i] && ws[i] < 0; },
                                  [&](long i, int at) { list[at] = (int)i; });
}

// One lane per wave raises a progress flag (every lane writing the same word
// serialises those writes).
__device__ __forceinline__ void wave_flag(bool f, int *flag)
{
    const unsigned long long m = __ballot(f);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) *flag = 1;
}

__global__ void k_jump_some(int n, const int *list, int *ws, int *flag)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int i = list[k];
        int v = ws[i];
        bool more = false;
        if (v < 0) {
            int t = ws[-v - 1];
            ws[i] = t;
            more = t < 0;
        }
        wave_flag(more, flag);
    }
}

// identify_watershedMFD (newflow.cu:469-487): pointer jumping; in-place reads
// of already-jumped targets only shorten chains
__global__ void k_jump(long N, const uint8_t *mask, int *ws, int *flag)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        int v = ws[i];
        bool more = false;
        if (v < 0 && mask[i]) {
            int t = ws[-v - 1];
            ws[i] = t;
            more = t < 0;
        }
        wave_flag(more, flag);
    }
}

// ------------------------------------------------------ edges, spill

#define HASH_EMPTY 0xFFFFFFFFFFFFFFFFULL

__device__ __forceinline__ unsigned long long mix64(unsigned long long k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
    return k;
}

// identify_watershed_boundary_write_backMFD (newflow.cu:787-877) without the
// fixed-size table: every mask cell c and in-grid neighbour n with different
// labels and z_c >= z_n contributes edge {ws_c, ws_n} at height z_c, and the
// edge list keeps the minimum per unordered pair (order-free, exact).  The
// cells of a block share most of their basin pairs (a boundary runs through
// many cells): pairs are first merged in an LDS table (minimum weight per
// pair), and each block appends its distinct pairs to a list with one atomic;
// pairs that find no LDS slot are appended one by one.  Blocks of 64 x 16
// cells with a 1024-entry table (16 KB) keep eight blocks resident per
// compute unit.  The list is then sorted by pair and reduced to the minimum
// weight per pair.  Keys pack
// (lo << B) | hi with B bits per label.  The count may exceed the capacity:
// the caller then grows the list and runs again.
#define EDGE_LDS 1024
#define EDGE_RG 4   // row groups of 4 per block: 64 x (4 * EDGE_RG) cells
__global__ void __launch_bounds__(256) k_edges_emit(int W, int H, const uint8_t *__restrict__ mask,
                                                    const int *__restrict__ ws, const double *__restrict__ z,
                                                    const uint8_t *__restrict__ demid, unsigned long long *keys,
                                                    unsigned long long *vals, long cap, int *count, int B)
{
    __shared__ unsigned long long lk[EDGE_LDS], lv[EDGE_LDS];
    __shared__ int s_wave[4];
    __shared__ int s_base;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int q = tid; q < EDGE_LDS; q += 256) { lk[q] = HASH_EMPTY; lv[q] = HASH_EMPTY; }
    __syncthreads();
    const int icol = blockIdx.x * blockDim.x + threadIdx.x;
    for (int g = 0; g < EDGE_RG; g++) {
        const int irow = (blockIdx.y * EDGE_RG + g) * blockDim.y + threadIdx.y;
        const long c = (long)irow * W + icol;
        if (icol < W && irow < H && mask[c]) {
            int me = ws[c];
            double zc = z[c];
            unsigned long long wenc = ord_enc(zc);
            // the eight neighbours' labels, heights and DEM ids in flight at
            // once (a wait per neighbour inside the branches below would
            // serialise them); an off-grid neighbour reads the cell itself,
            // whose label never makes an edge
            int oth[8];
            double zn[8];
            uint32_t dn[8];   // 32-bit: byte registers would be packed, with a wait per load
#pragma unroll
            for (int d = 0; d < 8; d++) {
                const int cx = icol + cdx(d), cy = irow + cdy(d);
                const long nb = (cx >= 0 && cx < W && cy >= 0 && cy < H) ? (long)cy * W + cx : c;
                oth[d] = ws[nb];
                zn[d] = z[nb];
                dn[d] = demid ? demid[nb] : 0;
            }
            const int mydem = demid ? demid[c] : 0;
            int seen[8];
            int ns = 0;
#pragma unroll
            for (int d = 0; d < 8; d++) {
                const int other = oth[d];
                if (other == me) continue;
                // a neighbour outside the cell's own DEM is off-grid for that DEM
                if (dn[d] != mydem) continue;
                if (!(zc >= zn[d])) continue;
                bool dup = false;
                for (int q = 0; q < ns; q++) dup |= (seen[q] == other);
                if (dup) continue;
                seen[ns++] = other;
                unsigned long long lo = (unsigned)(me < other ? me : other), hi = (unsigned)(me < other ? other : me);
                unsigned long long key = (lo << B) | hi;
                unsigned h = (unsigned)(mix64(key) & (EDGE_LDS - 1));
                bool done = false;
                for (int p = 0; p < 16 && !done; p++) {
                    unsigned long long prev = atomicCAS(&lk[h], HASH_EMPTY, key);
                    if (prev == HASH_EMPTY || prev == key) { atomicMin(&lv[h], wenc); done = true; }
                    h = (h + 1) & (EDGE_LDS - 1);
                }
                if (!done) {
                    int at = atomicAdd(count, 1);
                    if (at < cap) { keys[at] = key; vals[at] = wenc; }
                }
            }
        }
    }
    __syncthreads();
    constexpr int PERT = EDGE_LDS / 256;
    int mine = 0;
#pragma unroll
    for (int j = 0; j < PERT; j++) mine += lk[tid * PERT + j] != HASH_EMPTY;
    int pos = block_append_n(count, mine, s_wave, &s_base);
#pragma unroll
    for (int j = 0; j < PERT; j++) {
        unsigned long long key = lk[tid * PERT + j];
        if (key == HASH_EMPTY) continue;
        if (pos < cap) { keys[pos] = key; vals[pos] = lv[tid * PERT + j]; }
        pos++;
    }
}

// unique pairs (sorted) to the edge list; label flags as in the compaction
__global__ void k_edges_decode(int ne, const unsigned long long *ukeys, const unsigned long long *uw, int B,
                               uint32_t *elo, uint32_t *ehi, unsigned long long *ew, uint8_t *seen,
                               const uint8_t *labdem, DemScal *ds)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne) return;
    unsigned long long key = ukeys[k];
    uint32_t lo = (uint32_t)(key >> B), hi = (uint32_t)(key & ((1ULL << B) - 1));
    elo[k] = lo;
    ehi[k] = hi;
    ew[k] = uw[k];
    if (!seen[lo]) seen[lo] = 1;
    if (!seen[hi]) seen[hi] = 1;
    if (lo == 0 && !ds[labdem[hi]].zero_seen) ds[labdem[hi]].zero_seen = 1;
}

// distinct labels of each slot in the edge set (newflow.cu:1045: data->sinkcount)
__global__ void __launch_bounds__(256) k_count_labels(long n, const uint8_t *seen, const uint8_t *labdem, int ns,
                                                      DemScal *ds)
{
    __shared__ int hist[LEM_MAX_SLOTS + 1];
    for (int j = threadIdx.x; j < ns; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    for (long l = 1 + blockIdx.x * (long)blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x)
        if (seen[l]) atomicAdd(&hist[labdem[l]], 1);
    slot_hist_flush(hist, ds, (int)offsetof(DemScal, sinkcount), ns);
}

__global__ void k_slots_sinkcount_init(int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    ds[j].sinkcount = 0;
    ds[j].zero_seen = 0;
}

// label 0 counts as one of a slot's labels when the slot has an edge to it;
// a slot without sinks skipped flooding and reports its sink count
__global__ void k_slots_sinkcount_fin(const DemDesc *dd, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    if (ds[j].sinks_pass1 == 0) ds[j].sinkcount = 0;
    else ds[j].sinkcount += ds[j].zero_seen;
}

__global__ void k_minimax_init(long n, unsigned long long *r, uint8_t *seen)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        r[i] = (i == 0) ? ord_enc(-lm_inf()) : ord_enc(lm_inf());
        seen[i] = 0;
    }
}

// Bottleneck distance to label 0 over the basin graph:
// r[b] = min over edges (b,c) of max(w_bc, r[c]), r[0] = -inf.  This is the
// value the reference's Kruskal/union-find pass assigns (routeflow.cu:529-572):
// a basin is raised to the height at which it first joins label 0 in
// ascending-weight order.  Relaxation in any order reaches the same fixed point.
__global__ void k_minimax_relax(int ne, const uint32_t *elo, const uint32_t *ehi, const unsigned long long *ew,
                                unsigned long long *r, int *flag)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne) return;
    uint32_t a = elo[k], b = ehi[k];
    unsigned long long w = ew[k];
    unsigned long long ra = r[a], rb = r[b];
    unsigned long long ca = w > rb ? w : rb;
    unsigned long long cb = w > ra ? w : ra;
    if (ca < ra) { atomicMin(&r[a], ca); *flag = 1; }
    if (cb < rb) { atomicMin(&r[b], cb); *flag = 1; }
}

// ---- spill heights by Borůvka phases
//
// The bottleneck distance above (the relaxation's fixed point) equals the
// heaviest edge on the minimum spanning tree's path from label 0, with edges
// ordered by encoded weight and then by position in the label-sorted list (a
// strict total order, so the tree is unique).  Borůvka builds that tree in
// phases: every component takes its lightest outgoing edge, and the chosen
// edges join components.  In the tree of components the phases form (each
// component a child of the one it joins, that edge weighing what the
// component chose), the heaviest edge between two leaves equals the
// heaviest edge on the spanning tree's path between them (V. King, "A
// simpler minimum spanning tree verification algorithm", Algorithmica 18,
// 1997).  So a label v that joins label 0's component in phase p gets
//     r(v) = max(-inf, the weights v's components chose in phases 1..p,
//                the weights label 0's components chose in phases 1..p),
// both running maxima, one per label and one for label 0.  Edges of weight
// +inf or above (encoded) lower nothing in the relaxation and are left out;
// a label that never joins keeps +inf (the cut-off fallback below takes
// those).  About log2(labels) phases instead of one relaxation pass per
// link of the longest bottleneck path.
//
// Per phase (bw/be: the chosen weight and edge per component, double-buffered
// by phase parity; nxt: the component each root joins):
__global__ void k_bv_init(long n, uint32_t *comp, unsigned long long *mv, unsigned long long *bw0,
                          unsigned long long *bw1, uint32_t *be0, uint32_t *be1)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        comp[i] = (uint32_t)i;
        mv[i] = 0ULL;
        bw0[i] = bw1[i] = ~0ULL;
        be0[i] = be1[i] = ~0u;
    }
}

// the lightest outgoing weight of each component.  The edge list is sorted
// by (lo, hi), so the lanes of a wave whose edges leave one component on
// the lo side sit in runs: each run's weights are reduced in the wave
// (a minimum over edges of one component, in any grouping) and only the
// run's first lane updates the component, instead of one atomic per lane
// on the same address; the same for runs on the hi side.
__global__ void __launch_bounds__(256) k_bv_min_w(int ne, const uint32_t *__restrict__ lo,
                                                  const uint32_t *__restrict__ hi,
                                                  const unsigned long long *__restrict__ w,
                                                  const uint32_t *__restrict__ comp, unsigned long long *bw)
{
    const unsigned long long INF = ord_enc(lm_inf());
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool live = false;
    unsigned long long wk = ~0ULL;
    uint32_t ca = ~0u, cb = ~0u;
    if (k < ne) {
        wk = w[k];
        ca = comp[lo[k]];
        cb = comp[hi[k]];
        live = ca != cb && wk < INF;
    }
    // lanes without an edge to count carry a key no component has; the hi
    // side's components repeat in runs too once components have merged
    uint32_t key = live ? ca : ~0u, key2 = live ? cb : ~0u;
    unsigned long long v = live ? wk : ~0ULL, v2 = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long ov = __shfl_down(v, off, 64), ov2 = __shfl_down(v2, off, 64);
        const uint32_t okey = __shfl_down(key, off, 64), okey2 = __shfl_down(key2, off, 64);
        if (lane + off < 64 && okey == key && ov < v) v = ov;
        if (lane + off < 64 && okey2 == key2 && ov2 < v2) v2 = ov2;
    }
    // every lane takes part in the shuffles (a shuffle under a branch would
    // read the inactive lane 0 as 0, component 0's key)
    const uint32_t prev = __shfl_up(key, 1, 64), prev2 = __shfl_up(key2, 1, 64);
    const bool head = lane == 0 || prev != key;
    const bool head2 = lane == 0 || prev2 != key2;
    if (live && head && v < bw[ca]) atomicMin(&bw[ca], v);
    if (live && head2 && v2 < bw[cb]) atomicMin(&bw[cb], v2);
}

// the first edge of that weight
__global__ void k_bv_min_e(int ne, const uint32_t *__restrict__ lo, const uint32_t *__restrict__ hi,
                           const unsigned long long *__restrict__ w, const uint32_t *__restrict__ comp,
                           const unsigned long long *__restrict__ bw, uint32_t *be)
{
    const unsigned long long INF = ord_enc(lm_inf());
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne) return;
    const unsigned long long wk = w[k];
    const uint32_t ca = comp[lo[k]], cb = comp[hi[k]];
    if (ca == cb || wk >= INF) return;
    if (wk == bw[ca]) atomicMin(&be[ca], (uint32_t)k);
    if (wk == bw[cb]) atomicMin(&be[cb], (uint32_t)k);
}

// a component with a chosen edge points at the component across it (every
// other label at itself)
__global__ void k_bv_hook(long n, const uint32_t *__restrict__ lo, const uint32_t *__restrict__ hi,
                          const uint32_t *__restrict__ comp, const uint32_t *__restrict__ be, uint32_t *nxt)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t e = be[i];
        uint32_t t = (uint32_t)i;
        if (e != ~0u) {
            const uint32_t a = comp[lo[e]], b = comp[hi[e]];
            t = a == (uint32_t)i ? b : a;
        }
        nxt[i] = t;
    }
}

// Two components that chose the same edge point at each other (under a
// strict order no longer cycle forms); the smaller becomes the root.  Only
// the smaller writes, and the entry it reads is the larger's, which the
// larger never writes here.
__global__ void k_bv_break(long n, const uint32_t *__restrict__ be, uint32_t *nxt)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (be[i] == ~0u) continue;
        const uint32_t o = nxt[i];
        if (o > (uint32_t)i && nxt[o] == (uint32_t)i) nxt[i] = (uint32_t)i;
    }
}

// pointer jumping towards the roots of the joined components (a fixed
// number of passes; the readers below walk what is left)
__global__ void k_bv_jump(long n, uint32_t *nxt)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t t = nxt[i];
        const uint32_t tt = nxt[t];
        if (tt != t) nxt[i] = tt;
    }
}

// The root of x's joined component: the pointers form trees whose roots
// point at themselves (the only cycles Borůvka's choices make under a
// strict order are pairs, and k_bv_break turns those into roots), and after
// the jumping passes few links are left.  A walk longer than any tree could
// be raises the error flag instead of running on.
#define BV_WALK_MAX (1u << 20)
__device__ __forceinline__ uint32_t bv_root(const uint32_t *nxt, uint32_t x, int *err)
{
    for (uint32_t h = 0; h < BV_WALK_MAX; h++) {
        const uint32_t y = nxt[x];
        if (y == x) return x;
        x = y;
    }
    *err = 1;
    return x;
}

// label 0's running maximum and its new component: st[0], st[1]
__global__ void k_bv_root0(const uint32_t *comp, const unsigned long long *bw, const uint32_t *be,
                           const uint32_t *nxt, unsigned long long *st, int *err)
{
    const uint32_t c0 = comp[0];
    if (be[c0] != ~0u && bw[c0] > st[0]) st[0] = bw[c0];
    st[1] = bv_root(nxt, nxt[c0], err);
}

// every label: its running maximum, its new component, and its spill height
// if that component is label 0's; the other parity's entries are reset for
// the next phase (each label its own)
__global__ void k_bv_labels(long n, uint32_t *comp, const unsigned long long *__restrict__ bw,
                            const uint32_t *__restrict__ be, const uint32_t *__restrict__ nxt,
                            unsigned long long *mv, unsigned long long *r, const unsigned long long *st,
                            unsigned long long *bwn, uint32_t *ben, int *err)
{
    const unsigned long long INF = ord_enc(lm_inf()), NEG = ord_enc(-lm_inf());
    const unsigned long long m0 = st[0];
    const uint32_t root0 = (uint32_t)st[1];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t c = comp[i];
        unsigned long long m = mv[i];
        if (be[c] != ~0u && bw[c] > m) m = bw[c];
        mv[i] = m;
        const uint32_t nc = bv_root(nxt, nxt[c], err);
        comp[i] = nc;
        bwn[i] = ~0ULL;
        ben[i] = ~0u;
        if (i != 0 && nc == root0 && r[i] == INF) {
            unsigned long long v = m > m0 ? m : m0;
            r[i] = v > NEG ? v : NEG;
        }
    }
}

// edges between different components, in order
__global__ void __launch_bounds__(256) k_bv_keep(int ne, const uint32_t *elo, const uint32_t *ehi,
                                                 const unsigned long long *ew, const uint32_t *comp, uint32_t *lo2,
                                                 uint32_t *hi2, unsigned long long *w2, int *count)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    const unsigned long long INF = ord_enc(lm_inf());
    compact_chunks_ordered<COMPACT_ITEMS>(
        ne, count, s_wave, &s_base, [&](long k) { return ew[k] < INF && comp[elo[k]] != comp[ehi[k]]; },
        [&](long k, int at) { lo2[at] = elo[k]; hi2[at] = ehi[k]; w2[at] = ew[k]; });
}

// Basins with no path to label 0 (only possible when NaN heights cut them
// off): the reference's union-find never marks them done, merges them in
// ascending order and leaves every member with the height of the component's
// last merge, i.e. the largest edge of the component's minimum spanning tree.
// That equals max over v of the bottleneck distance from any one member, so
// the component's smallest label seeds a second relaxation and the component
// takes the maximum.
__global__ void k_fb_unreached(int ne, const uint32_t *elo, const uint32_t *ehi, const unsigned long long *r, int *flag)
{
    const unsigned long long INF = ord_enc(lm_inf());
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ne && (r[elo[k]] == INF || r[ehi[k]] == INF)) *flag = 1;
}

__global__ void k_fb_comp_init(long n, uint32_t *comp, unsigned long long *cmax)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        comp[i] = (uint32_t)i;
        cmax[i] = 0ULL;
    }
}

__global__ void k_fb_comp_relax(int ne, const uint32_t *elo, const uint32_t *ehi, const unsigned long long *r,
                                uint32_t *comp, int *flag)
{
    const unsigned long long INF = ord_enc(lm_inf());
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne) return;
    uint32_t a = elo[k], b = ehi[k];
    if (r[a] != INF || r[b] != INF) return;
    uint32_t ca = comp[a], cb = comp[b];
    if (ca < cb) { atomicMin(&comp[b], ca); *flag = 1; }
    else if (cb < ca) { atomicMin(&comp[a], cb); *flag = 1; }
}

// seeds: the smallest label of each cut-off component; other cut-off labels
// restart from +inf in a separate array (cmax doubles as that array)
__global__ void k_fb_seed(long n, const unsigned long long *r, const uint32_t *comp, const uint8_t *seen,
                          unsigned long long *r2)
{
    const unsigned long long INF = ord_enc(lm_inf());
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r2[i] = (r[i] == INF && seen[i] && comp[i] == (uint32_t)i) ? ord_enc(-lm_inf()) : INF;
}

__global__ void k_fb_max(long n, const unsigned long long *r, const uint32_t *comp, const uint8_t *seen,
                         const unsigned long long *r2, unsigned long long *out)
{
    const unsigned long long INF = ord_enc(lm_inf());
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (r[i] == INF && seen[i] && r2[i] != INF && r2[i] != ord_enc(-lm_inf())) atomicMax(&out[comp[i]], r2[i]);
}

__global__ void k_fb_apply(long n, unsigned long long *r, const uint32_t *comp, const uint8_t *seen,
                           const unsigned long long *out)
{
    const unsigned long long INF = ord_enc(lm_inf());
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (r[i] == INF && seen[i] && out[comp[i]] != 0ULL) r[i] = out[comp[i]];
}

// raise_watershedMFD (newflow.cu:879-894) into a copy of the surface; label 0
// and labels without a path keep raise = 0 as in the reference's zeroed array
__global__ void k_raise(long N, const uint8_t *mask, const int *ws, const unsigned long long *r,
                        const double *z, const uint8_t *demid, const DemScal *ds, double *z2)
{
    const unsigned long long INF = ord_enc(lm_inf());
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double v = z[i];
        const uint8_t m = mask[i];   // the cell's own loads in flight together
        const int lab = ws[i];
        const int sl = slot_of(demid, i);
        // floodingDriver only runs for a DEM with sinks left (routeflow.cu:913)
        if (m && ds[sl].sinks_pass1 > 0) {
            unsigned long long e = r[lab];
            double rs = (lab == 0 || e == INF) ? 0.0 : ord_dec(e);
            if (v < rs) v = rs;
        }
        z2[i] = v;
    }
}

// routeflow.cu:938-953
__global__ void k_merge(long N, int flowtype, const uint8_t *mask, uint8_t *fd, const uint8_t *SFD)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        uint8_t f = fd[i];
        if (flowtype == 0) f = SFD[i];
        else if (f == 0 && mask[i]) f = SFD[i];
        fd[i] = f;
    }
}

// ------------------------------------------------------------ host side

static int grid1d(long n, int block = 256, int cap = 8192)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// 256-thread blocks over the columns, a row per block row (rows beyond the
// grid's height are taken in strides)
static dim3 grid2d(int W, int H)
{
    return dim3((unsigned)((W + 255) / 256), (unsigned)(H < 32768 ? (H > 0 ? H : 1) : 32768));
}

static int read_scalars(lemhip_ctx *c)
{
    LEM_CHECK(c, hipMemcpyAsync(c->hscal, c->scal, sizeof(LemScalars), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

static int clear_flag(lemhip_ctx *c)
{
    LEM_CHECK(c, hipMemsetAsync(&c->scal->flag, 0, sizeof(int), c->s));
    return LEMHIP_OK;
}

static int read_flag(lemhip_ctx *c, int *v)
{
    LEM_CHECK(c, hipMemcpyAsync(&c->hscal->flag, &c->scal->flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    *v = c->hscal->flag;
    return LEMHIP_OK;
}

// FlowDirs + plateau routing on surface zs; slopes == nullptr for pass 2.
// Leaves the candidate list in c->cand and its length in *ncand_out; returns
// the number of rounds that changed something in *rounds.
template <typename TS>
static int directions_and_plateaus(lemhip_ctx *c, const double *zs, TS *slopes, int *ncand_out, int *rounds)
{
    long N = c->N;
    int W = c->W, H = c->H;
    uint8_t *eqb = c->dmask;   // equal-height neighbour bits; the accumulation passes reuse the buffer later
    LEM_CHECK(c, hipMemsetAsync(&c->scal->ncand, 0, sizeof(int), c->s));
    dim3 blk(64, 4);
    dim3 grd((W + 63) / 64, (H + 63) / 64);
    prof_begin(c, PROF_FLOWDIRS);
    k_flowdirs<TS><<<grd, blk, 0, c->s>>>(W, H, (int)c->p.cellsize, zs, c->mask, slopes, c->NP, c->fd, c->SFD, eqb,
                                          c->cand, &c->scal->ncand);
    prof_end(c, PROF_FLOWDIRS);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hscal->ncand, &c->scal->ncand, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    int n = c->hscal->ncand;
    *ncand_out = n;
    *rounds = 0;
    if (n == 0) return LEMHIP_OK;
    prof_begin(c, PROF_PLATEAU);
    const int ntx = (W + PTW - 1) / PTW, nty = (H + PTH - 1) / PTH, ntiles = ntx * nty;
    int *tile_has = c->pl_tiles, *tstamp = c->pl_tiles + ntiles, *tver = c->pl_tiles + 5 * (long)ntiles;
    int *L[3] = {c->pl_tiles + 2 * (long)ntiles, c->pl_tiles + 3 * (long)ntiles, c->pl_tiles + 4 * (long)ntiles};
    PlatCtl *ctl = reinterpret_cast<PlatCtl *>(c->pl_ctl);
    float *sp_in = c->sp0, *sp_out = c->sp1;
    int launch = c->pl_launch;
    if (launch > (1 << 29)) {  // tile stamps hold launch numbers: restart them well before they wrap
        LEM_CHECK(c, hipMemsetAsync(tstamp, 0, ntiles * sizeof(int), c->s));
        launch = 0;
    }
    LEM_CHECK(c, hipMemsetAsync(tile_has, 0, ntiles * sizeof(int), c->s));
    LEM_CHECK(c, hipMemsetAsync(ctl, 0, sizeof(PlatCtl), c->s));
    k_sp_init<<<grid1d(N), 256, 0, c->s>>>(N, sp_in);
    k_plat_tver_init<<<(ntiles + 255) / 256, 256, 0, c->s>>>(ntiles, tver);
    k_sp_seed<<<grid1d(n, 256, 1 << 30), 256, 0, c->s>>>(W, c->cand, n, sp_in, ntx, tile_has);
    k_sp_seedmark<<<grid1d(n, 256, 1 << 30), 256, 0, c->s>>>(W, H, c->cand, n, sp_in, eqb, ntx, nty, tile_has);
    k_plat_first_list<<<grid1d(ntiles, 256, 1024), 256, 0, c->s>>>(ntiles, tile_has, L[launch % 3], ctl, launch % 3);
    LEM_LAUNCHED(c);
    int bpc = 0;
    hipError_t oe = hipSuccess;
    switch (c->plat_k) {
    case 4: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_plat_block<4, TS, false>, PLAT_NT, 0); break;
    case 16: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_plat_block<16, TS, false>, PLAT_NT, 0); break;
    default: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_plat_block<8, TS, false>, PLAT_NT, 0); break;
    }
    if (oe != hipSuccess || bpc < 1) bpc = 1;
    int G = bpc * c->ncu;
    if (G > ntiles) G = ntiles;
    const int BATCH = 16;   // launches between host checks (an empty launch costs less than a read)
    int r = 0;  // rounds done in this pass
    long guard = 0;
    for (;;) {
        prof_begin(c, PROF_PLAT_K);
        for (int q = 0; q < BATCH; q++) {
            const int *a_cur = L[launch % 3];
            int *a_next = L[(launch + 1) % 3];
            switch (c->plat_k) {
#define PLAT_LAUNCH(K)                                                                                           \
    if (slopes)                                                                                                  \
        k_plat_block<K, TS, true><<<G, PLAT_NT, 0, c->s>>>(W, H, launch, r, a_cur, a_next, sp_in, sp_out, tver, eqb, \
                                                           c->fd, c->SFD, slopes, c->NP, tile_has, tstamp, ctl);      \
    else                                                                                                         \
        k_plat_block<K, TS, false><<<G, PLAT_NT, 0, c->s>>>(W, H, launch, r, a_cur, a_next, sp_in, sp_out, tver, eqb, \
                                                            c->fd, c->SFD, slopes, c->NP, tile_has, tstamp, ctl)
            case 4: PLAT_LAUNCH(4); break;
            case 16: PLAT_LAUNCH(16); break;
            default: PLAT_LAUNCH(8); break;
#undef PLAT_LAUNCH
            }
            ++launch;
            r += c->plat_k;
        }
        prof_end(c, PROF_PLAT_K, BATCH);
        LEM_LAUNCHED(c);
        LEM_CHECK(c, hipMemcpyAsync(&c->hint[24], ctl, 3 * sizeof(int), hipMemcpyDeviceToHost, c->s));
        LEM_CHECK(c, hipMemcpyAsync(&c->hint[27], &ctl->last_change, sizeof(int), hipMemcpyDeviceToHost, c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        if (c->hint[24 + launch % 3] == 0) break;   // nothing queued for the next launch
        // every launch with a moving tile changes at least one cell, and a
        // cell's distance only decreases through a finite set of path lengths
        if (++guard > (long)n + 2) { c->err = "plateau routing did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    c->pl_launch = launch;
    *rounds = c->hint[27];
    prof_end(c, PROF_PLATEAU, *rounds + 1);
    return LEMHIP_OK;
}

static int ensure_hash(lemhip_ctx *c, long want)
{
    long cap = 1024;
    while (cap < want) cap <<= 1;
    if (cap <= c->hcap) return LEMHIP_OK;
    if (c->hkeys) { (void)hipFree(c->hkeys); (void)hipFree(c->hvals); (void)hipFree(c->elo); (void)hipFree(c->ehi); (void)hipFree(c->ew); }
    for (int q = 0; q < 2; q++) {
        if (c->elo2[q]) { (void)hipFree(c->elo2[q]); (void)hipFree(c->ehi2[q]); (void)hipFree(c->ew2[q]); }
        c->elo2[q] = c->ehi2[q] = nullptr;
        c->ew2[q] = nullptr;
    }
    c->hkeys = c->hvals = c->ew = nullptr;
    c->elo = c->ehi = nullptr;
    c->hcap = 0;
    LEM_CHECK(c, hipMalloc(&c->hkeys, cap * 8));
    LEM_CHECK(c, hipMalloc(&c->hvals, cap * 8));
    LEM_CHECK(c, hipMalloc(&c->elo, cap * 4));
    LEM_CHECK(c, hipMalloc(&c->ehi, cap * 4));
    LEM_CHECK(c, hipMalloc(&c->ew, cap * 8));
    for (int q = 0; q < 2; q++) {
        LEM_CHECK(c, hipMalloc(&c->elo2[q], cap * 4));
        LEM_CHECK(c, hipMalloc(&c->ehi2[q], cap * 4));
        LEM_CHECK(c, hipMalloc(&c->ew2[q], cap * 8));
    }
    c->hcap = cap;
    c->ecap = cap;
    return LEMHIP_OK;
}

static int ensure_labels(lemhip_ctx *c, long nl)
{
    if (nl <= c->lcap) return LEMHIP_OK;
    if (c->rmm) { (void)hipFree(c->rmm); (void)hipFree(c->seen); (void)hipFree(c->comp); (void)hipFree(c->cmax); }
    c->rmm = nullptr;
    c->seen = nullptr;
    c->comp = nullptr;
    c->cmax = nullptr;
    long cap = nl + nl / 4 + 1024;
    LEM_CHECK(c, hipMalloc(&c->rmm, cap * 8));
    LEM_CHECK(c, hipMalloc(&c->seen, cap));
    LEM_CHECK(c, hipMalloc(&c->comp, cap * 4));
    LEM_CHECK(c, hipMalloc(&c->cmax, cap * 8));
    c->lcap = cap;
    return LEMHIP_OK;
}

static int cutoff_components(lemhip_ctx *c, int nlab, int ne)
{
    long n = (long)nlab + 1;
    // second relaxation array: reuse the hash value table (>= 8*(nlab+1) slots)
    unsigned long long *r2 = c->hvals;
    k_fb_comp_init<<<grid1d(n), 256, 0, c->s>>>(n, c->comp, c->cmax);
    LEM_LAUNCHED(c);
    for (int it = 0;; ++it) {
        LEM_TRY(clear_flag(c));
        k_fb_comp_relax<<<grid1d(ne, 256, 1 << 30), 256, 0, c->s>>>(ne, c->elo, c->ehi, c->rmm, c->comp, &c->scal->flag);
        LEM_LAUNCHED(c);
        int f = 0;
        LEM_TRY(read_flag(c, &f));
        if (!f) break;
        if (it > nlab + 2) { c->err = "component labelling did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    k_fb_seed<<<grid1d(n), 256, 0, c->s>>>(n, c->rmm, c->comp, c->seen, r2);
    LEM_LAUNCHED(c);
    for (int it = 0;; ++it) {
        LEM_TRY(clear_flag(c));
        k_minimax_relax<<<grid1d(ne, 256, 1 << 30), 256, 0, c->s>>>(ne, c->elo, c->ehi, c->ew, r2, &c->scal->flag);
        LEM_LAUNCHED(c);
        int f = 0;
        LEM_TRY(read_flag(c, &f));
        if (!f) break;
        if (it > nlab + 2) { c->err = "cut-off spill heights did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    k_fb_max<<<grid1d(n), 256, 0, c->s>>>(n, c->rmm, c->comp, c->seen, r2, c->cmax);
    k_fb_apply<<<grid1d(n), 256, 0, c->s>>>(n, c->rmm, c->comp, c->seen, c->cmax);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

// Spill heights of the labels into c->rmm (initialised: label 0 -inf, the
// rest +inf) by Borůvka phases over the label-sorted edge list.
static int spill_heights(lemhip_ctx *c, int nlab, int ne)
{
    const long n = (long)nlab + 1;
    if (ne <= 0) return LEMHIP_OK;
    // per-label scratch: running maxima, chosen weights (parity 1) and the two
    // scalars in the hash value table, chosen edges and roots in the hash key
    // table (both >= 8 * (nlab + 1) entries of 8 bytes), chosen weights
    // (parity 0) in cmax
    unsigned long long *mv = c->hvals, *bw[2] = {c->cmax, c->hvals + n}, *st = c->hvals + 2 * n;
    uint32_t *be[2] = {reinterpret_cast<uint32_t *>(c->hkeys), reinterpret_cast<uint32_t *>(c->hkeys) + n};
    uint32_t *nxt = reinterpret_cast<uint32_t *>(c->hkeys) + 2 * n;
    k_bv_init<<<grid1d(n), 256, 0, c->s>>>(n, c->comp, mv, bw[0], bw[1], be[0], be[1]);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemsetAsync(st, 0, 2 * sizeof(unsigned long long), c->s));
    const uint32_t *lo = c->elo, *hi = c->ehi;
    const unsigned long long *w = c->ew;
    int na = ne, buf = 0;
    LEM_TRY(clear_flag(c));   // raised only by a pointer walk that does not end
    // the profile slot times the kernels only, not the host reads between them
    for (int ph = 0; na > 0; ph++) {
        const int p = ph & 1;
        const int ge = grid1d(na, 256, 1 << 30);
        prof_begin(c, PROF_MM_K);
        k_bv_min_w<<<ge, 256, 0, c->s>>>(na, lo, hi, w, c->comp, bw[p]);
        k_bv_min_e<<<ge, 256, 0, c->s>>>(na, lo, hi, w, c->comp, bw[p], be[p]);
        k_bv_hook<<<grid1d(n), 256, 0, c->s>>>(n, lo, hi, c->comp, be[p], nxt);
        k_bv_break<<<grid1d(n), 256, 0, c->s>>>(n, be[p], nxt);
        // five jumping passes leave chains at most 1/32 as long; the root and
        // label kernels walk the rest (no host read per phase for the jumps)
        for (int q = 0; q < 5; q++) k_bv_jump<<<grid1d(n), 256, 0, c->s>>>(n, nxt);
        prof_end(c, PROF_MM_K, 9);
        LEM_LAUNCHED(c);
        uint32_t *lo2 = c->elo2[buf], *hi2 = c->ehi2[buf];
        unsigned long long *w2 = c->ew2[buf];
        LEM_CHECK(c, hipMemsetAsync(&c->scal->nedges, 0, sizeof(int), c->s));
        prof_begin(c, PROF_MM_K);
        k_bv_root0<<<1, 1, 0, c->s>>>(c->comp, bw[p], be[p], nxt, st, &c->scal->flag);
        k_bv_labels<<<grid1d(n), 256, 0, c->s>>>(n, c->comp, bw[p], be[p], nxt, mv, c->rmm, st, bw[p ^ 1],
                                                  be[p ^ 1], &c->scal->flag);
        k_bv_keep<<<grid_chunks(na), 256, 0, c->s>>>(na, lo, hi, w, c->comp, lo2, hi2, w2, &c->scal->nedges);
        prof_end(c, PROF_MM_K, 3);
        LEM_LAUNCHED(c);
        LEM_TRY(read_scalars(c));
        if (c->hscal->flag) { c->err = "spill-height component pointers form a cycle"; return LEMHIP_E_NOT_CONVERGED; }
        const int nk = c->hscal->nedges;
        // every component with an outgoing edge joins another each phase, so
        // the components with edges at least halve
        if (nk >= na && ph > 0) { c->err = "spill-height phases made no progress"; return LEMHIP_E_NOT_CONVERGED; }
        na = nk;
        lo = lo2; hi = hi2; w = w2;
        buf ^= 1;
        if (ph > 64) { c->err = "spill-height phases did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    return LEMHIP_OK;
}

// floodingDriver (routeflow.cu:202-676) up to the raised surface in c->z2
static int flooding(lemhip_ctx *c, int ncand1)
{
    long N = c->N;
    int W = c->W, H = c->H;
    // labels
    k_sink_flags<<<grid2d(W, H), 256, 0, c->s>>>(W, H, c->mask, c->SFD, c->demid, c->ddesc, c->scan_tmp);
    LEM_LAUNCHED(c);
    size_t need = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, c->scan_tmp, c->ws, (int)N, c->s);
    if (need > c->cub_bytes) {
        if (c->cub_tmp) (void)hipFree(c->cub_tmp);
        c->cub_tmp = nullptr;
        LEM_CHECK(c, hipMalloc(&c->cub_tmp, need));
        c->cub_bytes = need;
    }
    // prefix lands in ws, then k_init_ws rewrites ws in place from it: use z2 as the prefix buffer instead
    int *prefix = reinterpret_cast<int *>(c->z2);
    LEM_CHECK(c, hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, need, c->scan_tmp, prefix, (int)N, c->s));
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[0], c->scan_tmp + (N - 1), sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[1], prefix + (N - 1), sizeof(int), hipMemcpyDeviceToHost, c->s));
    k_init_ws<<<grid2d(W, H), 256, 0, c->s>>>(W, H, c->mask, c->SFD, prefix, c->demid, c->ddesc, c->ws, c->labdem);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    int nlab = c->hint[0] + c->hint[1];
    c->h_nlab = nlab;
    // sink merge over the sinks that have an equal-height neighbour (list in the
    // prefix buffer, free after k_init_ws)
    {
        int *mlist = reinterpret_cast<int *>(c->z2);
        LEM_CHECK(c, hipMemsetAsync(&c->scal->ncand, 0, sizeof(int), c->s));
        k_merge_list<<<grid1d(ncand1, 256, 4096), 256, 0, c->s>>>(c->cand, ncand1, c->SFD, c->dmask, mlist,
                                                                 &c->scal->ncand);
        LEM_LAUNCHED(c);
        LEM_TRY(read_scalars(c));
        int nm = c->hscal->ncand;
        for (int it = 0; nm > 0; it += 4) {
            LEM_TRY(clear_flag(c));
            for (int q = 0; q < 4; q++)
                k_sink_merge<<<grid1d(nm, 256, 1 << 30), 256, 0, c->s>>>(W, H, mlist, nm, c->SFD, c->dem, c->ws,
                                                                         &c->scal->flag);
            LEM_LAUNCHED(c);
            int f = 0;
            LEM_TRY(read_flag(c, &f));
            if (!f) break;
            if (it / 4 > nm + 2) { c->err = "sink merge did not converge"; return LEMHIP_E_NOT_CONVERGED; }
        }
    }
    // pointer jumping: four passes over every cell, then only the cells still pointing
    for (int it = 0;; it += 4) {
        LEM_TRY(clear_flag(c));
        for (int q = 0; q < 4; q++) k_jump<<<grid1d(N), 256, 0, c->s>>>(N, c->mask, c->ws, &c->scal->flag);
        LEM_LAUNCHED(c);
        int f = 0;
        LEM_TRY(read_flag(c, &f));
        if (!f) break;
        int *jlist = c->scan_tmp;
        LEM_CHECK(c, hipMemsetAsync(&c->scal->ncand, 0, sizeof(int), c->s));
        k_jump_list<<<grid_chunks(N), 256, 0, c->s>>>(N, c->mask, c->ws, jlist, &c->scal->ncand);
        LEM_LAUNCHED(c);
        LEM_TRY(read_scalars(c));
        int nj = c->hscal->ncand;
        for (int it2 = 0; nj > 0; it2 += 4) {
            LEM_TRY(clear_flag(c));
            for (int q = 0; q < 4; q++)
                k_jump_some<<<grid1d(nj, 256, 8192), 256, 0, c->s>>>(nj, jlist, c->ws, &c->scal->flag);
            LEM_LAUNCHED(c);
            LEM_TRY(read_flag(c, &f));
            if (!f) break;
            if (it2 > 128) { c->err = "watershed pointer jumping did not converge"; return LEMHIP_E_NOT_CONVERGED; }
        }
        break;
    }
    // edges: block-distinct pairs appended, sorted by pair, reduced to the
    // minimum weight per pair; the list comes out sorted by label (the
    // Borůvka phases read neighbouring labels' components from the same
    // lines, and their ties are broken by position in this order)
    LEM_TRY(ensure_labels(c, (long)nlab + 1));
    int B = 1;
    while ((1L << B) <= (long)nlab) B++;
    long want = 8L * (nlab + 1);
    if (want < 1024) want = 1024;
    long np = 0;
    for (int attempt = 0;; ++attempt) {
        LEM_TRY(ensure_hash(c, want));
        LEM_CHECK(c, hipMemsetAsync(&c->scal->nedges, 0, sizeof(int), c->s));
        dim3 blk(64, 4), grd((W + 63) / 64, (H + 4 * EDGE_RG - 1) / (4 * EDGE_RG));
        k_edges_emit<<<grd, blk, 0, c->s>>>(W, H, c->mask, c->ws, c->dem, c->demid, c->hkeys, c->hvals, c->hcap,
                                            &c->scal->nedges, B);
        LEM_LAUNCHED(c);
        LEM_TRY(read_scalars(c));
        np = c->hscal->nedges;
        if (np <= c->hcap) break;
        want = np + np / 4;
        if (attempt > 8) { c->err = "edge list overflow"; return LEMHIP_E_NOMEM; }
    }
    k_minimax_init<<<grid1d(nlab + 1), 256, 0, c->s>>>(nlab + 1, c->rmm, c->seen);
    k_slots_sinkcount_init<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->nslot, c->dsc);
    LEM_LAUNCHED(c);
    int ne = 0;
    if (np > 0) {
        unsigned long long *skeys = c->ew2[0], *svals = c->ew2[1];
        size_t need3 = 0, need4 = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, need3, c->hkeys, skeys, c->hvals, svals, (int)np, 0, 2 * B,
                                                 c->s);
        (void)hipcub::DeviceReduce::ReduceByKey(nullptr, need4, skeys, c->hkeys, svals, c->hvals, &c->scal->nedges,
                                                hipcub::Min(), (int)np, c->s);
        if (need4 > need3) need3 = need4;
        if (need3 > c->cub_bytes) {
            if (c->cub_tmp) (void)hipFree(c->cub_tmp);
            c->cub_tmp = nullptr;
            LEM_CHECK(c, hipMalloc(&c->cub_tmp, need3));
            c->cub_bytes = need3;
        }
        size_t t3 = c->cub_bytes;
        LEM_CHECK(c, hipcub::DeviceRadixSort::SortPairs(c->cub_tmp, t3, c->hkeys, skeys, c->hvals, svals, (int)np, 0,
                                                         2 * B, c->s));
        t3 = c->cub_bytes;
        LEM_CHECK(c, hipcub::DeviceReduce::ReduceByKey(c->cub_tmp, t3, skeys, c->hkeys, svals, c->hvals,
                                                        &c->scal->nedges, hipcub::Min(), (int)np, c->s));
        LEM_TRY(read_scalars(c));
        ne = c->hscal->nedges;
        k_edges_decode<<<(ne + 255) / 256, 256, 0, c->s>>>(ne, c->hkeys, c->hvals, B, c->elo, c->ehi, c->ew, c->seen,
                                                          c->labdem, c->dsc);
        LEM_LAUNCHED(c);
    }
    k_count_labels<<<grid1d(nlab + 1, 256, 1024), 256, 0, c->s>>>(nlab + 1, c->seen, c->labdem, c->nslot, c->dsc);
    LEM_LAUNCHED(c);
    c->h_nedges = ne;

    LEM_TRY(spill_heights(c, nlab, ne));
    LEM_TRY(clear_flag(c));
    if (ne > 0) k_fb_unreached<<<grid1d(ne, 256, 1 << 30), 256, 0, c->s>>>(ne, c->elo, c->ehi, c->rmm, &c->scal->flag);
    LEM_LAUNCHED(c);
    int cut = 0;
    LEM_TRY(read_flag(c, &cut));
    if (cut) LEM_TRY(cutoff_components(c, nlab, ne));
    k_raise<<<grid1d(N), 256, 0, c->s>>>(N, c->mask, c->ws, c->rmm, c->dem, c->demid, c->dsc, c->z2);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

// Times the plateau pass with every rounds-per-launch candidate on the same
// surface (the pass is idempotent: it rebuilds directions, distances and marks
// from the surface) and keeps the fastest.
template <typename TS>
static int tune_plateau(lemhip_ctx *c, const double *zs, TS *slopes)
{
    static const int KS[3] = {4, 8, 16};
    // each choice timed twice, in the order 8 4 16 16 4 8, keeping its
    // faster time: a context tuned beside others sees their phases change
    // between its timings, and one timing per choice picked 4 on some boxes
    static const int ORDER[6] = {1, 0, 2, 2, 0, 1};
    hipEvent_t e0, e1;
    LEM_CHECK(c, hipEventCreate(&e0));
    LEM_CHECK(c, hipEventCreate(&e1));
    for (int q = 0; q < 3; q++) c->tune_ms[8 + q] = 1e30;
    for (int o = 0; o < 6; o++) {
        const int q = ORDER[o];
        c->plat_k = KS[q];
        int n = 0, r = 0;
        LEM_CHECK(c, hipEventRecord(e0, c->s));
        LEM_TRY(directions_and_plateaus<TS>(c, zs, slopes, &n, &r));
        LEM_CHECK(c, hipEventRecord(e1, c->s));
        LEM_CHECK(c, hipEventSynchronize(e1));
        float ms = 0.f;
        LEM_CHECK(c, hipEventElapsedTime(&ms, e0, e1));
        if (ms < c->tune_ms[8 + q]) c->tune_ms[8 + q] = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    float best = 1e30f;
    int bestk = 8;
    for (int q = 0; q < 3; q++)
        if ((float)c->tune_ms[8 + q] < best) { best = (float)c->tune_ms[8 + q]; bestk = KS[q]; }
    // 8 rounds per launch fit the tile's region best (one region row per
    // wave, five cells per thread): another count only when clearly faster,
    // since contexts tuned side by side time each other's traffic too
    if (bestk != 8 && best > 0.9f * (float)c->tune_ms[9]) bestk = 8;
    c->plat_k = bestk;
    c->tune_plat_pending = false;
    return LEMHIP_OK;
}

template <typename TS>
static int fd_phase_t(lemhip_ctx *c, int iter)
{
    (void)iter;
    long N = c->N;
    // lowest cell for the outlet check, taken on the unadjusted surface
    prof_begin(c, PROF_FD);
    LEM_TRY(lem_slot_lowest(c, c->dem, 0, c->nslot));
    // pass 1 on the unraised surface; slopes keep these values
    int n1 = 0, r1 = 0;
    LEM_TRY(directions_and_plateaus<TS>(c, c->dem, reinterpret_cast<TS *>(c->slopes), &n1, &r1));
    k_slots_sinks_init<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->nslot, c->dsc);
    if (n1 > 0) {
        k_count_sinks<<<grid1d(n1, 256, 2048), 256, 0, c->s>>>(c->cand, n1, c->SFD, c->demid, c->nslot, c->dsc);
        LEM_LAUNCHED(c);
    }
    LEM_TRY(lem_read_slots(c));
    int sinks = 0;
    for (int j = 0; j < c->nslot; j++) sinks += c->hdesc[j].w > 0 ? c->hdsc[j].sinks_pass1 : 0;

    c->h_sinks1 = sinks;
    c->h_rounds1 = r1;
    c->h_rounds2 = 0;
    c->h_nlab = 0;
    c->h_nedges = 0;
    c->h_sinkcount = sinks;
    if (sinks == 0) k_slots_sinkcount_init<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->nslot, c->dsc);
    if (sinks > 0) {
        prof_begin(c, PROF_FLOOD);
        LEM_TRY(flooding(c, n1));
        k_slots_sinkcount_fin<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->ddesc, c->nslot, c->dsc);
        prof_end(c, PROF_FLOOD);
        int n2 = 0, r2 = 0;
        if (c->tune_plat_pending) LEM_TRY(tune_plateau<TS>(c, c->z2, (TS *)nullptr));
        LEM_TRY(directions_and_plateaus<TS>(c, c->z2, (TS *)nullptr, &n2, &r2));
        c->h_rounds2 = r2;
    }
    k_merge<<<grid1d(N), 256, 0, c->s>>>(N, c->p.flowdirectiontype, c->mask, c->fd, c->SFD);
    k_slots_fd_outlet<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->fd, c->ddesc, c->nslot, c->dsc);
    k_slots_outlet_adjust<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->dem, c->ddesc, c->nslot, c->dsc);
    LEM_LAUNCHED(c);
    prof_end(c, PROF_FD);
    return LEMHIP_OK;
}

int lem_fd_phase(lemhip_ctx *c, int iter)
{
    if (c->prec == LEMHIP_FP64) return fd_phase_t<double>(c, iter);
    return fd_phase_t<float>(c, iter);
}
