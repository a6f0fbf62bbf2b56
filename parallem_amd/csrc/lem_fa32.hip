// lem_fa32.hip — flow accumulation of the FP32 form under single flow
// directions (CalcMFD with one receiver per cell, mfd_simple.cu:33-171) as
// exact integer subtree sums.
//
// With one receiver per cell the donor graph is a forest, and the
// reference's value of a completed cell, w + max(0, sum of donor fa), is the
// sum of w over the cell's subtree (every w >= 0, so the clamp never
// acts).  Each cell's weight is rounded once to a 64-bit fixed-point integer
// q = round(w * 2^F), with F chosen per DEM from its total runoff so that no
// sum can exceed 2^61; integer addition is associative, so the subtree sums
// are exact and do not depend on the order, the tiling or the other DEMs of
// an atlas.  The only rounding is that of q (relative 2^-F-ish, far below the
// FP32 form's 1e-6 bound) and of the final conversion to the float pair.
//
// Two passes over tiles of TW x TH cells with a small global solve between:
//   A. each tile sums its own subtrees in LDS by pointer doubling (round t
//      adds every cell's partial sum to its 2^t-th in-tile ancestor), finds
//      each cell's in-tile root, and lists its exits (cells whose receiver is
//      in another tile) with their in-tile sums;
//   B. the exits form a forest of their own (an exit's parent is the exit
//      its receiver's in-tile path leads to); the same doubling over that
//      forest gives every exit its final value;
//   C. each tile sums its subtrees again with the final values of the exits
//      that flow into it added at their receivers, and writes fa.
//
// Cycles: fd[outlet] = WEST can close a cycle through the outlet (the only
// possible one: every other link goes strictly downhill or down a plateau's
// distance field).  Cells on a cycle never complete in the reference and keep
// fa = w.  The outlet's link is cut and its west neighbour carries a tag of
// 2^62: tagged cells are exactly the path from the west neighbour to its
// root.  If the outlet is tagged, that path is the cycle (fa = w); otherwise
// the path lies downstream of the outlet and gets the outlet's sum added.
#include "lem_ctx.h"

#define FX_TAG (1ULL << 62)
#define FX_NONE 0xFFFFu

// direction index of a single ESRI code, -1 for 0 or a multi-bit code
__device__ __forceinline__ int fx_dir(uint8_t code)
{
    switch (code) {
    case 64: return 0; case 128: return 1; case 1: return 2; case 2: return 3;
    case 4: return 4; case 8: return 5; case 16: return 6; case 32: return 7;
    default: return -1;
    }
}

__device__ __forceinline__ unsigned long long fx_quant(float w, int shift)
{
    return (unsigned long long)llrint(ldexp((double)w, shift));
}

__device__ __forceinline__ float2 fx_value(unsigned long long s, int shift)
{
    const double v = ldexp((double)s, -shift);
    const float hi = (float)v;
    return make_float2(hi, (float)(v - (double)hi));
}

// the tagged cell of a slot: the west neighbour of an outlet whose link to
// it exists (outlet valid, not in column 0, west neighbour valid)
__device__ __forceinline__ bool fx_tagged(long i, int W, const uint8_t *mask, int outlet)
{
    return outlet >= 1 && i == (long)outlet - 1 && (outlet % W) != 0 && mask[outlet] && mask[outlet - 1];
}

struct FxArgs {
    int W, H;
    const uint8_t *mask, *fd, *demid;
    const float *w;
    const DemScal *ds;
    const int *shift;            // per slot
    // pass A outputs / pass C inputs
    int *exit_id;                // per cell (exits only)
    int *xroot;                  // per cell (tile border cells only): exit its in-tile path leads to, or -1
    unsigned long long *ex_val;  // per exit: in-tile sum (A), final value after the forest solve
    int *ex_recv;                // per exit: receiver cell
    int *ex_count;
    int ex_cap;
    // pass C outputs
    float2 *fa;
    int *tag_cell;
    unsigned long long *tag_sum;
    int *tag_count;
    int tag_cap;
    DemScal *dsw;
};

// n items of this thread appended to a global list with one atomic per
// block; returns this thread's first index (every thread must call it)
__device__ __forceinline__ int fx_block_append(int *counter, int n, int *s_wave, int *s_base)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            const int cnt = s_wave[w];
            s_wave[w] = tot;
            tot += cnt;
        }
        *s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const int pos = *s_base + s_wave[wave] + x - n;
    __syncthreads();
    return pos;
}

// padded (TW+2) x (TH+2) index of tile cell (lr, lc); the ring holds the
// neighbours in other tiles (or off the grid)
template <int TW> __device__ __forceinline__ int fx_pad(int lr, int lc) { return (lr + 1) * (TW + 2) + lc + 1; }

template <int TW, int TH, bool FINAL>
__global__ void __launch_bounds__(256) k_fx_tile(FxArgs a)
{
    constexpr int NC = TW * TH, PER = NC / 256, PW = TW + 2, NP = (TW + 2) * (TH + 2);
    constexpr int NPPER = (NP + 255) / 256;
    static_assert(NC % 256 == 0, "tile shape");
    __shared__ unsigned long long x[NC];
    __shared__ uint16_t J[NC];
    __shared__ uint16_t R[FINAL ? 1 : NC];
    __shared__ uint8_t sfd[NP], smk[NP];
    __shared__ uint8_t isexit[FINAL ? 1 : NC];
    __shared__ int s_wave[4], s_base;
    const int W = a.W, H = a.H;
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH;
    for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
        const int tx0 = (tile % ntx) * TW, ty0 = (tile / ntx) * TH;
        __syncthreads();
        // directions and mask of the tile and its ring, every load in flight
        // before the first store (off-grid positions: no mask, no code)
        {
            uint32_t vf[NPPER], vm[NPPER];   // 32-bit: byte registers would be packed, with a wait per load
#pragma unroll
            for (int q = 0; q < NPPER; q++) {
                const int k = threadIdx.x + q * 256;
                const int gy = ty0 - 1 + k / PW, gx = tx0 - 1 + k % PW;
                const bool in = k < NP && gy >= 0 && gy < H && gx >= 0 && gx < W;
                const long n = (long)gy * W + gx;
                vf[q] = in ? a.fd[n] : 0;
                vm[q] = in ? a.mask[n] : 0;
            }
#pragma unroll
            for (int q = 0; q < NPPER; q++) {
                const int k = threadIdx.x + q * 256;
                if (k < NP) {
                    sfd[k] = vf[q];
                    smk[k] = vm[q];
                }
            }
        }
        __syncthreads();
        int nexit = 0;
        uint32_t exitbits = 0;
        int slot[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int li = threadIdx.x + q * 256;
            const int lr = li / TW, lc = li % TW;
            const int r = ty0 + lr, col = tx0 + lc;
            const int p = fx_pad<TW>(lr, lc);
            slot[q] = 0;
            unsigned long long v = 0;
            uint16_t j = FX_NONE;
            bool ex = false;
            if (smk[p]) {
                const long i = (long)r * W + col;
                const int s = a.demid ? (int)a.demid[i] : 0;
                slot[q] = s;
                const int outlet = a.ds[s].outlet;
                v = fx_quant(a.w[i], a.shift[s]);
                if (fx_tagged(i, W, a.mask, outlet)) v += FX_TAG;
                const int d = fx_dir(sfd[p]);
                if (d >= 0 && i != outlet) {
                    const int pn = p + cdy(d) * PW + cdx(d);
                    if (smk[pn]) {
                        const int nr = lr + cdy(d), nc = lc + cdx(d);
                        if (nr >= 0 && nr < TH && nc >= 0 && nc < TW) j = (uint16_t)(nr * TW + nc);
                        else ex = true;
                    }
                }
                if (FINAL && (lr == 0 || lr == TH - 1 || lc == 0 || lc == TW - 1)) {
                    // exits of other tiles that drain into this cell
#pragma unroll
                    for (int dd = 0; dd < 8; dd++) {
                        const int tr = lr + cdy(dd), tc = lc + cdx(dd);
                        if (tr >= 0 && tr < TH && tc >= 0 && tc < TW) continue;   // inside the tile
                        const int pn = p + cdy(dd) * PW + cdx(dd);
                        if (!smk[pn] || sfd[pn] != donor_code(dd)) continue;
                        const long y = (long)(r + cdy(dd)) * W + (col + cdx(dd));
                        if (y == outlet) continue;                                // cut link
                        v += a.ex_val[a.exit_id[y]];
                    }
                }
            }
            x[li] = v;
            J[li] = j;
            if (!FINAL) {
                R[li] = j == FX_NONE ? (uint16_t)li : j;
                isexit[li] = ex ? 1 : 0;
            }
            if (ex) {
                nexit++;
                exitbits |= 1u << q;
            }
        }
        __syncthreads();
        // pointer doubling: after round t every cell holds the sum over the
        // cells within 2^(t+1) - 1 links upstream in the tile
        for (;;) {
            unsigned long long v[PER];
            uint16_t jj[PER], rr[PER];
            int act = 0;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int li = threadIdx.x + q * 256;
                const uint16_t j = J[li];
                if (j != FX_NONE) {
                    act |= 1 << q;
                    v[q] = x[li];
                    jj[q] = J[j];
                    if (!FINAL) rr[q] = R[R[li]];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < PER; q++) {
                if (!((act >> q) & 1)) continue;
                const int li = threadIdx.x + q * 256;
                atomicAdd(&x[J[li]], v[q]);
                J[li] = jj[q];
                if (!FINAL) R[li] = rr[q];
            }
            if (!__syncthreads_or(act)) break;
        }
        // outputs
        if (!FINAL) {
            int k = fx_block_append(a.ex_count, nexit, s_wave, &s_base);
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int li = threadIdx.x + q * 256;
                const int lr = li / TW, lc = li % TW;
                const int r = ty0 + lr, col = tx0 + lc;
                if (r >= H || col >= W) continue;
                const long i = (long)r * W + col;
                if ((exitbits >> q) & 1) {
                    if (k < a.ex_cap) {
                        const int d = fx_dir(sfd[fx_pad<TW>(lr, lc)]);
                        a.exit_id[i] = k;
                        a.ex_val[k] = x[li];
                        a.ex_recv[k] = (int)((long)(r + cdy(d)) * W + (col + cdx(d)));
                    }
                    k++;
                }
                if (lr == 0 || lr == TH - 1 || lc == 0 || lc == TW - 1) {
                    const int root = R[li];
                    a.xroot[i] = isexit[root] ? (int)((long)(ty0 + root / TW) * W + (tx0 + root % TW)) : -1;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int li = threadIdx.x + q * 256;
                const int lr = li / TW, lc = li % TW;
                const int r = ty0 + lr, col = tx0 + lc;
                if (r >= H || col >= W) continue;
                const long i = (long)r * W + col;
                if (!smk[fx_pad<TW>(lr, lc)]) {
                    a.fa[i] = make_float2(0.f, 0.f);
                    continue;
                }
                const unsigned long long s = x[li];
                const int s2 = slot[q];
                if (i == a.ds[s2].outlet) a.dsw[s2].fx_outlet_sum = s;
                if (s >= FX_TAG) {
                    const int k = atomicAdd(a.tag_count, 1);
                    if (k < a.tag_cap) {
                        a.tag_cell[k] = (int)i;
                        a.tag_sum[k] = s;
                    }
                } else {
                    a.fa[i] = fx_value(s, a.shift[s2]);
                }
            }
        }
    }
}

// parent of each exit in the exit forest
__global__ void k_fx_parent(int n, const int *ex_recv, const int *xroot, const int *exit_id, int *P)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int xr = xroot[ex_recv[k]];
        P[k] = xr >= 0 ? exit_id[xr] : -1;
    }
}

// one doubling round over the exit forest in two launches: snapshot, then add
__global__ void k_fx_gather(int n, const unsigned long long *val, const int *P, unsigned long long *tv, int *tp)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int p = P[k];
        tv[k] = val[k];
        tp[k] = p >= 0 ? P[p] : -1;
    }
}

// Late rounds point many exits at the same few roots: the lanes of a wave
// with equal parents sit in runs, and each run's values are summed in the
// wave (a sum over the run's own lanes only, so every value counts once)
// and added by the run's first lane with one atomic.  Integer sums: the
// totals do not depend on the grouping.
__global__ void __launch_bounds__(256) k_fx_scatter(int n, unsigned long long *val, int *P,
                                                    const unsigned long long *tv, const int *tp, int *flag)
{
    const int lane = threadIdx.x & 63;
    bool more = false;
    // whole waves go round the loop together (the shuffles need every lane)
    for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += gridDim.x * blockDim.x) {
        const int k = k0 + (int)threadIdx.x;
        int p = -1;
        unsigned long long v = 0;
        if (k < n) {
            p = P[k];
            if (p >= 0) {
                v = tv[k];
                const int q = tp[k];
                P[k] = q;
                more |= q >= 0;
            }
        }
        const int pprev = __shfl_up(p, 1, 64);
        const bool head = lane == 0 || pprev != p;
        const unsigned long long heads = __ballot(head);
        const unsigned long long above = lane == 63 ? 0ULL : heads & (~0ULL << (lane + 1));
        const int end = above ? __ffsll((long long)above) - 1 : 64;   // first lane of the next run
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long o = __shfl_down(v, off, 64);
            if (lane + off < end) v += o;
        }
        if (head && p >= 0) atomicAdd(&val[p], v);
    }
    if (__any(more) && lane == 0) *flag = 1;
}

// tagged cells: the cycle through the outlet (fa = w, never completed), or
// the path below the outlet (the outlet's sum added)
__global__ void k_fx_tags(const int *count, int cap, const int *tag_cell, const unsigned long long *tag_sum,
                          const uint8_t *demid, const DemScal *ds, const int *shift, const float *w, float2 *fa,
                          int *uncomputed)
{
    const int n = *count < cap ? *count : cap;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int i = tag_cell[k];
        const int s = demid ? (int)demid[i] : 0;
        const unsigned long long o = ds[s].fx_outlet_sum;
        if (o >= FX_TAG) {
            fa[i] = make_float2(w[i], 0.f);
            atomicAdd(uncomputed, 1);
        } else {
            fa[i] = fx_value(tag_sum[k] - FX_TAG + o, shift[s]);
        }
    }
}

static int fx_grid(long n)
{
    long g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

// 0: done; 1: inputs the integer form cannot represent exactly (a negative
// or non-finite runoff weight, or too many tagged cells) — the caller runs
// the dependency engine instead
int lem_fa32_sfd(lemhip_ctx *c, int *fallback)
{
    *fallback = 0;
    LEM_TRY(lem_read_slots(c));
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[30], &c->scal->w_bad, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    if (c->hint[30]) { *fallback = 1; return LEMHIP_OK; }
    // per-slot shift: sums stay below 2^61 (the rounding adds at most N/2)
    for (int j = 0; j < c->nslot; j++) {
        const double tw = c->hdsc[j].totwater;
        int sh = 61;
        if (c->hdesc[j].w > 0) {
            if (!(tw >= 0.0) || tw > 1e300) { *fallback = 1; return LEMHIP_OK; }
            if (tw > 0.0) sh = 61 - (int)ceil(log2(tw * (1 + 1e-12)));
            if (sh > 1000) sh = 1000;
        }
        c->hfx_shift[j] = sh;
    }
    LEM_CHECK(c, hipMemcpyAsync(c->fx_shift, c->hfx_shift, sizeof(int) * c->nslot, hipMemcpyHostToDevice, c->s));
    const long N = c->N;
    FxArgs a;
    a.W = c->W;
    a.H = c->H;
    a.mask = c->mask;
    a.fd = c->fd;
    a.demid = c->demid;
    a.w = (const float *)c->w;
    a.ds = c->dsc;
    a.shift = c->fx_shift;
    a.exit_id = reinterpret_cast<int *>(c->scan_tmp);
    a.xroot = c->ws;
    unsigned long long *tv = reinterpret_cast<unsigned long long *>(c->z2);
    a.ex_val = tv + N / 2;
    a.ex_recv = c->cand;
    a.ex_count = &c->scal->nedges;
    a.ex_cap = (int)(N / 2);
    a.fa = reinterpret_cast<float2 *>(c->fa);
    a.tag_cell = c->fx_tag_cell;
    a.tag_sum = c->fx_tag_sum;
    a.tag_count = &c->scal->flag;
    a.tag_cap = (int)c->fx_tag_cap;
    a.dsw = c->dsc;
    constexpr int TW = 64, TH = 32;
    const long ntiles = (long)((c->W + TW - 1) / TW) * ((c->H + TH - 1) / TH);
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fx_tile<TW, TH, false>, 256, 0) != hipSuccess || bpc < 1)
        bpc = 2;
    int G = bpc * c->ncu;
    if (G > ntiles) G = (int)ntiles;
    // A: in-tile sums, exits, in-tile roots of border cells
    LEM_CHECK(c, hipMemsetAsync(a.ex_count, 0, sizeof(int), c->s));
    k_fx_tile<TW, TH, false><<<G, 256, 0, c->s>>>(a);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[31], a.ex_count, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    const int nex = c->hint[31];
    // the exit values and their snapshot share the 8-byte scratch of N cells;
    // exits are tile border cells (9% of a full tile), so only grids of a few
    // cells could have more
    if ((long)nex > N / 2) { *fallback = 1; return LEMHIP_OK; }
    // B: the exit forest
    if (nex > 0) {
        int *P = reinterpret_cast<int *>(c->sp0), *tp = reinterpret_cast<int *>(c->sp1);
        k_fx_parent<<<fx_grid(nex), 256, 0, c->s>>>(nex, a.ex_recv, a.xroot, a.exit_id, P);
        LEM_LAUNCHED(c);
        for (int round = 0;; round++) {
            LEM_CHECK(c, hipMemsetAsync(&c->scal->nlabels, 0, sizeof(int), c->s));
            for (int q = 0; q < 2; q++) {
                k_fx_gather<<<fx_grid(nex), 256, 0, c->s>>>(nex, a.ex_val, P, tv, tp);
                k_fx_scatter<<<fx_grid(nex), 256, 0, c->s>>>(nex, a.ex_val, P, tv, tp, &c->scal->nlabels);
            }
            LEM_LAUNCHED(c);
            LEM_CHECK(c, hipMemcpyAsync(&c->hint[32], &c->scal->nlabels, sizeof(int), hipMemcpyDeviceToHost, c->s));
            LEM_CHECK(c, hipStreamSynchronize(c->s));
            if (!c->hint[32]) break;
            if (round > 32) { c->err = "exit forest did not converge"; return LEMHIP_E_NOT_CONVERGED; }
        }
    }
    // C: final sums, fa, tagged cells
    LEM_CHECK(c, hipMemsetAsync(a.tag_count, 0, sizeof(int), c->s));
    k_fx_tile<TW, TH, true><<<G, 256, 0, c->s>>>(a);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemsetAsync(&c->scal->fa_uncomputed, 0, sizeof(int), c->s));
    k_fx_tags<<<64, 256, 0, c->s>>>(a.tag_count, a.tag_cap, a.tag_cell, a.tag_sum, c->demid, c->dsc, c->fx_shift,
                                    a.w, a.fa, &c->scal->fa_uncomputed);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[33], a.tag_count, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[34], &c->scal->fa_uncomputed, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    if (c->hint[33] > a.tag_cap) { *fallback = 1; return LEMHIP_OK; }
    c->h_fa_unc = c->hint[34];
    c->dag_rounds_last = 0;
    return LEMHIP_OK;
}
