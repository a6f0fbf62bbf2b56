// lem_dag.hip — the two fixed-point passes over the donor graph:
//   flow accumulation (CalcMFD, mfd_simple.cu:33-171, loops :204-222, :293-305)
//   sediment routing  (MFDsedbud, mfd_accum.cu:32-254, loops :294-321, :413-436)
//
// Both compute a cell once every donor is final, summing donor terms in the
// reference's fixed order E, SE, S, SW, W, NW, N, NE.  The reference does this
// with one full-grid launch per dependency level plus host round trips; here
// each workgroup owns a tile of cells, runs the tile's local dependency chains
// to completion in LDS-synchronised iterations, and only chains that cross a
// tile edge wait for the next round.  A cell's status word records the round
// in which it completed, so a workgroup trusts another tile's value only when
// it was written in an earlier launch (kernel boundary = visibility point).
#include "lem_ctx.h"
#include <string.h>

#define DAG_TILE_DONE 0x7FFFFFFF   // wake stamp of a tile with no pending cell
// status word: DAG_PENDING, r >= 0 done in round r (0 = before the first
// round), -(r+1) dead (can never complete) since round r

template <typename T> struct FaStore;
template <> struct FaStore<double> { using type = double; };
template <> struct FaStore<float> { using type = float2; };

// float-float (double-word) arithmetic for the compensated FP32 form.  A sum
// past the float range saturates to (inf, 0): the error terms would form
// inf - inf = NaN, while the FP64 form keeps a huge finite value (MFD path
// counts), which every consumer clamps the same way as inf (overland flow and
// incision caps, the gelifluction cut-off, the weathering cap)
__device__ __forceinline__ float2 ff_add(float2 a, float2 b)
{
    float s = a.x + b.x;
    if (__builtin_isinf(s)) return make_float2(s, 0.f);
    float v = s - a.x;
    float e = (a.x - (s - v)) + (b.x - v);
    e = e + (a.y + b.y);
    float hi = s + e;
    float lo = e - (hi - s);
    return make_float2(hi, lo);
}

// Packed LDS word of a tile cell during a round:
//   bits  0-7   donors in the tile not yet resolved; W_NEVER for a cell that
//               no decrement may hand over (not pending, parked, or dead
//               from the start): at most 8 decrements never reach zero
//   bits  8-15  donor bits (pull order)      bits 16-23  receiver bits (fd)
//   bits 24-27  dead donors seen             bits 28-29  state
//   bit  30     the sediment stone update applies
// A resolved cell hands its receivers one atomic add each (-1, or -1 plus one
// dead donor), and the add that brings a counter to zero returns the whole
// word of the receiver, so the next link of the chain starts with its donor
// and receiver bits and its dead flag already in registers.
#define W_CNT 0xFFu
#define W_NEVER 0x80u
#define W_DEAD1 (1u << 24)
#define W_STSH 28
#define W_UPD (1u << 30)

// ------------------------------------------------------------ cell bodies
//
// A cell body caches in LDS, per tile cell, what its receivers in the same
// tile will read, plus the same values for the one-cell ring around the tile
// (donors in other tiles, completed in an earlier launch), and evaluates one
// cell once all donors are final.  Walks then never wait on global loads.
// Values that receivers read are kept in a padded (TW+2) x (TH+2) layout that
// holds the ring too, so a donor is one LDS read at a fixed offset from the
// cell's padded index; the cell's own values use the same layout.

// padded index of neighbour d of the cell at padded index p
template <int TW> __host__ __device__ constexpr int poff(int d) { return cdy(d) * (TW + 2) + cdx(d); }
template <int TW> __device__ __forceinline__ int pidx(int lr, int lc) { return (lr + 1) * (TW + 2) + lc + 1; }

// padded offset of direction (k + 2) & 7 for a bit index k = 0..7 (E, SE, S,
// SW, W, NW, N, NE): bit k of a donor mask in pull order, and bit k of an
// ESRI code, both name that direction.  Two 2-bit tables, no branches.
__host__ __device__ constexpr uint32_t kpack(bool ydir)
{
    uint32_t r = 0;
    for (int k = 0; k < 8; k++) r |= (uint32_t)((ydir ? cdy((k + 2) & 7) : cdx((k + 2) & 7)) + 1) << (2 * k);
    return r;
}
// the eight offsets as signed bytes of one 64-bit constant (tiles up to 124
// cells wide): one shift and one sign extension
template <int TW> __host__ __device__ constexpr uint64_t kbytes()
{
    uint64_t r = 0;
    for (int k = 0; k < 8; k++)
        r |= (uint64_t)(uint8_t)(int8_t)(cdy((k + 2) & 7) * (TW + 2) + cdx((k + 2) & 7)) << (8 * k);
    return r;
}
template <int TW> __device__ __forceinline__ int koff(int k)
{
    if constexpr (TW + 3 <= 127) {
        constexpr uint64_t KB = kbytes<TW>();
        return (int)(int8_t)(uint8_t)(KB >> (8 * k));
    } else {
        constexpr uint32_t KX = kpack(false), KY = kpack(true);
        const int dx = (int)((KX >> (2 * k)) & 3u) - 1;
        const int dy = (int)((KY >> (2 * k)) & 3u) - 1;
        return dy * (TW + 2) + dx;
    }
}

// position (nr, nc) relative to the tile of ring cell h of a TW x TH tile: top
// row, bottom row (both with corners), left column, right column
template <int TW, int TH>
__device__ __forceinline__ void halo_pos(int h, int *nr, int *nc)
{
    if (h < TW + 2) { *nr = -1; *nc = h - 1; }
    else if (h < 2 * (TW + 2)) { *nr = TH; *nc = h - (TW + 2) - 1; }
    else if (h < 2 * (TW + 2) + TH) { *nr = h - 2 * (TW + 2); *nc = -1; }
    else { *nr = h - 2 * (TW + 2) - TH; *nc = TW; }
}

__device__ __forceinline__ void fa_acc(double &a, double v) { a += v * 1.0; }
__device__ __forceinline__ void fa_acc(float2 &a, float2 v) { a = ff_add(a, v); }
__device__ __forceinline__ void fa_clamp(double &a) { if (a < 0.0) a = 0.0; }
__device__ __forceinline__ void fa_clamp(float2 &a) { if (a.x < 0.0f) a = make_float2(0.f, 0.f); }
__device__ __forceinline__ double fa_final(double w, double a) { double v = w; v += a; return v; }
__device__ __forceinline__ float2 fa_final(float w, float2 a) { return ff_add(make_float2(w, 0.f), a); }

// Cell body interface used by the round kernel (p: padded index; every LDS
// array of a body is padded, so one index serves all of them):
//   Reg  load(i)        the cell's own inputs (tile cells, loaded in phase 1)
//   HReg load_halo(i)   what a receiver reads of a donor in another tile
//   Fin  store(L, p, r, &extra)  put the walk-visible values into LDS and
//        return what the cell writes to HBM when it completes that does not
//        depend on its donors (kept in registers); extra: word flags
//   compute(L, p, bits, upd)  the chain part, LDS only; the donors are the
//        set bits of the pull-order mask, summed lowest bit first, which is
//        the order E..NE of the reference (absent donors add nothing)
//   finish(L, p, i, fin, upd)  HBM writes of a cell completed in this
//        round, after the walks (coalesced, off the chains)
template <typename T, int TW, int TH>
struct FaCell {
    using FS = typename FaStore<T>::type;
    static constexpr int NC = TW * TH;
    static constexpr int NP = (TW + 2) * (TH + 2);
    static constexpr bool kCanDie = false;
    static constexpr int kWaves = 6;   // waves per SIMD the register budget is cut for
    const T *w;
    FS *fa;
    struct Lds {
        FS val[NP];
        T w[NP];
    };
    // loads are issued unconditionally and stored to LDS after one wait; the
    // value of a cell that is not final yet is never read
    struct Reg {
        T w;
        FS v;
    };
    using HReg = FS;
    struct Fin {};
    __device__ __forceinline__ Reg load(long i) const { return Reg{w[i], fa[i]}; }
    __device__ __forceinline__ HReg load_halo(long i) const { return fa[i]; }
    __device__ __forceinline__ Fin store(Lds &L, int p, const Reg &r, uint32_t *) const
    {
        L.w[p] = r.w;
        L.val[p] = r.v;
        return Fin{};
    }
    __device__ __forceinline__ void store_halo(Lds &L, int p, const HReg &r) const { L.val[p] = r; }
    // fa = w + max(0, sum of donor fa in the order E..NE) (mfd_simple.cu:71-168;
    // prop = 1 for every mask donor, runoffweight.cu:126-134, so fa*prop == fa)
    __device__ __forceinline__ void compute(Lds &L, int p, unsigned bits, bool) const
    {
        // one read per donor (chain cells mostly have one or two donors, and
        // a compensated add costs more than the read)
        const T wv = L.w[p];
        FS accum{};
        while (bits) {
            const int k = __ffs(bits) - 1;
            bits &= bits - 1;
            fa_acc(accum, L.val[p + koff<TW>(k)]);
        }
        fa_clamp(accum);
        L.val[p] = fa_final(wv, accum);
    }
    __device__ __forceinline__ void finish(const Lds &L, int p, long i, const Fin &, bool) const { fa[i] = L.val[p]; }
};

// MFDsedbud body (mfd_accum.cu:63-252) split by what it reads:
//  * deposition from the donors' erosion, incision and gelifluction, which
//    are inputs of this pass: computed for every cell by the prep kernel
//    into a scratch array (dps);
//  * the new soil thickness and the weights of the stone update, from the
//    cell's own values: computed when the tile is loaded;
//  * the stone update, the only term that reads donors' results:
//      stone = a + b * (sum of the donors' updated stone, order E..NE)
//    with a = old * stone and b = 1 - old (old = soilT / sediment thickness),
//    the same operations and roundings as the body written in one piece.
template <typename T, int TW, int TH>
struct SedCell {
    static constexpr int NC = TW * TH;
    static constexpr int NP = (TW + 2) * (TH + 2);
    static constexpr bool kCanDie = true;
    static constexpr int kWaves = 4;   // fewer waves, but no spills in the FP32 form
    const T *ero, *inci, *geli, *dps;
    T *depo, *stone, *fines;
    double *soilT;
    struct Lds {
        T stone[NP];
        T a[NP], b[NP];
    };
    struct Reg {
        T e, in, g, s, dp;
        double sT;
    };
    using HReg = T;
    struct Fin {
        T dp;
        double sTn;
    };
    __device__ __forceinline__ Reg load(long i) const { return Reg{ero[i], inci[i], geli[i], stone[i], dps[i], soilT[i]}; }
    __device__ __forceinline__ HReg load_halo(long i) const { return stone[i]; }
    __device__ __forceinline__ Fin store(Lds &L, int p, const Reg &r, uint32_t *extra) const
    {
        L.stone[p] = r.s;
        double sT = r.sT;
        double sedthick = sT + (double)r.dp;
        T a = 0, b = 0;
        if (sedthick != 0.0) {
            double old = sT / sedthick;
            double input = 1 - old;
            if (sizeof(T) == 8) { a = (T)(old * (double)r.s); b = (T)input; }
            else { a = (T)old * r.s; b = (T)input; }
            *extra |= W_UPD;
        }
        L.a[p] = a;
        L.b[p] = b;
        double toterosion;
        if (sizeof(T) == 8) toterosion = (double)r.e + (double)r.in + (double)r.g;
        else toterosion = (double)(r.e + r.in + r.g);
        Fin f{r.dp, sT};
        if (toterosion >= sedthick) f.sTn = 0.0;
        if (sedthick > toterosion) f.sTn = sedthick - toterosion;
        return f;
    }
    __device__ __forceinline__ void store_halo(Lds &L, int p, const HReg &r) const { L.stone[p] = r; }
    // the reference sums all eight terms with zeros for absent donors; adding
    // only the donors' terms from +0 gives the same value up to the sign of a
    // zero sum, which the clamp below erases
    __device__ __forceinline__ void compute(Lds &L, int p, unsigned bits, bool upd) const
    {
        if (!upd) return;
        const T one = (T)1;
        const T a = L.a[p], b = L.b[p];
        // all eight neighbours in flight at once (the padded layout keeps the
        // reads inside the array), then the donors' terms in order
        T v8[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v8[k] = L.stone[p + poff<TW>(cpull(k))];
        T propStonefrom = (T)0;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (bits & (1u << k)) propStonefrom = propStonefrom + one * v8[k];
        T st = a + b * propStonefrom;
        if (st < (T)0.01) st = (T)0.01;
        else if (st > (T)99.9) st = (T)99.9;
        L.stone[p] = st;
    }
    __device__ __forceinline__ void finish(const Lds &L, int p, long i, const Fin &f, bool upd) const
    {
        depo[i] = f.dp;
        if (upd) {
            T st = L.stone[p];
            stone[i] = st;
            fines[i] = (T)100 - st;
        }
        soilT[i] = f.sTn;
    }
};

// ------------------------------------------------------------ prep kernels

// Both prep kernels read each cell's 3x3 neighbourhood.  A 64 x 16 block of
// cells stages its 66 x 18 region (rows coalesced) in LDS once; off-grid
// positions are staged as nodata, which is how every rule treats them.
#define PREP_RG 4   // rows per thread: the block covers 64 x (4 * PREP_RG) cells
#define PREP_RW 66
#define PREP_RH (4 * PREP_RG + 2)

// donor bits in pull order; status: FA computes every mask cell, nodata and
// off-grid neighbours are complete from the start and contribute nothing
__global__ void __launch_bounds__(256) k_fa_prep(int W, int H, const uint8_t *__restrict__ mask,
                                                 const uint8_t *__restrict__ fd, uint8_t *__restrict__ dm,
                                                 int *__restrict__ st)
{
    __shared__ uint8_t smk[PREP_RH * PREP_RW], sfd[PREP_RH * PREP_RW];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * (4 * PREP_RG);
    constexpr int SN = PREP_RH * PREP_RW, SPER = (SN + 255) / 256;
    uint32_t vm[SPER], vf[SPER];   // 32-bit: byte registers would be packed, with a wait per load
#pragma unroll
    for (int j = 0; j < SPER; j++) {
        const int q = tid + j * 256;
        const int gy = y0 - 1 + q / PREP_RW, gx = x0 - 1 + q % PREP_RW;
        const bool in = q < SN && gx >= 0 && gx < W && gy >= 0 && gy < H;
        const long n = (long)gy * W + gx;
        vm[j] = in ? mask[n] : 0;
        vf[j] = in ? fd[n] : 0;
    }
#pragma unroll
    for (int j = 0; j < SPER; j++) {
        const int q = tid + j * 256;
        if (q < SN) { smk[q] = vm[j]; sfd[q] = vf[j]; }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < PREP_RG; g++) {
        const int lr = g * 4 + threadIdx.y;
        const int icol = x0 + threadIdx.x, irow = y0 + lr;
        if (icol >= W || irow >= H) continue;
        const long i = (long)irow * W + icol;
        const int p = (lr + 1) * PREP_RW + threadIdx.x + 1;
        unsigned bits = 0;
        if (smk[p]) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int d = cpull(k);
                const int pn = p + cdy(d) * PREP_RW + cdx(d);
                if (smk[pn] && (sfd[pn] & donor_code(d))) bits |= 1u << k;
            }
        }
        dm[i] = (uint8_t)bits;
        st[i] = smk[p] ? DAG_PENDING : 0;
    }
}

// MFDsedbud never computes row 0 / column 0 (mfd_accum.cu:42) and nodata
// cells are never ok (mfd_accum.cu:390-394); nodata carries fd = -9999, whose
// low byte sets E, W, NW, N, NE, so a nodata (or off-grid) neighbour at
// E, SE, S, SW or W is a donor that never completes (sediment blocking).  Also the
// deposition of every cell (mfd_accum.cu:63-200): the donors' erosion,
// incision and gelifluction terms in the order E..NE plus the cell's own, each
// weighted by its deposition fraction; it reads only inputs of the pass.
template <typename T>
__global__ void __launch_bounds__(256) k_sed_prep(int W, int H, const uint8_t *__restrict__ mask,
                                                  const uint8_t *__restrict__ fd, const uint8_t *__restrict__ demid,
                                                  const DemDesc *__restrict__ dd, uint8_t *__restrict__ dm,
                                                  int *__restrict__ st, const T *__restrict__ ero,
                                                  const T *__restrict__ inci, const T *__restrict__ geli, double ddk,
                                                  double dck, double dgk, T *__restrict__ dps,
                                                  const double *__restrict__ soilT, const T *__restrict__ stone,
                                                  int *bad)
{
    __shared__ uint8_t smk[PREP_RH * PREP_RW], sfd[PREP_RH * PREP_RW];
    __shared__ T se[PREP_RH * PREP_RW], si[PREP_RH * PREP_RW], sg[PREP_RH * PREP_RW];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * (4 * PREP_RG);
    // every load of the region in flight before the first LDS store
    constexpr int SN = PREP_RH * PREP_RW, SPER = (SN + 255) / 256;
    uint32_t vm[SPER], vf[SPER];   // 32-bit: byte registers would be packed, with a wait per load
    T ve[SPER], vi[SPER], vg[SPER];
#pragma unroll
    for (int j = 0; j < SPER; j++) {
        const int q = tid + j * 256;
        const int gy = y0 - 1 + q / PREP_RW, gx = x0 - 1 + q % PREP_RW;
        const bool in = q < SN && gx >= 0 && gx < W && gy >= 0 && gy < H;
        const long n = (long)gy * W + gx;
        vm[j] = in ? mask[n] : 0;
        vf[j] = in ? fd[n] : 0;
        ve[j] = in ? ero[n] : (T)0;
        vi[j] = in ? inci[n] : (T)0;
        vg[j] = in ? geli[n] : (T)0;
    }
#pragma unroll
    for (int j = 0; j < SPER; j++) {
        const int q = tid + j * 256;
        if (q >= SN) continue;
        smk[q] = vm[j];
        sfd[q] = vf[j];
        se[q] = vm[j] ? ve[j] : (T)0;
        si[q] = vm[j] ? vi[j] : (T)0;
        sg[q] = vm[j] ? vg[j] : (T)0;
    }
    __syncthreads();
    const T one = (T)1;
    const T kd = (T)(1 - ddk), kc = (T)(1 - dck), kg = (T)(1 - dgk);
#pragma unroll
    for (int g = 0; g < PREP_RG; g++) {
        const int lr = g * 4 + threadIdx.y;
        const int icol = x0 + threadIdx.x, irow = y0 + lr;
        if (icol >= W || irow >= H) continue;
        const long i = (long)irow * W + icol;
        const int p = (lr + 1) * PREP_RW + threadIdx.x + 1;
        unsigned bits = 0;
        int s = 0;
        if (smk[p]) {
            // row 0 or column 0 of the cell's DEM: the north or west
            // neighbour is outside the DEM's rectangle (off the grid, or in
            // another DEM's rectangle or none)
            bool dead;
            if (demid) {
                const uint8_t me = demid[i];
                const uint8_t nn = irow > 0 ? demid[i - W] : DEM_NONE, ww = icol > 0 ? demid[i - 1] : DEM_NONE;
                dead = nn != me || ww != me;
            } else {
                dead = irow == 0 || icol == 0;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int d = cpull(k);
                const int pn = p + cdy(d) * PREP_RW + cdx(d);
                if (!smk[pn]) {
                    if ((-9999) & donor_code(d)) dead = true;
                    continue;
                }
                if (sfd[pn] & donor_code(d)) bits |= 1u << k;
            }
            s = dead ? -1 : DAG_PENDING;
            T accum_d = 0, accum_c = 0, accum_g = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (!(bits & (1u << k))) continue;
                const int pn = p + cdy(cpull(k)) * PREP_RW + cdx(cpull(k));
                accum_d += one * (se[pn] * kd);
                accum_c += one * (si[pn] * kc);
                accum_g += one * (sg[pn] * kg);
            }
            const T dp = ((accum_d + se[p]) * (T)ddk) + ((accum_c + si[p]) * (T)dck) + ((accum_g + sg[p]) * (T)dgk);
            dps[i] = dp;
            // the stone-mixing weights of a cell that can compute, as the
            // engine's cell body forms them: the contraction of the FP32 form
            // needs them finite (bad is set otherwise)
            if (bad && !dead) {
                const double sT = soilT[i];
                const T sv = stone[i];
                const double sedthick = sT + (double)dp;
                T a = sv, b = 0;
                if (sedthick != 0.0) {
                    const double old = sT / sedthick;
                    a = (T)old * sv;
                    b = (T)(1 - old);
                }
                if (!__builtin_isfinite(a) || !__builtin_isfinite(b)) *bad = 1;
            }
        }
        dm[i] = (uint8_t)bits;
        st[i] = s;
    }
}

// ------------------------------------------------------------ tiled rounds

// Round bookkeeping: the woken tiles of round r are list L[r % 3] with count
// cnt[r % 3]; round r appends the tiles it wakes to L[(r+1) % 3] (each once:
// wake[tile] holds the last round it was queued for) and clears cnt[(r+2) % 3],
// last read by round r-1.  flag records progress for the host.
struct DagCtl {
    int cnt[3];
    int pad0[29];
    int flag;
    int pad1[31];
};


// threads of a workgroup that walk chains (whole waves)
#ifndef DAG_WALKERS
#define DAG_WALKERS 64
#endif

// Inside a tile, donors are read from LDS only (values, states, counters);
// HBM copies are read by other tiles in later launches.  So ordering a
// resolved cell's LDS writes before the counter update that can hand it to
// another thread needs only the LDS queue drained, not the global stores.

__device__ __forceinline__ void lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Appends a workgroup's woken tiles to the next round's list with one
// atomic (called by every thread of the workgroup, after a barrier).
#define DAG_QCAP 512
__device__ __forceinline__ void dag_flush(int *qbuf, int *qn, int *next, int *count)
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

// One round over the woken tiles.  Inside a tile every cell gets a counter of
// donors still unresolved in the tile; cells whose donors in other tiles are
// not all final from earlier launches are parked for a later round.  A thread
// resolving a cell decrements its receivers in the tile, and the thread that
// brings a receiver's counter to zero continues along it, so a dependency
// chain runs inside the tile at LDS speed with no barrier per link.  Extra
// ready receivers (MFD) go to an LDS spill list processed after a barrier.
// A chain leaving the tile only marks the neighbour tile in an LDS mask; the
// tile wakes each marked neighbour once, after its walks.  A cell with a dead
// donor dies (sediment blocking).  Work of a cell that does not depend on
// its donors' results (the cell body's pre-phase) is done for all cells of
// the tile in parallel before the walks.
// SFD: every cell has at most one receiver (single flow directions), so the
// walk has no list of extra ready receivers
template <int TW, int TH, class Cell, bool SFD>
__global__ void __launch_bounds__(256, Cell::kWaves) k_dag_round(int W, int H, int round, Cell cell, const uint8_t *__restrict__ dm,
                                                   const uint8_t *__restrict__ fdv, int *__restrict__ st,
                                                   const int *__restrict__ act,
                                                   int *__restrict__ next, int *__restrict__ wake, DagCtl *ctl)
{
    constexpr int NC = TW * TH;
    constexpr int PER = NC / 256;
    constexpr int NH = 2 * (TW + 2) + 2 * TH;
    constexpr int HPER = (NH + 255) / 256;
    constexpr int NP = (TW + 2) * (TH + 2);
    static_assert(NC % 256 == 0, "tile shape");
    // packed words (W_*) in the padded layout; a ring cell's word has state 3,
    // a counter no decrement brings to zero, and in bits 8-16 the one-hot 3x3
    // position of the neighbour tile it belongs to, so a hand-over to it
    // marks that tile for waking without a bounds test
    __shared__ uint32_t wd[NP];
    // states at the start of the round, padded layout: tile cells ST_PENDING,
    // ST_DONE, ST_DEAD; ring cells the same, with a donor in another tile
    // that is not final from an earlier launch as ST_RING_WAIT
    __shared__ uint8_t pst[(TW + 2) * (TH + 2)];
    __shared__ typename Cell::Lds lds;
    __shared__ uint16_t spill[2][NC];
    __shared__ int nspill[2];
    __shared__ int take;
    __shared__ int any_pending;
    __shared__ unsigned exits;
    __shared__ int qbuf[DAG_QCAP];   // tiles woken for the next round, not yet appended
    __shared__ int qn;
    bool wg_prog = false;            // thread 0: some visit of this workgroup made progress
    if (threadIdx.x == 0) qn = 0;
    // the list count and this workgroup's first entry are read together (the
    // grid never exceeds the list's capacity; an entry past the count is
    // never used)
    const int n = ctl->cnt[round % 3];
    const int first = act[blockIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->cnt[(round + 2) % 3] = 0;
    const int ntx = (W + TW - 1) / TW, nty = (H + TH - 1) / TH;
    // the next visit's tile is read during this one
    int tile_next = first;
    for (int q0 = blockIdx.x; q0 < n; q0 += gridDim.x) {
        const int tile = tile_next;
        tile_next = q0 + (int)gridDim.x < n ? act[q0 + gridDim.x] : 0;
        const int tx0 = (tile % ntx) * TW, ty0 = (tile / ntx) * TH;
        __syncthreads();
        if (threadIdx.x == 0) { any_pending = 0; nspill[0] = 0; nspill[1] = 0; take = 0; exits = 0; }
        // 1. load states and cached values of the tile and of its ring (all
        //    loads in flight before the first wait); a ring cell counts as
        //    final only if it completed in an earlier launch
        int vs[PER];
        uint32_t vb[PER], vf[PER];   // 32-bit: byte registers would be packed, with a wait per load
        typename Cell::Reg vr[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int li = threadIdx.x + q * 256;
            int r = ty0 + li / TW, c = tx0 + li % TW;
            vs[q] = 0;   // off-grid: done before the first round, never a donor or receiver
            vb[q] = 0;
            vf[q] = 0;
            vr[q] = typename Cell::Reg{};
            if (r < H && c < W) {
                long i = (long)r * W + c;
                vs[q] = st[i];
                vb[q] = dm[i];
                vf[q] = fdv[i];
                vr[q] = cell.load(i);
            }
        }
        int hv[HPER];
        typename Cell::HReg hr[HPER];
#pragma unroll
        for (int q = 0; q < HPER; q++) {
            int hh = threadIdx.x + q * 256;
            hv[q] = 0;   // off-grid ring cells are never donors
            hr[q] = typename Cell::HReg{};
            if (hh < NH) {
                int nr, nc;
                halo_pos<TW, TH>(hh, &nr, &nc);
                int r = ty0 + nr, c = tx0 + nc;
                if (r >= 0 && r < H && c >= 0 && c < W) {
                    long i = (long)r * W + c;
                    hv[q] = st[i];
                    hr[q] = cell.load_halo(i);
                }
            }
        }
        typename Cell::Fin fin[PER];
        uint32_t xw[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int li = threadIdx.x + q * 256;
            int v = vs[q];
            uint32_t sv = (v == DAG_PENDING) ? ST_PENDING : (v >= 0 ? ST_DONE : ST_DEAD);
            xw[q] = 0;
            const int p = pidx<TW>(li / TW, li % TW);
            fin[q] = cell.store(lds, p, vr[q], &xw[q]);
            wd[p] = (sv << W_STSH) | ((uint32_t)vf[q] << 16) | ((uint32_t)vb[q] << 8) | W_NEVER;
            pst[p] = (uint8_t)sv;
        }
#pragma unroll
        for (int q = 0; q < HPER; q++) {
            int hh = threadIdx.x + q * 256;
            if (hh < NH) {
                int v = hv[q];
                uint8_t hs;
                if (v == DAG_PENDING) hs = ST_RING_WAIT;
                else if (v >= 0) hs = (v < round) ? ST_DONE : ST_RING_WAIT;
                else hs = ((-v - 1) < round) ? ST_DEAD : ST_RING_WAIT;
                int nr, nc;
                halo_pos<TW, TH>(hh, &nr, &nc);
                const int p = pidx<TW>(nr, nc);
                pst[p] = hs;
                cell.store_halo(lds, p, hr[q]);
                const int ty = nr < 0 ? 0 : (nr >= TH ? 2 : 1), tx = nc < 0 ? 0 : (nc >= TW ? 2 : 1);
                wd[p] = (3u << W_STSH) | (1u << (8 + ty * 3 + tx)) | W_NEVER;
            }
        }
        __syncthreads();
        // 2. counters: donors unresolved in the tile; park cells waiting on
        //    other tiles
        uint32_t myw[PER];
        int seeds = 0;   // bit q: cell q of this thread starts a walk
        int pend0 = 0;   // bit q: cell q was pending when the round started
#pragma unroll
        for (int q = 0; q < PER; q++) {
            int li = threadIdx.x + q * 256;
            myw[q] = 0;
            if (vs[q] != DAG_PENDING) continue;
            pend0 |= 1 << q;
            const int p = pidx<TW>(li / TW, li % TW);
            unsigned bits = vb[q];
            // all eight states in flight at once (the padded layout keeps
            // them inside the array), then a one-hot code per donor
            uint32_t s8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) s8[k] = pst[p + poff<TW>(cpull(k))];
            int nin = 0;
            uint32_t seen = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t oh = ((bits >> k) & 1u) << s8[k];
                nin += (int)(oh & (1u << ST_PENDING));
                seen |= oh;
            }
            const bool dead = (seen >> ST_DEAD) & 1u, parked = (seen >> ST_RING_WAIT) & 1u;
            uint32_t base = ((uint32_t)ST_PENDING << W_STSH) | ((uint32_t)vf[q] << 16) | (bits << 8) | xw[q];
            if (dead) { myw[q] = base | W_NEVER | W_DEAD1; seeds |= 1 << q; }
            else if (!parked) { myw[q] = base | (uint32_t)nin; if (nin == 0) seeds |= 1 << q; }
            else myw[q] = base | W_NEVER;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int li = threadIdx.x + q * 256;
            const int p = pidx<TW>(li / TW, li % TW);
            if ((pend0 >> q) & 1) wd[p] = myw[q];
            // the seeds form the work list of the first pass (list 1)
            const bool sd = (seeds >> q) & 1;
            const unsigned long long m = __ballot(sd);
            int base = 0;
            if ((threadIdx.x & 63) == 0 && m) base = atomicAdd(&nspill[1], __popcll(m));
            base = __shfl(base, 0);
            if (sd) spill[1][base + __popcll(m & ((1ULL << (threadIdx.x & 63)) - 1))] = (uint16_t)p;
        }
        __syncthreads();
        // 3. walks.  Pass p works through list (p - 1) & 1 and collects the
        //    extra ready receivers (MFD) in list p & 1.  Only the first
        //    DAG_WALKERS threads walk: each takes the next list entry whenever
        //    its chain ends, so its lanes stay busy while the list holds work,
        //    and the other waves leave the SIMDs to other workgroups.
        unsigned my_exits = 0;
        for (int pass = 0;; pass++) {
            const int nwork = nspill[(pass - 1) & 1];
            if (nwork == 0) break;
            if (threadIdx.x < DAG_WALKERS) {
                const uint16_t *lst = spill[(pass - 1) & 1];
                int cur = -1;
                uint32_t word = 0;
                for (;;) {
                    if (cur < 0) {
                        const int idx = atomicAdd(&take, 1);
                        if (idx >= nwork) break;
                        cur = lst[idx];
                        word = wd[cur];   // only dead-count and counter bits of a dead seed can have moved
                    }
                    const bool dead = Cell::kCanDie && ((word >> 24) & 0xFu) != 0;
                    // receivers: the set bits of the fd code (bit k is direction
                    // (k + 2) & 7); the first one's address is ready before the
                    // wait below.  A cell with no receiver adds 0 to its own
                    // word instead, whose counter is not 1 once it is resolved
                    // (0, or what is left of W_NEVER), so nothing is handed over
                    unsigned f = (word >> 16) & 0xFFu;
                    int rp0 = f ? cur + koff<TW>(__ffs(f) - 1) : cur;
                    uint32_t ns = ST_DEAD;
                    if (!dead) {
                        cell.compute(lds, cur, (word >> 8) & 0xFFu, (word & W_UPD) != 0);
                        ns = ST_DONE;
                    }
                    wd[cur] = (word & ~(3u << W_STSH)) | (ns << W_STSH);
                    asm volatile("" : "+v"(rp0));
                    lds_order();
                    // unsigned adds: -1, or -1 plus one dead donor (the counter
                    // of a receiver is at least 1 here, so no borrow leaves it)
                    const uint32_t delta = dead ? (W_DEAD1 - 1u) : 0xFFFFFFFFu;
                    int nxt;
                    uint32_t nword;
                    {
                        // the first (under single flow directions the only)
                        // receiver, without branches: the walk continues with
                        // it when the add hands it over
                        const uint32_t d0 = f ? delta : 0u;
                        f &= f - 1;
                        const uint32_t old = atomicAdd(&wd[rp0], d0);
                        // a ring word carries the neighbour tile's position
                        my_exits |= (((old >> W_STSH) & 3u) == 3u) ? ((old >> 8) & 0x1FFu) : 0u;
                        nxt = ((old & W_CNT) == 1u) ? rp0 : -1;
                        nword = old + d0;
                    }
                    while (!SFD && f) {   // further receivers (multiple flow directions)
                        const int k = __ffs(f) - 1;
                        f &= f - 1;
                        const int rp = cur + koff<TW>(k);
                        const uint32_t old = atomicAdd(&wd[rp], delta);
                        my_exits |= (((old >> W_STSH) & 3u) == 3u) ? ((old >> 8) & 0x1FFu) : 0u;
                        if ((old & W_CNT) == 1u) {
                            if (nxt < 0) { nxt = rp; nword = old + delta; }
                            else spill[pass & 1][atomicAdd(&nspill[pass & 1], 1)] = (uint16_t)rp;
                        }
                    }
                    cur = nxt;
                    word = nword;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) { nspill[(pass + 1) & 1] = 0; take = 0; }
            __syncthreads();
        }
        if (my_exits) atomicOr(&exits, my_exits);
        __syncthreads();
        // wake each neighbour the chains flow into, once per round: gathered
        // in LDS and appended to the next list with one atomic per full
        // buffer (workgroups adding to one counter serialise).  Before the
        // HBM writes below: the wait for the atomic's returned value would
        // otherwise also wait for those stores (the vector memory counter
        // counts stores as well), and the next visit's loads overlap them
        if (threadIdx.x < 9 && ((exits >> threadIdx.x) & 1)) {
            int ty = tile / ntx + (int)threadIdx.x / 3 - 1, tx = tile % ntx + (int)threadIdx.x % 3 - 1;
            if (ty >= 0 && ty < nty && tx >= 0 && tx < ntx) {
                const int t2 = ty * ntx + tx;
                // not queued already, nor finished
                if (atomicMax(&wake[t2], round + 1) < round + 1) qbuf[atomicAdd(&qn, 1)] = t2;
            }
        }
        // 4. HBM writes of the cells resolved in this round, and bookkeeping
        bool pend = false;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            if (!((pend0 >> q) & 1)) continue;
            int li = threadIdx.x + q * 256;
            const int p = pidx<TW>(li / TW, li % TW);
            uint32_t wv = wd[p];
            uint32_t s2 = (wv >> W_STSH) & 3u;
            if (s2 == ST_PENDING) { pend = true; continue; }
            long i = (long)(ty0 + li / TW) * W + (tx0 + li % TW);
            if (s2 == ST_DONE) {
                cell.finish(lds, p, i, fin[q], (wv & W_UPD) != 0);
                st[i] = round;
            } else {
                st[i] = -(round + 1);
            }
        }
        if (pend) any_pending = 1;
        int prog = __syncthreads_or(seeds != 0);
        if (threadIdx.x == 0) {
            if (prog) wg_prog = true;
            if (!any_pending) atomicExch(&wake[tile], DAG_TILE_DONE);
        }
        __syncthreads();
        if (qn > DAG_QCAP - 9) dag_flush(qbuf, &qn, next, &ctl->cnt[(round + 1) % 3]);
    }
    __syncthreads();
    if (qn > 0) dag_flush(qbuf, &qn, next, &ctl->cnt[(round + 1) % 3]);
    if (threadIdx.x == 0 && wg_prog) ctl->flag = 1;
}

__global__ void k_count_pending(long N, const uint8_t *mask, const int *st, int *out)
{
    __shared__ int sh[256];
    int acc = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        const uint8_t m = mask[i];   // both loads in flight together
        const int s = st[i];
        acc += (m && (s == DAG_PENDING || s < 0)) ? 1 : 0;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && sh[0]) atomicAdd(out, sh[0]);
}

static int grid1d_dag(long n)
{
    long g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}

__global__ void k_iota(int n, int *a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// first_rounds > 0: only that many rounds (the first visits every tile),
// with no wait for the device; *uncomputed is not set
template <int TW, int TH, class Cell, bool SFD>
static int run_rounds(lemhip_ctx *c, Cell cell, int *uncomputed, int prof_slot, int first_rounds = 0)
{
    int ntiles = ((c->W + TW - 1) / TW) * ((c->H + TH - 1) / TH);
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    int *L[3] = {c->dag_list[0], c->dag_list[1], c->dag_list[2]};
    DagCtl *ctl = reinterpret_cast<DagCtl *>(c->dag_ctl);
    int r = c->dag_round;
    if (r > (1 << 29)) r = 0;   // status words hold round numbers
    // wake stamps start below every round number of the pass
    LEM_CHECK(c, hipMemsetAsync(c->dag_wake, 0, sizeof(int) * ntiles, c->s));
    // the first round visits every tile
    k_iota<<<(ntiles + 255) / 256, 256, 0, c->s>>>(ntiles, L[(r + 1) % 3]);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemsetAsync(ctl, 0, sizeof(DagCtl), c->s));
    c->hint[16] = ntiles;
    LEM_CHECK(c, hipMemcpyAsync(&ctl->cnt[(r + 1) % 3], &c->hint[16], sizeof(int), hipMemcpyHostToDevice, c->s));
    const int BATCH = c->dag_batch;
    // as many workgroups as the CUs hold at once (the first rounds visit every tile)
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_dag_round<TW, TH, Cell, SFD>, 256, 0) != hipSuccess || bpc < 1)
        bpc = 4;
    int G = bpc * c->ncu;
    if (G > ntiles) G = ntiles;
    if (first_rounds > 0) {
        prof_begin(c, prof_slot);
        for (int q = 0; q < first_rounds; q++) {
            ++r;
            k_dag_round<TW, TH, Cell, SFD><<<G, 256, 0, c->s>>>(c->W, c->H, r, cell, c->dmask, c->fd, st, L[r % 3],
                                                           L[(r + 1) % 3], c->dag_wake, ctl);
        }
        prof_end(c, prof_slot, first_rounds);
        LEM_LAUNCHED(c);
        c->dag_rounds_last = first_rounds;
        c->dag_round = r;
        return LEMHIP_OK;
    }
    long guard = 0;
    for (;;) {
        LEM_CHECK(c, hipMemsetAsync(&ctl->flag, 0, sizeof(int), c->s));
        prof_begin(c, prof_slot);
        for (int q = 0; q < BATCH; q++) {
            ++r;
            k_dag_round<TW, TH, Cell, SFD><<<G, 256, 0, c->s>>>(c->W, c->H, r, cell, c->dmask, c->fd, st, L[r % 3],
                                                           L[(r + 1) % 3], c->dag_wake, ctl);
        }
        prof_end(c, prof_slot, BATCH);
        LEM_LAUNCHED(c);
        LEM_CHECK(c, hipMemcpyAsync(&c->hint[20], &ctl->flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        if (!c->hint[20]) break;
        guard += 1;  // batches that made progress; each completes at least one cell
        if (guard > c->N + 2) { c->err = "dependency rounds did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    c->dag_rounds_last = r - c->dag_round;
    c->dag_round = r;
    LEM_CHECK(c, hipMemsetAsync(&c->scal->flag, 0, sizeof(int), c->s));
    k_count_pending<<<grid1d_dag(c->N), 256, 0, c->s>>>(c->N, c->mask, st, &c->scal->flag);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[3], &c->scal->flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    *uncomputed = c->hint[3];
    return LEMHIP_OK;
}

// Tile shapes of the dependency engine (the autotuner picks one per context).
// Every shape gives identical results: the per-cell arithmetic and its order
// do not depend on the tiling, only the number of rounds does.
#define DAG_SHAPES(X) X(0, 64, 16) X(1, 32, 32) X(2, 16, 64) X(3, 128, 8) X(4, 64, 32) X(5, 32, 64)

template <typename T, int TW, int TH>
static int fa_shape(lemhip_ctx *c)
{
    FaCell<T, TW, TH> cell{reinterpret_cast<const T *>(c->w), reinterpret_cast<typename FaStore<T>::type *>(c->fa)};
    if (c->p.flowdirectiontype == 0) return run_rounds<TW, TH, decltype(cell), true>(c, cell, &c->h_fa_unc, PROF_FA_K);
    return run_rounds<TW, TH, decltype(cell), false>(c, cell, &c->h_fa_unc, PROF_FA_K);
}

template <typename T, int TW, int TH>
static int sed_shape(lemhip_ctx *c)
{
    SedCell<T, TW, TH> cell{reinterpret_cast<const T *>(c->ero), reinterpret_cast<const T *>(c->inci),
                            reinterpret_cast<const T *>(c->geli), reinterpret_cast<const T *>(c->z2),
                            reinterpret_cast<T *>(c->depo), reinterpret_cast<T *>(c->stone),
                            reinterpret_cast<T *>(c->fines), c->soilT};
    if (c->p.flowdirectiontype == 0) return run_rounds<TW, TH, decltype(cell), true>(c, cell, &c->h_sed_unc, PROF_SED_K);
    return run_rounds<TW, TH, decltype(cell), false>(c, cell, &c->h_sed_unc, PROF_SED_K);
}

template <typename T>
int lem_fa_dag(lemhip_ctx *c)
{
    dim3 blk(64, 4), grd((c->W + 63) / 64, (c->H + 4 * PREP_RG - 1) / (4 * PREP_RG));
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    k_fa_prep<<<grd, blk, 0, c->s>>>(c->W, c->H, c->mask, c->fd, c->dmask, st);
    LEM_LAUNCHED(c);
    switch (c->dag_shape) {
#define X(id, tw, th) case id: return fa_shape<T, tw, th>(c);
    DAG_SHAPES(X)
#undef X
    default: c->err = "unknown tile shape"; return LEMHIP_E_ARG;
    }
}

// donor bits (dmask), states (scan_tmp) and deposition (z2 as T) of the
// sediment pass; bad (optional) is set when a computable cell's
// stone-mixing weights are not finite
template <typename T>
int lem_sed_prep(lemhip_ctx *c, int *bad)
{
    dim3 blk(64, 4), grd((c->W + 63) / 64, (c->H + 4 * PREP_RG - 1) / (4 * PREP_RG));
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    double ddk = c->p.ddk, dck = c->p.dck, dgk = c->p.dgk;
    if (ddk > 0.95) ddk = 0.95;   // mfd_accum.cu:406-407
    if (dck > 0.9) dck = 0.9;
    // deposition goes to the raised-surface buffer of the direction phase,
    // free until the next step
    k_sed_prep<T><<<grd, blk, 0, c->s>>>(c->W, c->H, c->mask, c->fd, c->demid, c->ddesc, c->dmask, st,
                                         (const T *)c->ero, (const T *)c->inci, (const T *)c->geli, ddk, dck, dgk,
                                         (T *)c->z2, c->soilT, (const T *)c->stone, bad);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

template <typename T>
int lem_sed_dag(lemhip_ctx *c)
{
    c->h_sed_open = 0;
    LEM_TRY(lem_sed_prep<T>(c, nullptr));
    switch (c->dag_shape) {
#define X(id, tw, th) case id: return sed_shape<T, tw, th>(c);
    DAG_SHAPES(X)
#undef X
    default: c->err = "unknown tile shape"; return LEMHIP_E_ARG;
    }
}

// The first rounds of the engine's sediment pass in the FP32 form under
// single flow directions, after lem_sed_prep.  Round 1 visits every 32 x 32
// tile and completes each cell whose upstream lies inside its tile (on the
// fBm workloads about 82% of the cells); each later round visits the tiles
// woken by the last one's exits and completes the cells whose upstream
// crosses into them only from cells completed before (after round 2 about
// 91%).  The rest stays pending for the contraction (lem_sed32.hip).  The tile shape is fixed, so
// which cells complete here depends on the cell's position modulo 32 only.
int lem_sed_first(lemhip_ctx *c, int rounds)
{
    SedCell<float, 32, 32> cell{reinterpret_cast<const float *>(c->ero), reinterpret_cast<const float *>(c->inci),
                                reinterpret_cast<const float *>(c->geli), reinterpret_cast<const float *>(c->z2),
                                reinterpret_cast<float *>(c->depo), reinterpret_cast<float *>(c->stone),
                                reinterpret_cast<float *>(c->fines), c->soilT};
    return run_rounds<32, 32, decltype(cell), true>(c, cell, nullptr, PROF_SED_K, rounds);
}

// Times flow accumulation with every tile shape, then the best shape with
// several launch batches, on the current directions and runoff weights (the
// pass is idempotent), and keeps the fastest.  Results do not depend on the
// choice.
int lem_tune_dag(lemhip_ctx *c)
{
    hipEvent_t e0, e1;
    LEM_CHECK(c, hipEventCreate(&e0));
    LEM_CHECK(c, hipEventCreate(&e1));
    auto time_one = [&](float *ms) -> int {
        LEM_CHECK(c, hipEventRecord(e0, c->s));
        LEM_TRY(c->prec == LEMHIP_FP64 ? lem_fa_dag<double>(c) : lem_fa_dag<float>(c));
        LEM_CHECK(c, hipEventRecord(e1, c->s));
        LEM_CHECK(c, hipEventSynchronize(e1));
        LEM_CHECK(c, hipEventElapsedTime(ms, e0, e1));
        return LEMHIP_OK;
    };
    float best = 1e30f;
    int bs = c->dag_shape;
    c->dag_batch = 16;
    for (int k = 0; k < lem_dag_shape_count(); k++) {
        c->dag_shape = k;
        float ms = 0.f;
        LEM_TRY(time_one(&ms));
        c->tune_ms[k] = ms;
        if (ms < best) { best = ms; bs = k; }
    }
    c->dag_shape = bs;
    static const int BATCHES[3] = {4, 8, 32};
    int bb = 16;
    for (int q = 0; q < 3; q++) {
        c->dag_batch = BATCHES[q];
        float ms = 0.f;
        LEM_TRY(time_one(&ms));
        c->tune_ms[11 + q] = ms;
        if (ms < best) { best = ms; bb = BATCHES[q]; }
    }
    c->dag_batch = bb;
    c->tune_dag_pending = false;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return LEMHIP_OK;
}

// The same choice where the engine serves only sediment routing (FP32 form
// under single flow directions, flow accumulation in integers): sediment
// routing is timed with every tile shape and then with several launch
// batches.  The pass updates stone, soil thickness, deposition and fines in
// place, so they are saved first and restored before every timed run and
// after the last (the step then routes sediment from the state it had).
int lem_tune_sed(lemhip_ctx *c)
{
    const long N = c->N;
    const size_t fb = (size_t)N * sizeof(float), db = (size_t)N * sizeof(double);
    char *save = nullptr;
    LEM_CHECK(c, hipMalloc((void **)&save, 3 * fb + db));
    struct Field { void *p; size_t bytes; } fields[4] = {{c->stone, fb}, {c->depo, fb}, {c->fines, fb}, {c->soilT, db}};
    auto copy_all = [&](bool out) -> int {
        size_t off = 0;
        for (const Field &f : fields) {
            void *dst = out ? (void *)(save + off) : f.p;
            const void *src = out ? f.p : (const void *)(save + off);
            LEM_CHECK(c, hipMemcpyAsync(dst, src, f.bytes, hipMemcpyDeviceToDevice, c->s));
            off += f.bytes;
        }
        return LEMHIP_OK;
    };
    hipEvent_t e0, e1;
    LEM_CHECK(c, hipEventCreate(&e0));
    LEM_CHECK(c, hipEventCreate(&e1));
    auto time_one = [&](float *ms) -> int {
        LEM_TRY(copy_all(false));
        LEM_CHECK(c, hipEventRecord(e0, c->s));
        LEM_TRY(lem_sed_dag<float>(c));
        LEM_CHECK(c, hipEventRecord(e1, c->s));
        LEM_CHECK(c, hipEventSynchronize(e1));
        LEM_CHECK(c, hipEventElapsedTime(ms, e0, e1));
        return LEMHIP_OK;
    };
    LEM_TRY(copy_all(true));
    float best = 1e30f;
    int bs = c->dag_shape;
    c->dag_batch = 16;
    for (int k = 0; k < lem_dag_shape_count(); k++) {
        c->dag_shape = k;
        float ms = 0.f;
        LEM_TRY(time_one(&ms));
        c->tune_ms[k] = ms;
        if (ms < best) { best = ms; bs = k; }
    }
    c->dag_shape = bs;
    static const int BATCHES[3] = {4, 8, 32};
    int bb = 16;
    for (int q = 0; q < 3; q++) {
        c->dag_batch = BATCHES[q];
        float ms = 0.f;
        LEM_TRY(time_one(&ms));
        c->tune_ms[11 + q] = ms;
        if (ms < best) { best = ms; bb = BATCHES[q]; }
    }
    c->dag_batch = bb;
    LEM_TRY(copy_all(false));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    c->tune_dag_pending = false;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(save);
    return LEMHIP_OK;
}

int lem_dag_shape_count() { return 6; }
void lem_dag_shape_dims(int id, int *tw, int *th)
{
    switch (id) {
#define X(i, a, b) case i: *tw = a; *th = b; return;
    DAG_SHAPES(X)
#undef X
    default: *tw = 0; *th = 0;
    }
}

template int lem_fa_dag<double>(lemhip_ctx *);
template int lem_fa_dag<float>(lemhip_ctx *);
template int lem_sed_dag<double>(lemhip_ctx *);
template int lem_sed_dag<float>(lemhip_ctx *);
template int lem_sed_prep<float>(lemhip_ctx *, int *);
