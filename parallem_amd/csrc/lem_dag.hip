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

#define DAG_PENDING 0x7FFFFFFF
// status word: DAG_PENDING, r >= 0 done in round r (0 = before the first
// round), -(r+1) dead (can never complete) since round r

template <typename T> struct FaStore;
template <> struct FaStore<double> { using type = double; };
template <> struct FaStore<float> { using type = float2; };

// float-float (double-word) arithmetic for the compensated FP32 form
__device__ __forceinline__ float2 ff_add(float2 a, float2 b)
{
    float s = a.x + b.x;
    float v = s - a.x;
    float e = (a.x - (s - v)) + (b.x - v);
    e = e + (a.y + b.y);
    float hi = s + e;
    float lo = e - (hi - s);
    return make_float2(hi, lo);
}

// ------------------------------------------------------------ cell bodies

template <typename T>
struct FaCell {
    using FS = typename FaStore<T>::type;
    const T *w;
    FS *fa;
    // fa = w + max(0, sum of donor fa) (mfd_simple.cu:71-168; prop = 1 for
    // every mask donor, runoffweight.cu:126-134, so fa*prop == fa)
    __device__ __forceinline__ void compute(long i, int W, unsigned bits) const;
};

template <>
__device__ __forceinline__ void FaCell<double>::compute(long i, int W, unsigned bits) const
{
    double accum = 0.;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (bits & (1u << k)) {
            int d = kPULL[k];
            long n = i + (long)kDY[d] * W + kDX[d];
            accum += fa[n] * 1.0;
        }
    }
    if (accum < 0.0) accum = 0.0;
    double v = w[i];
    v += accum;
    fa[i] = v;
}

template <>
__device__ __forceinline__ void FaCell<float>::compute(long i, int W, unsigned bits) const
{
    float2 accum = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (bits & (1u << k)) {
            int d = kPULL[k];
            long n = i + (long)kDY[d] * W + kDX[d];
            accum = ff_add(accum, fa[n]);
        }
    }
    if (accum.x < 0.0f) accum = make_float2(0.f, 0.f);
    fa[i] = ff_add(make_float2(w[i], 0.f), accum);
}

template <typename T>
struct SedCell {
    const T *ero, *inci, *geli;
    T *depo, *stone, *fines;
    double *soilT;
    double ddk, dck, dgk;
    // MFDsedbud body (mfd_accum.cu:63-252)
    __device__ __forceinline__ void compute(long i, int W, unsigned bits) const
    {
        T accum_d = 0, accum_c = 0, accum_g = 0;
        T from[8];
        const T one = (T)1;
        const T kd = (T)(1 - ddk), kc = (T)(1 - dck), kg = (T)(1 - dgk);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            from[k] = 0;
            if (bits & (1u << k)) {
                int d = kPULL[k];
                long n = i + (long)kDY[d] * W + kDX[d];
                accum_d += one * (ero[n] * kd);
                accum_c += one * (inci[n] * kc);
                accum_g += one * (geli[n] * kg);
                from[k] = one * stone[n];
            }
        }
        T dp = ((accum_d + ero[i]) * (T)ddk) + ((accum_c + inci[i]) * (T)dck) + ((accum_g + geli[i]) * (T)dgk);
        depo[i] = dp;
        double sT = soilT[i];
        double sedthick = sT + (double)dp;
        if (sedthick != 0.0) {
            double old = sT / sedthick;
            double input = 1 - old;
            T propStonefrom = (from[0] + from[1] + from[2] + from[3] + from[4] + from[5] + from[6] + from[7]);
            T st = (T)((old * (double)stone[i]) + (input * (double)propStonefrom));
            if (sizeof(T) == 4) st = (T)(((T)old * stone[i]) + ((T)input * propStonefrom));
            if (st < (T)0.01) st = (T)0.01;
            else if (st > (T)99.9) st = (T)99.9;
            stone[i] = st;
            fines[i] = (T)100 - st;
        }
        double toterosion = (double)ero[i] + (double)inci[i] + (double)geli[i];
        if (sizeof(T) == 4) toterosion = (double)(ero[i] + inci[i] + geli[i]);
        if (toterosion >= sedthick) soilT[i] = 0.0;
        if (sedthick > toterosion) soilT[i] = sedthick - toterosion;
    }
};

// ------------------------------------------------------------ prep kernels

// donor bits in pull order; status: FA computes every mask cell, nodata and
// off-grid neighbours are complete from the start and contribute nothing
__global__ void k_fa_prep(int W, int H, const uint8_t *__restrict__ mask, const uint8_t *__restrict__ fd,
                          uint8_t *__restrict__ dm, int *__restrict__ st)
{
    int icol = blockIdx.x * blockDim.x + threadIdx.x;
    int irow = blockIdx.y * blockDim.y + threadIdx.y;
    if (icol >= W || irow >= H) return;
    long i = (long)irow * W + icol;
    unsigned bits = 0;
    if (mask[i]) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int d = kPULL[k];
            int cx = icol + kDX[d], cy = irow + kDY[d];
            if (cx < 0 || cx >= W || cy < 0 || cy >= H) continue;
            long n = (long)cy * W + cx;
            if (mask[n] && (fd[n] & donor_code(d))) bits |= 1u << k;
        }
    }
    dm[i] = (uint8_t)bits;
    st[i] = mask[i] ? DAG_PENDING : 0;
}

// MFDsedbud never computes row 0 / column 0 (mfd_accum.cu:42) and nodata
// cells are never ok (mfd_accum.cu:390-394); nodata carries fd = -9999, whose
// low byte sets E, W, NW, N, NE, so a nodata (or, D11, off-grid) neighbour at
// E, SE, S, SW or W is a donor that never completes (D9)
__global__ void k_sed_prep(int W, int H, const uint8_t *__restrict__ mask, const uint8_t *__restrict__ fd,
                           uint8_t *__restrict__ dm, int *__restrict__ st)
{
    int icol = blockIdx.x * blockDim.x + threadIdx.x;
    int irow = blockIdx.y * blockDim.y + threadIdx.y;
    if (icol >= W || irow >= H) return;
    long i = (long)irow * W + icol;
    unsigned bits = 0;
    int s = 0;
    if (mask[i]) {
        bool dead = (irow == 0 || icol == 0);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int d = kPULL[k];
            int cx = icol + kDX[d], cy = irow + kDY[d];
            bool off = (cx < 0 || cx >= W || cy < 0 || cy >= H);
            long n = (long)cy * W + cx;
            if (off || !mask[n]) {
                if ((-9999) & donor_code(d)) dead = true;
                continue;
            }
            if (fd[n] & donor_code(d)) bits |= 1u << k;
        }
        s = dead ? -1 : DAG_PENDING;
    }
    dm[i] = (uint8_t)bits;
    st[i] = s;
}

// ------------------------------------------------------------ tiled rounds

template <int TW, int TH, class Cell>
__global__ void __launch_bounds__(256) k_dag_round(int W, int H, int round, Cell cell, const uint8_t *__restrict__ dm,
                                                   int *__restrict__ st, int *__restrict__ tile_done, int *progress)
{
    constexpr int NC = TW * TH;
    constexpr int PER = NC / 256;
    __shared__ uint8_t lst[NC];
    __shared__ int any_pending;
    int tile = blockIdx.x;
    if (tile_done[tile]) return;
    int ntx = (W + TW - 1) / TW;
    int tx0 = (tile % ntx) * TW, ty0 = (tile / ntx) * TH;
    if (threadIdx.x == 0) any_pending = 0;
    // local state: 0 pending, 1 done, 2 dead (off-grid slots count as done)
    uint8_t mybits[PER];
    long myidx[PER];
    bool myact[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int li = threadIdx.x + q * 256;
        int r = ty0 + li / TW, c = tx0 + li % TW;
        bool in = r < H && c < W;
        long i = (long)r * W + c;
        myidx[q] = i;
        uint8_t s = 1;
        uint8_t b = 0;
        if (in) {
            int v = st[i];
            s = (v == DAG_PENDING) ? 0 : (v >= 0 ? 1 : 2);
            b = dm[i];
        }
        lst[li] = s;
        mybits[q] = b;
        myact[q] = in && s == 0;
    }
    __syncthreads();
    bool progressed = false;
    for (;;) {
        uint8_t nxt[PER];
        bool any = false;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            nxt[q] = 0;
            if (!myact[q]) continue;
            int li = threadIdx.x + q * 256;
            int lr = li / TW, lc = li % TW;
            unsigned bits = mybits[q];
            bool ready = true, dead = false;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (!(bits & (1u << k))) continue;
                int d = kPULL[k];
                int nr = lr + kDY[d], nc = lc + kDX[d];
                uint8_t s;
                if (nr >= 0 && nr < TH && nc >= 0 && nc < TW) {
                    s = lst[nr * TW + nc];
                } else {
                    int v = st[myidx[q] + (long)kDY[d] * W + kDX[d]];
                    if (v == DAG_PENDING) s = 0;
                    else if (v >= 0) s = (v < round) ? 1 : 0;
                    else s = ((-v - 1) < round) ? 2 : 0;
                }
                if (s == 2) dead = true;
                if (s != 1) ready = false;
            }
            if (dead) nxt[q] = 2;
            else if (ready) {
                cell.compute(myidx[q], W, bits);
                nxt[q] = 1;
            }
            any |= (nxt[q] != 0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; q++) {
            if (nxt[q]) {
                int li = threadIdx.x + q * 256;
                lst[li] = nxt[q];
                st[myidx[q]] = (nxt[q] == 1) ? round : -(round + 1);
                myact[q] = false;
            }
        }
        progressed |= any;
        if (!__syncthreads_or(any)) break;
    }
    bool pend = false;
#pragma unroll
    for (int q = 0; q < PER; q++) pend |= myact[q];
    if (pend) any_pending = 1;
    if (__syncthreads_or(progressed) && threadIdx.x == 0) *progress = 1;
    if (threadIdx.x == 0 && !any_pending) tile_done[tile] = 1;
}

__global__ void k_count_pending(long N, const uint8_t *mask, const int *st, int *out)
{
    __shared__ int sh[256];
    int acc = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        acc += (mask[i] && (st[i] == DAG_PENDING || st[i] < 0)) ? 1 : 0;
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

template <int TW, int TH, class Cell>
static int run_rounds(lemhip_ctx *c, Cell cell, int *uncomputed)
{
    int ntiles = ((c->W + TW - 1) / TW) * ((c->H + TH - 1) / TH);
    LEM_CHECK(c, hipMemsetAsync(c->tile_flags, 0, sizeof(int) * ntiles, c->s));
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    for (int round = 1;; ++round) {
        LEM_CHECK(c, hipMemsetAsync(&c->scal->flag, 0, sizeof(int), c->s));
        k_dag_round<TW, TH, Cell><<<ntiles, 256, 0, c->s>>>(c->W, c->H, round, cell, c->dmask, st, c->tile_flags,
                                                            &c->scal->flag);
        LEM_LAUNCHED(c);
        LEM_CHECK(c, hipMemcpyAsync(&c->hint[2], &c->scal->flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        if (!c->hint[2]) break;
        if (round > c->N + 2) { c->err = "dependency rounds did not converge"; return LEMHIP_E_NOT_CONVERGED; }
    }
    LEM_CHECK(c, hipMemsetAsync(&c->scal->flag, 0, sizeof(int), c->s));
    k_count_pending<<<grid1d_dag(c->N), 256, 0, c->s>>>(c->N, c->mask, st, &c->scal->flag);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemcpyAsync(&c->hint[3], &c->scal->flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    *uncomputed = c->hint[3];
    return LEMHIP_OK;
}

template <typename T>
int lem_fa_dag(lemhip_ctx *c)
{
    dim3 blk(64, 4), grd((c->W + 63) / 64, (c->H + 3) / 4);
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    k_fa_prep<<<grd, blk, 0, c->s>>>(c->W, c->H, c->mask, c->fd, c->dmask, st);
    LEM_LAUNCHED(c);
    FaCell<T> cell{reinterpret_cast<const T *>(c->w), reinterpret_cast<typename FaStore<T>::type *>(c->fa)};
    return run_rounds<64, 16>(c, cell, &c->h_fa_unc);
}

template <typename T>
int lem_sed_dag(lemhip_ctx *c)
{
    dim3 blk(64, 4), grd((c->W + 63) / 64, (c->H + 3) / 4);
    int *st = reinterpret_cast<int *>(c->scan_tmp);
    k_sed_prep<<<grd, blk, 0, c->s>>>(c->W, c->H, c->mask, c->fd, c->dmask, st);
    LEM_LAUNCHED(c);
    double ddk = c->p.ddk, dck = c->p.dck, dgk = c->p.dgk;
    if (ddk > 0.95) ddk = 0.95;   // mfd_accum.cu:406-407
    if (dck > 0.9) dck = 0.9;
    SedCell<T> cell{reinterpret_cast<const T *>(c->ero), reinterpret_cast<const T *>(c->inci),
                    reinterpret_cast<const T *>(c->geli), reinterpret_cast<T *>(c->depo),
                    reinterpret_cast<T *>(c->stone), reinterpret_cast<T *>(c->fines), c->soilT, ddk, dck, dgk};
    return run_rounds<64, 16>(c, cell, &c->h_sed_unc);
}

template int lem_fa_dag<double>(lemhip_ctx *);
template int lem_fa_dag<float>(lemhip_ctx *);
template int lem_sed_dag<double>(lemhip_ctx *);
template int lem_sed_dag<float>(lemhip_ctx *);
