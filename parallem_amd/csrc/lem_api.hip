// lem_api.hip — the extern "C" boundary (include/lemhip.h): context
// lifecycle, the reference's host-side initialisation performed on the
// device, host-layout upload/download, and the timestep phases.
#include "lem_ctx.h"
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <vector>
#include <type_traits>
#include <thread>
#include <mutex>

static int grid1d(long n, int block = 256, int cap = 8192)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ------------------------------------------------------------ init kernels

// Initial state of one DEM rectangle D of the grid (dem_in == nullptr: the
// rectangle becomes nodata).  Inputs are in the DEM's own row-major layout.
// memory.cu:207-208 (soil thickness), memory.cu:229-237 (mask),
// Data.cpp:131-185 (process matrices), Data.cpp:231-271 (climate, first call)
template <typename T, typename FS>
__global__ void k_init_rect(int W, DemDesc D, int slot, lem_params p, const double *__restrict__ dem_in,
                            const double *__restrict__ rain_in, int soilT_mode, uint8_t *demid, double *dem,
                            uint8_t *mask, double *soilT, T *rain, T *temp, T *last_rain, T *last_temp, T *soilM,
                            T *stone, T *fines, T *nut, T *soilB, T *TotB, T *w, T *slopeptr, T *ero, T *inci,
                            T *geli, T *depo, T *dz, T *wC, T *wP, uint8_t *fd, uint8_t *SFD, T *slopes, FS *fa,
                            long N, int *soilneg)
{
    int neg = 0;
    double rainmax = p.rain * 2;                     // Data.cpp:246-250
    double rainmin = p.rain;
    double increment = (rainmax - rainmin) / D.h;
    double minheight = 467.228;
    double rescale = (10000 / (p.cellsize * p.cellsize));
    long n = (long)D.w * D.h;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        int row = (int)(k / D.w), col = (int)(k % D.w);
        long x = (long)(D.y0 + row) * W + (D.x0 + col);
        double z = dem_in ? dem_in[k] : -9999.0;
        dem[x] = z;
        if (demid) demid[x] = dem_in ? (uint8_t)slot : (uint8_t)DEM_NONE;
        bool m = !(z == -9999);                      // memory.cu:229-237
        mask[x] = m ? 1 : 0;
        double sT = (z > 0) ? ((z - 400) / 1400) * 5 : 0.0; // memory.cu:207-208
        neg += (m && sT < 0.0) ? 1 : 0;
        if (soilT_mode == 1 && sT < 0.0) sT = 0.0;
        double r, t;
        if (m) {
            r = rain_in ? rain_in[k] : rainmin + (row * increment);
            t = p.temp - ((z - minheight) * 0.0065);
        } else {
            r = -9999.;
            t = -9999.;
        }
        rain[x] = (T)r; temp[x] = (T)t; last_rain[x] = (T)r; last_temp[x] = (T)t;
        if (m) {                                     // Data.cpp:131-152
            fd[x] = 0; SFD[x] = 0; slopeptr[x] = 0; dz[x] = 0; w[x] = (T)1.0;
            soilT[x] = sT;
            stone[x] = (T)p.stoneval; fines[x] = (T)p.fineval; soilM[x] = (T)p.soilMval;
            nut[x] = (T)p.nutval; soilB[x] = (T)p.soilBval;
            double tempeffect = 100 + (100 / (t + 15));
            TotB[x] = (T)((3000.0 + (5.0 * (r - tempeffect))) / rescale);
        } else {                                     // Data.cpp:155-170
            fd[x] = 0; SFD[x] = 0; dz[x] = 0; slopeptr[x] = 0; w[x] = 0;
            soilT[x] = 0.0;
            stone[x] = 0; fines[x] = 0; soilM[x] = 0; nut[x] = 0; soilB[x] = 0; TotB[x] = 0;
        }
        ero[x] = 0; geli[x] = 0; inci[x] = 0; depo[x] = 0; wC[x] = 0; wP[x] = 0;
        fa[x] = FS{};
        for (int d = 0; d < 8; d++) slopes[d * N + x] = 0;   // Data.cpp:181-185
    }
    if (soilneg) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) neg += __shfl_down(neg, o);
        if ((threadIdx.x & 63) == 0 && neg) atomicAdd(soilneg, neg);
    }
}

__global__ void k_slot_scal_init(DemScal *ds, int slot)
{
    DemScal z = {};
    ds[slot] = z;
}

// copytolastclimate (memory_dev.cu:363-373)
template <typename T>
__global__ void k_copy_last(long N, const T *rain, const T *temp, T *last_rain, T *last_temp)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        last_rain[i] = rain[i];
        last_temp[i] = temp[i];
    }
}

// SetClimateMatrices(iteration >= 0) (Data.cpp:231-271) on one DEM rectangle
template <typename T>
__global__ void k_set_climate(int W, DemDesc D, double rain_s, double temp_s, const double *dem, const uint8_t *mask,
                              T *rain, T *temp)
{
    double rainmax = rain_s * 2, rainmin = rain_s;
    double increment = (rainmax - rainmin) / D.h;
    long n = (long)D.w * D.h;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        int row = (int)(k / D.w);
        long x = (long)(D.y0 + row) * W + (D.x0 + (int)(k % D.w));
        if (mask[x]) {
            rain[x] = (T)(rainmin + (row * increment));
            temp[x] = (T)(temp_s - ((dem[x] - 467.228) * 0.0065));
        } else {
            rain[x] = (T)-9999.;
            temp[x] = (T)-9999.;
        }
    }
}

__global__ void k_slots_first_iter(int ns, const int *flags, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) ds[j].first_iter = flags[j];
}

// ---------------------------------------------------------- atlas packing

// Every DEM starts at a column and a row that are multiples of LEM_ALIGN:
// the fixed 32 x 32 tiling of the FP32 sediment pass's first rounds then
// lies in the same place relative to each DEM as in a single-DEM context.
// A DEM's rectangle with its nodata column takes its width rounded up to a
// multiple of LEM_ALIGN (less at the grid's right edge), a shelf its height
// rounded up.
#define LEM_ALIGN 32
static inline int align_up(int v) { return (v + LEM_ALIGN - 1) / LEM_ALIGN * LEM_ALIGN; }

bool ShelfPack::alloc(int w, int h, int *x, int *y)
{
    int pw = w + 1, ph = h + 1;
    int best = -1, bestk = -1, waste = 1 << 30;
    for (int s = 0; s < (int)shelves.size(); s++) {
        Shelf &S = shelves[s];
        if (S.h < ph || S.h - ph >= waste) continue;
        for (int k = 0; k < (int)S.free.size(); k++)
            if (S.free[k].second - S.free[k].first >= pw) { best = s; bestk = k; waste = S.h - ph; break; }
    }
    if (best < 0) {
        if (pw > W || top + ph > H) return false;
        Shelf S;
        S.y = top;
        S.h = ph;
        S.free.push_back({0, W});
        shelves.push_back(S);
        top = std::min(H, top + align_up(ph));
        best = (int)shelves.size() - 1;
        bestk = 0;
    }
    Shelf &S = shelves[best];
    *x = S.free[bestk].first;
    *y = S.y;
    // an interval ends at the grid's edge or at another DEM's (aligned) start
    S.free[bestk].first += std::min(align_up(pw), S.free[bestk].second - S.free[bestk].first);
    if (S.free[bestk].first == S.free[bestk].second) S.free.erase(S.free.begin() + bestk);
    return true;
}

void ShelfPack::release(int x, int y, int w, int h)
{
    (void)h;
    int pw = std::min(align_up(w + 1), W - x);   // what alloc took
    for (int s = 0; s < (int)shelves.size(); s++) {
        Shelf &S = shelves[s];
        if (S.y != y) continue;
        auto &F = S.free;
        size_t k = 0;
        while (k < F.size() && F[k].first < x) k++;
        F.insert(F.begin() + k, {x, x + pw});
        for (size_t q = 0; q + 1 < F.size();) {
            if (F[q].second == F[q + 1].first) { F[q].second = F[q + 1].second; F.erase(F.begin() + q + 1); }
            else q++;
        }
        // drop empty shelves from the top so their rows can take a taller shelf
        while (!shelves.empty()) {
            Shelf &T = shelves.back();
            if (T.free.size() == 1 && T.free[0].first == 0 && T.free[0].second == W) {
                top = T.y;
                shelves.pop_back();
            } else break;
        }
        return;
    }
}

// ------------------------------------------------------------- lifecycle

static void free_all(lemhip_ctx *c)
{
    void *ptrs[] = {c->dem, c->soilT, c->mask, c->slopes, c->soilM, c->stone, c->fines, c->nut, c->soilB, c->TotB,
                    c->rain, c->temp, c->last_rain, c->last_temp, c->fd, c->SFD, c->w, c->fa, c->slopeptr, c->ero,
                    c->inci, c->geli, c->depo, c->dz, c->wC, c->wP, c->sp0, c->sp1, c->cand, c->ws, c->scan_tmp,
                    c->z2, c->pl_tiles, c->pl_ctl, c->cub_tmp,
                    c->hkeys, c->hvals, c->elo, c->ehi, c->ew, c->rmm, c->seen, c->status, c->comp, c->cmax,
                    c->dmask, c->dag_list[0], c->dag_list[1], c->dag_list[2], c->dag_wake,
                    c->dag_ctl, c->partials, c->scal, c->demid, c->ddesc, c->dsc, c->slot_flags, c->labdem,
                    c->dstage, c->tcache, c->fx_shift, c->fx_tag_cell, c->fx_tag_sum, c->cs_rec};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (int q = 0; q < 2; q++) {
        if (c->elo2[q]) (void)hipFree(c->elo2[q]);
        if (c->ehi2[q]) (void)hipFree(c->ehi2[q]);
        if (c->ew2[q]) (void)hipFree(c->ew2[q]);
    }
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hdsc) (void)hipHostFree(c->hdsc);
    if (c->hflags) (void)hipHostFree(c->hflags);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->hint) (void)hipHostFree(c->hint);
    for (GraphSeg &g : c->gseg) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    if (c->prof_ready)
        for (int k = 0; k < PROF_COUNT; k++)
            for (int q = 0; q < PROF_MAXP; q++) {
                (void)hipEventDestroy(c->prof[k].a[q]);
                (void)hipEventDestroy(c->prof[k].b[q]);
            }
    if (c->s) (void)hipStreamDestroy(c->s);
}

static int upload_desc(lemhip_ctx *c)
{
    int ns = 0, kb = 1;
    for (int j = 0; j < LEM_MAX_SLOTS; j++)
        if (c->hdesc[j].w > 0) {
            ns = j + 1;
            if (slot_kb(c->hdesc[j].h) > kb) kb = slot_kb(c->hdesc[j].h);
        }
    c->nslot = ns > 0 ? ns : 1;
    c->slot_kbmax = kb;
    LEM_CHECK(c, hipMemcpyAsync(c->ddesc, c->hdesc, sizeof(DemDesc) * (LEM_MAX_SLOTS + 1), hipMemcpyHostToDevice,
                                c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

static int alloc_ctx(lemhip_ctx *c, const lem_params *p, int W, int H, int precision, int device_id)
{
    c->p = *p;
    c->p.width = W;
    c->p.height = H;
    c->prec = precision;
    c->dev = device_id;
    c->W = W;
    c->H = H;
    c->N = (long)W * H;
    c->Hfull = H;
    c->NP = c->N;
    c->tsz = precision == LEMHIP_FP64 ? 8 : 4;
    if (hipSetDevice(device_id) != hipSuccess) { c->err = "hipSetDevice failed"; return LEMHIP_E_HIP; }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device_id) == hipSuccess && ncu > 0)
            c->ncu = ncu;
    }
    LEM_CHECK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    long N = c->N;
    size_t T = c->tsz;
    auto A = [&](void **ptr, size_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes < 64 ? 64 : bytes);
        if (e != hipSuccess) { c->err = std::string("hipMalloc: ") + hipGetErrorString(e); return LEMHIP_E_NOMEM; }
        return LEMHIP_OK;
    };
    LEM_TRY(A((void **)&c->dem, N * 8));
    LEM_TRY(A((void **)&c->soilT, N * 8));
    LEM_TRY(A((void **)&c->mask, N));
    LEM_TRY(A(&c->slopes, 8 * N * T));
    void **tf[] = {&c->soilM, &c->stone, &c->fines, &c->nut, &c->soilB, &c->TotB, &c->rain, &c->temp,
                   &c->last_rain, &c->last_temp, &c->w, &c->slopeptr, &c->ero, &c->inci, &c->geli, &c->depo,
                   &c->dz, &c->wC, &c->wP};
    for (void **q : tf) LEM_TRY(A(q, N * T));
    LEM_TRY(A((void **)&c->fd, N));
    LEM_TRY(A((void **)&c->SFD, N));
    LEM_TRY(A(&c->fa, N * 8));  // double, or float2 hi/lo
    LEM_TRY(A((void **)&c->sp0, N * 4));
    LEM_TRY(A((void **)&c->sp1, N * 4));
    long ptiles = (long)((W + PTW - 1) / PTW) * ((H + PTH - 1) / PTH);
    LEM_TRY(A((void **)&c->pl_tiles, 6 * ptiles * 4));
    LEM_TRY(A(&c->pl_ctl, 256));
    LEM_CHECK(c, hipMemset(c->pl_tiles, 0, 6 * ptiles * 4));
    LEM_TRY(A((void **)&c->cand, N * 4));
    LEM_TRY(A((void **)&c->ws, N * 4));
    LEM_TRY(A(&c->tcache, 2 * N * T));
    LEM_TRY(A((void **)&c->scan_tmp, N * 4));
    LEM_TRY(A((void **)&c->z2, N * 8));
    LEM_TRY(A((void **)&c->dmask, N));
    LEM_TRY(A((void **)&c->labdem, N + 1));
    long ntiles = 0;
    for (int k = 0; k < lem_dag_shape_count(); k++) {
        int tw, th;
        lem_dag_shape_dims(k, &tw, &th);
        long nt = ((c->W + tw - 1) / tw) * (long)((c->H + th - 1) / th);
        if (nt > ntiles) ntiles = nt;
    }
    c->dag_tiles_max = ntiles;
    for (int q = 0; q < 3; q++) LEM_TRY(A((void **)&c->dag_list[q], ntiles * 4));
    LEM_TRY(A((void **)&c->dag_wake, ntiles * 4));
    LEM_CHECK(c, hipMemset(c->dag_wake, 0, ntiles * 4));
    LEM_TRY(A(&c->dag_ctl, 512));
    LEM_TRY(A((void **)&c->partials, 3L * SLOT_KB * (LEM_MAX_SLOTS + 1) * 8));
    LEM_TRY(A((void **)&c->fx_shift, sizeof(int) * (LEM_MAX_SLOTS + 1)));
    // tagged cells of the integer flow accumulation: one path per DEM (longer
    // paths fall back to the dependency engine)
    c->fx_tag_cap = precision == LEMHIP_FP64 ? 1 : (N / 16 + 4096);
    LEM_TRY(A((void **)&c->fx_tag_cell, c->fx_tag_cap * sizeof(int)));
    LEM_TRY(A((void **)&c->fx_tag_sum, c->fx_tag_cap * sizeof(unsigned long long)));
    LEM_TRY(A((void **)&c->scal, sizeof(LemScalars)));
    LEM_CHECK(c, hipMemset(c->scal, 0, sizeof(LemScalars)));
    LEM_TRY(A((void **)&c->ddesc, sizeof(DemDesc) * (LEM_MAX_SLOTS + 1)));
    LEM_TRY(A((void **)&c->dsc, sizeof(DemScal) * (LEM_MAX_SLOTS + 1)));
    LEM_CHECK(c, hipMemset(c->dsc, 0, sizeof(DemScal) * (LEM_MAX_SLOTS + 1)));
    LEM_TRY(A((void **)&c->slot_flags, sizeof(int) * (LEM_MAX_SLOTS + 1)));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hscal, sizeof(LemScalars), hipHostMallocDefault));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hdsc, sizeof(DemScal) * (LEM_MAX_SLOTS + 1), hipHostMallocDefault));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hflags, sizeof(int) * (LEM_MAX_SLOTS + 1), hipHostMallocDefault));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hint, 64 * sizeof(int), hipHostMallocDefault));
    memset(c->hscal, 0, sizeof(LemScalars));
    memset(c->hdsc, 0, sizeof(DemScal) * (LEM_MAX_SLOTS + 1));
    memset(c->hflags, 0, sizeof(int) * (LEM_MAX_SLOTS + 1));
    return LEMHIP_OK;
}

static int check_create(const lem_params *p, int W, int H, int precision)
{
    if (!p || W <= 0 || H <= 0 || p->cellsize <= 0) return LEMHIP_E_ARG;
    if (precision != LEMHIP_FP64 && precision != LEMHIP_FP32_KAHAN) return LEMHIP_E_ARG;
    if ((long)W * H >= 0x7FFFFFFFL) return LEMHIP_E_ARG;  // labels and cell pointers are int32
    return LEMHIP_OK;
}

extern "C" int lemhip_create(const lem_params *p, int precision, int device_id, lemhip_ctx **out)
{
    if (!out) return LEMHIP_E_ARG;
    *out = nullptr;
    if (!p) return LEMHIP_E_ARG;
    LEM_TRY(check_create(p, p->width, p->height, precision));
    lemhip_ctx *c = new (std::nothrow) lemhip_ctx();
    if (!c) return LEMHIP_E_NOMEM;
    *out = c;
    LEM_TRY(alloc_ctx(c, p, p->width, p->height, precision, device_id));
    c->hrect[0] = DemDesc{0, 0, c->W, c->H};
    c->hdesc[0] = c->hrect[0];
    c->hstate[0] = 2;
    return upload_desc(c);
}

extern "C" int lemhip_create_atlas(const lem_params *p, int atlas_w, int atlas_h, int precision, int device_id,
                                   lemhip_ctx **out)
{
    if (!out) return LEMHIP_E_ARG;
    *out = nullptr;
    LEM_TRY(check_create(p, atlas_w, atlas_h, precision));
    lemhip_ctx *c = new (std::nothrow) lemhip_ctx();
    if (!c) return LEMHIP_E_NOMEM;
    *out = c;
    LEM_TRY(alloc_ctx(c, p, atlas_w, atlas_h, precision, device_id));
    c->atlas = true;
    c->pack.W = atlas_w;
    c->pack.H = atlas_h;
    LEM_CHECK(c, hipMalloc((void **)&c->demid, c->N));
    LEM_CHECK(c, hipMemset(c->demid, DEM_NONE, c->N));
    // every cell starts as nodata
    DemDesc all{0, 0, c->W, c->H};
    if (precision == LEMHIP_FP64)
        k_init_rect<double, double><<<grid1d(c->N), 256, 0, c->s>>>(
            c->W, all, DEM_NONE, c->p, nullptr, nullptr, 0, c->demid, c->dem, c->mask, c->soilT, (double *)c->rain,
            (double *)c->temp, (double *)c->last_rain, (double *)c->last_temp, (double *)c->soilM, (double *)c->stone,
            (double *)c->fines, (double *)c->nut, (double *)c->soilB, (double *)c->TotB, (double *)c->w,
            (double *)c->slopeptr, (double *)c->ero, (double *)c->inci, (double *)c->geli, (double *)c->depo,
            (double *)c->dz, (double *)c->wC, (double *)c->wP, c->fd, c->SFD, (double *)c->slopes, (double *)c->fa,
            c->N, nullptr);
    else
        k_init_rect<float, float2><<<grid1d(c->N), 256, 0, c->s>>>(
            c->W, all, DEM_NONE, c->p, nullptr, nullptr, 0, c->demid, c->dem, c->mask, c->soilT, (float *)c->rain,
            (float *)c->temp, (float *)c->last_rain, (float *)c->last_temp, (float *)c->soilM, (float *)c->stone,
            (float *)c->fines, (float *)c->nut, (float *)c->soilB, (float *)c->TotB, (float *)c->w,
            (float *)c->slopeptr, (float *)c->ero, (float *)c->inci, (float *)c->geli, (float *)c->depo,
            (float *)c->dz, (float *)c->wC, (float *)c->wP, c->fd, c->SFD, (float *)c->slopes, (float2 *)c->fa, c->N,
            nullptr);
    LEM_LAUNCHED(c);
    return upload_desc(c);
}

extern "C" int lemhip_destroy(lemhip_ctx *c)
{
    if (!c) return LEMHIP_E_ARG;
    (void)hipSetDevice(c->dev);
    if (c->s) (void)hipStreamSynchronize(c->s);
    free_all(c);
    delete c;
    return LEMHIP_OK;
}

extern "C" const char *lemhip_last_error(const lemhip_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int stage_reserve(lemhip_ctx *c, size_t bytes)
{
    if (bytes <= c->stage_bytes) return LEMHIP_OK;
    if (c->stage) (void)hipHostFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    LEM_CHECK(c, hipHostMalloc(&c->stage, bytes, hipHostMallocDefault));
    c->stage_bytes = bytes;
    return LEMHIP_OK;
}

static int dstage_reserve(lemhip_ctx *c, size_t bytes)
{
    if (bytes <= c->dstage_bytes) return LEMHIP_OK;
    if (c->dstage) (void)hipFree(c->dstage);
    c->dstage = nullptr;
    c->dstage_bytes = 0;
    LEM_CHECK(c, hipMalloc(&c->dstage, bytes));
    c->dstage_bytes = bytes;
    return LEMHIP_OK;
}

static int slot_rect(lemhip_ctx *c, int slot, int need_state, DemDesc *D)
{
    if (slot < 0 || slot >= LEM_MAX_SLOTS || c->hstate[slot] < need_state) {
        c->err = "slot not open";
        return LEMHIP_E_ARG;
    }
    *D = c->hrect[slot];
    return LEMHIP_OK;
}

template <typename T>
static void launch_init_rect(lemhip_ctx *c, const DemDesc &D, int slot, const double *dem_d, const double *rain_d,
                             int soilT_mode)
{
    using FS = typename std::conditional<sizeof(T) == 8, double, float2>::type;
    long n = (long)D.w * D.h;
    k_init_rect<T, FS><<<grid1d(n), 256, 0, c->s>>>(
        c->W, D, slot, c->p, dem_d, rain_d, soilT_mode, c->demid, c->dem, c->mask, c->soilT, (T *)c->rain,
        (T *)c->temp, (T *)c->last_rain, (T *)c->last_temp, (T *)c->soilM, (T *)c->stone, (T *)c->fines,
        (T *)c->nut, (T *)c->soilB, (T *)c->TotB, (T *)c->w, (T *)c->slopeptr, (T *)c->ero, (T *)c->inci,
        (T *)c->geli, (T *)c->depo, (T *)c->dz, (T *)c->wC, (T *)c->wP, c->fd, c->SFD, (T *)c->slopes, (FS *)c->fa,
        c->NP, slot == DEM_NONE ? nullptr : &c->dsc[slot].soilT_negative);
}

// Valid cells of one DEM (host layout w x h) that the reference's label
// kernels also reach from threads past a row end (newflow.cu:358-360: the
// block of 16 threads covering the last columns continues into the next
// row's first P columns).  The count of valid cells with a negative formula
// soil thickness (memory.cu:207-208) is taken by the initialisation kernel.
static int phantom_cells(const double *dem, int w, int h)
{
    const int P = (w / 16 + 1) * 16 - w;
    int ph = 0;
    for (int r = 1; r < h; r++)
        for (int q = 0; q < P && q < w; q++)
            if (dem[(long)r * w + q] != -9999.0) ph++;
    return ph;
}

// Data.cpp:38-66 (outlet), memory.cu:200-240 and Data.cpp:120-271 (initial
// matrices) for one DEM rectangle
static int init_slot(lemhip_ctx *c, int slot, const double *dem, const double *rain, int soilT_mode)
{
    c->tcache_ok = false;   // the slot's temperatures are set
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 1, &D));
    c->hphantom[slot] = phantom_cells(dem, D.w, D.h);
    long n = (long)D.w * D.h;
    LEM_TRY(dstage_reserve(c, (rain ? 2 : 1) * n * 8));
    double *dem_d = (double *)c->dstage;
    double *rain_d = rain ? dem_d + n : nullptr;
    LEM_CHECK(c, hipMemcpyAsync(dem_d, dem, n * 8, hipMemcpyHostToDevice, c->s));
    if (rain) LEM_CHECK(c, hipMemcpyAsync(rain_d, rain, n * 8, hipMemcpyHostToDevice, c->s));
    k_slot_scal_init<<<1, 1, 0, c->s>>>(c->dsc, slot);
    if (c->prec == LEMHIP_FP64) launch_init_rect<double>(c, D, slot, dem_d, rain_d, soilT_mode);
    else launch_init_rect<float>(c, D, slot, dem_d, rain_d, soilT_mode);
    LEM_LAUNCHED(c);
    c->hdesc[slot] = D;
    c->hstate[slot] = 2;
    LEM_TRY(upload_desc(c));
    LEM_TRY(lem_slot_lowest(c, c->dem, slot, 1));
    LEM_TRY(lem_slot_first_lowest(c, c->dem, slot, 1));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_init_state(lemhip_ctx *c, const double *dem, const double *rain, int soilT_mode)
{
    if (!c || !dem) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_init"; return LEMHIP_E_STATE; }
    if (soilT_mode != 0 && soilT_mode != 1) { c->err = "soilT_mode must be 0 or 1"; return LEMHIP_E_ARG; }
    LEM_CHECK(c, hipSetDevice(c->dev));
    return init_slot(c, 0, dem, rain, soilT_mode);
}

extern "C" int lemhip_slot_open(lemhip_ctx *c, int width, int height, int *slot)
{
    if (!c || !slot || width <= 0 || height <= 0) return LEMHIP_E_ARG;
    *slot = -1;
    if (!c->atlas) { c->err = "not an atlas context"; return LEMHIP_E_STATE; }
    int j = 0;
    while (j < LEM_MAX_SLOTS && c->hstate[j] != 0) j++;
    if (j == LEM_MAX_SLOTS) { c->err = "all slots in use"; return LEMHIP_E_NOMEM; }
    int x = 0, y = 0;
    if (!c->pack.alloc(width, height, &x, &y)) { c->err = "no room in the atlas"; return LEMHIP_E_NOMEM; }
    c->hrect[j] = DemDesc{x, y, width, height};
    c->hstate[j] = 1;
    *slot = j;
    return LEMHIP_OK;
}

extern "C" int lemhip_slot_init(lemhip_ctx *c, int slot, const double *dem, const double *rain, int soilT_mode)
{
    if (!c || !dem) return LEMHIP_E_ARG;
    if (soilT_mode != 0 && soilT_mode != 1) { c->err = "soilT_mode must be 0 or 1"; return LEMHIP_E_ARG; }
    LEM_CHECK(c, hipSetDevice(c->dev));
    return init_slot(c, slot, dem, rain, soilT_mode);
}

extern "C" int lemhip_slot_close(lemhip_ctx *c, int slot)
{
    if (!c) return LEMHIP_E_ARG;
    if (!c->atlas) { c->err = "not an atlas context"; return LEMHIP_E_STATE; }
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 1, &D));
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (c->prec == LEMHIP_FP64) launch_init_rect<double>(c, D, DEM_NONE, nullptr, nullptr, 0);
    else launch_init_rect<float>(c, D, DEM_NONE, nullptr, nullptr, 0);
    LEM_LAUNCHED(c);
    c->hdesc[slot] = DemDesc{0, 0, 0, 0};
    c->hstate[slot] = 0;
    c->pack.release(D.x0, D.y0, D.w, D.h);
    return upload_desc(c);
}

extern "C" int lemhip_slot_rect(lemhip_ctx *c, int slot, int *x0, int *y0, int *w, int *h)
{
    if (!c) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 1, &D));
    if (x0) *x0 = D.x0;
    if (y0) *y0 = D.y0;
    if (w) *w = D.w;
    if (h) *h = D.h;
    return LEMHIP_OK;
}

// ------------------------------------------------------- upload / download

static void *field_ptr(lemhip_ctx *c, int f, int *kind)
{
    // kind: 0 = T bulk, 1 = double, 2 = u8 code field, 3 = mask, 4 = slopes, 5 = fa
    switch (f) {
    case LEM_F_DEM: *kind = 1; return c->dem;
    case LEM_F_SOILT: *kind = 1; return c->soilT;
    case LEM_F_MASK: *kind = 3; return c->mask;
    case LEM_F_SLOPES: *kind = 4; return c->slopes;
    case LEM_F_FD: *kind = 2; return c->fd;
    case LEM_F_SFD: *kind = 2; return c->SFD;
    case LEM_F_FA: *kind = 5; return c->fa;
    case LEM_F_SLOPEPTR: *kind = 0; return c->slopeptr;
    case LEM_F_SOILM: *kind = 0; return c->soilM;
    case LEM_F_STONE: *kind = 0; return c->stone;
    case LEM_F_FINES: *kind = 0; return c->fines;
    case LEM_F_NUT: *kind = 0; return c->nut;
    case LEM_F_SOILB: *kind = 0; return c->soilB;
    case LEM_F_TOTB: *kind = 0; return c->TotB;
    case LEM_F_RAIN: *kind = 0; return c->rain;
    case LEM_F_TEMP: *kind = 0; return c->temp;
    case LEM_F_LAST_RAIN: *kind = 0; return c->last_rain;
    case LEM_F_LAST_TEMP: *kind = 0; return c->last_temp;
    case LEM_F_RUNOFF: *kind = 0; return c->w;
    case LEM_F_ERO: *kind = 0; return c->ero;
    case LEM_F_INCI: *kind = 0; return c->inci;
    case LEM_F_GELI: *kind = 0; return c->geli;
    case LEM_F_DEPO: *kind = 0; return c->depo;
    case LEM_F_DZ: *kind = 0; return c->dz;
    case LEM_F_WC: *kind = 0; return c->wC;
    case LEM_F_WP: *kind = 0; return c->wP;
    default: return nullptr;
    }
}

static size_t kind_elem(const lemhip_ctx *c, int kind)
{
    switch (kind) {
    case 0: case 4: return c->tsz;
    case 1: case 5: return 8;
    default: return 1;
    }
}

// rectangle D of plane `plane` of a grid field <-> contiguous w*h block at `buf`
static int rect_copy(lemhip_ctx *c, void *buf, void *base, int plane, size_t es, const DemDesc &D, bool down)
{
    char *dev = (char *)base + ((size_t)plane * c->NP + (size_t)D.y0 * c->W + D.x0) * es;
    char *blk = (char *)buf + (size_t)plane * D.w * D.h * es;
    if (down)
        LEM_CHECK(c, hipMemcpy2DAsync(blk, D.w * es, dev, c->W * es, D.w * es, D.h, hipMemcpyDeviceToHost, c->s));
    else
        LEM_CHECK(c, hipMemcpy2DAsync(dev, c->W * es, blk, D.w * es, D.w * es, D.h, hipMemcpyHostToDevice, c->s));
    return LEMHIP_OK;
}

static int slot_download(lemhip_ctx *c, int slot, int field, void *host)
{
    int kind = -1;
    void *src = field_ptr(c, field, &kind);
    if (!src) { c->err = "unknown field"; return LEMHIP_E_ARG; }
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    LEM_CHECK(c, hipSetDevice(c->dev));
    long n = (long)D.w * D.h;
    size_t es = kind_elem(c, kind);
    int planes = kind == 4 ? 8 : 1;
    size_t bytes = planes * n * es + (kind == 2 ? n : 0);
    LEM_TRY(stage_reserve(c, bytes));
    for (int q = 0; q < planes; q++) LEM_TRY(rect_copy(c, c->stage, src, q, es, D, true));
    uint8_t *mstage = (uint8_t *)c->stage + planes * n * es;
    if (kind == 2) LEM_TRY(rect_copy(c, mstage, c->mask, 0, 1, D, true));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    if (kind == 1 || (kind == 0 && c->prec == LEMHIP_FP64) || (kind == 5 && c->prec == LEMHIP_FP64)) {
        memcpy(host, c->stage, n * 8);
    } else if (kind == 0) {
        const float *s = (const float *)c->stage;
        double *d = (double *)host;
        for (long i = 0; i < n; i++) d[i] = s[i];
    } else if (kind == 5) {
        const float *s = (const float *)c->stage;
        double *d = (double *)host;
        for (long i = 0; i < n; i++) d[i] = (double)s[2 * i] + (double)s[2 * i + 1];
    } else if (kind == 3) {
        const uint8_t *s = (const uint8_t *)c->stage;
        int *d = (int *)host;
        for (long i = 0; i < n; i++) d[i] = s[i];
    } else if (kind == 2) {
        // nodata cells carry -9999 in the reference's fd/SFD (newflow.cu:74-78)
        const uint8_t *s = (const uint8_t *)c->stage;
        int *d = (int *)host;
        for (long i = 0; i < n; i++) d[i] = mstage[i] ? (int)s[i] : -9999;
    } else if (kind == 4) {
        double *d = (double *)host;
        if (c->prec == LEMHIP_FP64) {
            const double *s = (const double *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < n; i++) d[8 * i + k] = s[k * n + i];
        } else {
            const float *s = (const float *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < n; i++) d[8 * i + k] = s[k * n + i];
        }
    }
    return LEMHIP_OK;
}

static int slot_upload(lemhip_ctx *c, int slot, int field, const void *host)
{
    c->tcache_ok = false;   // any field: temperatures among them
    int kind = -1;
    void *dst = field_ptr(c, field, &kind);
    if (!dst) { c->err = "unknown field"; return LEMHIP_E_ARG; }
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    LEM_CHECK(c, hipSetDevice(c->dev));
    long n = (long)D.w * D.h;
    size_t es = kind_elem(c, kind);
    int planes = kind == 4 ? 8 : 1;
    LEM_TRY(stage_reserve(c, planes * n * es));
    if (kind == 1 || (kind == 0 && c->prec == LEMHIP_FP64) || (kind == 5 && c->prec == LEMHIP_FP64)) {
        memcpy(c->stage, host, n * 8);
    } else if (kind == 0) {
        float *d = (float *)c->stage;
        const double *s = (const double *)host;
        for (long i = 0; i < n; i++) d[i] = (float)s[i];
    } else if (kind == 5) {
        float *d = (float *)c->stage;
        const double *s = (const double *)host;
        for (long i = 0; i < n; i++) {
            float hi = (float)s[i];
            d[2 * i] = hi;
            d[2 * i + 1] = (float)(s[i] - (double)hi);
        }
    } else if (kind == 2 || kind == 3) {
        uint8_t *d = (uint8_t *)c->stage;
        const int *s = (const int *)host;
        for (long i = 0; i < n; i++) d[i] = (uint8_t)(s[i] < 0 ? 0 : s[i]);
    } else if (kind == 4) {
        const double *s = (const double *)host;
        if (c->prec == LEMHIP_FP64) {
            double *d = (double *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < n; i++) d[k * n + i] = s[8 * i + k];
        } else {
            float *d = (float *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < n; i++) d[k * n + i] = (float)s[8 * i + k];
        }
    }
    for (int q = 0; q < planes; q++) LEM_TRY(rect_copy(c, c->stage, dst, q, es, D, false));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_download(lemhip_ctx *c, int field, void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_download"; return LEMHIP_E_STATE; }
    return slot_download(c, 0, field, host);
}

extern "C" int lemhip_upload(lemhip_ctx *c, int field, const void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_upload"; return LEMHIP_E_STATE; }
    return slot_upload(c, 0, field, host);
}

extern "C" int lemhip_slot_download(lemhip_ctx *c, int slot, int field, void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    return slot_download(c, slot, field, host);
}

extern "C" int lemhip_slot_upload(lemhip_ctx *c, int slot, int field, const void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    return slot_upload(c, slot, field, host);
}

// outlet cell index in the DEM's own row-major numbering
extern "C" int lemhip_slot_get_outlet(lemhip_ctx *c, int slot, int *outlet)
{
    if (!c || !outlet) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    LEM_TRY(lem_read_slots(c));
    int o = c->hdsc[slot].outlet;
    *outlet = (o / c->W - D.y0) * D.w + (o % c->W - D.x0);
    return LEMHIP_OK;
}

extern "C" int lemhip_slot_set_outlet(lemhip_ctx *c, int slot, int outlet)
{
    if (!c) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    if (outlet < 0 || outlet >= D.w * D.h) return LEMHIP_E_ARG;
    c->hint[8] = (D.y0 + outlet / D.w) * c->W + D.x0 + outlet % D.w;
    LEM_CHECK(c, hipMemcpyAsync(&c->dsc[slot].outlet, &c->hint[8], sizeof(int), hipMemcpyHostToDevice, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_get_outlet(lemhip_ctx *c, int *outlet)
{
    if (!c || !outlet) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_get_outlet"; return LEMHIP_E_STATE; }
    return lemhip_slot_get_outlet(c, 0, outlet);
}

extern "C" int lemhip_set_outlet(lemhip_ctx *c, int outlet)
{
    if (!c) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_set_outlet"; return LEMHIP_E_STATE; }
    return lemhip_slot_set_outlet(c, 0, outlet);
}

// --------------------------------------------------------------- phases

// mfd_simple.cu:336: the outlet moves to the FA maximum on a DEM's first
// iteration only
static int set_first_iter(lemhip_ctx *c, const int *iters, int iter)
{
    int first = c->p.start_iter - c->p.runiniter;
    for (int j = 0; j < c->nslot; j++) {
        int it = iters ? iters[j] : iter;
        c->hflags[j] = (c->hdesc[j].w > 0 && it == first) ? 1 : 0;
    }
    LEM_CHECK(c, hipMemcpyAsync(c->slot_flags, c->hflags, sizeof(int) * c->nslot, hipMemcpyHostToDevice, c->s));
    k_slots_first_iter<<<(c->nslot + 255) / 256, 256, 0, c->s>>>(c->nslot, c->slot_flags, c->dsc);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

extern "C" int lemhip_flow_direction(lemhip_ctx *c, int iter)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    LEM_TRY(lem_fd_phase(c, iter));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_runoff(lemhip_ctx *c)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    LEM_TRY(lem_runoff_phase(c));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_flow_accum(lemhip_ctx *c, int iter)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    LEM_TRY(set_first_iter(c, nullptr, iter));
    LEM_TRY(lem_fa_phase(c, iter));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_erosion(lemhip_ctx *c, int iter)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    LEM_TRY(lem_mod_phase(c, iter));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

static int prof_resolve(lemhip_ctx *c)
{
    if (!c->prof_on) return LEMHIP_OK;
    for (int k = 0; k < PROF_COUNT; k++) {
        ProfSlot &p = c->prof[k];
        for (int q = 0; q < p.n; q++) {
            float t = 0.f;
            LEM_CHECK(c, hipEventElapsedTime(&t, p.a[q], p.b[q]));
            p.ms += t;
            p.events++;
        }
        p.n = 0;
    }
    return LEMHIP_OK;
}

// An atlas steps only the rows its shelves use (rectangles are packed from
// row 0); the context's H/N describe that band for the duration of the call.
struct ActiveRows {
    lemhip_ctx *c;
    explicit ActiveRows(lemhip_ctx *ctx) : c(ctx)
    {
        if (c->atlas) {
            int h = c->pack.top < 1 ? 1 : (c->pack.top > c->Hfull ? c->Hfull : c->pack.top);
            c->H = h;
            c->N = (long)c->W * h;
        }
    }
    ~ActiveRows()
    {
        c->H = c->Hfull;
        c->N = c->NP;
    }
};

static int step_impl(lemhip_ctx *c, const int *iters, int iter)
{
    LEM_CHECK(c, hipSetDevice(c->dev));
    ActiveRows band(c);
    prof_begin(c, PROF_STEP);
    LEM_TRY(set_first_iter(c, iters, iter));
    LEM_TRY(lem_fd_phase(c, iter));
    LEM_TRY(lem_runoff_phase(c));
    LEM_TRY(lem_fa_phase(c, iter));
    LEM_TRY(lem_mod_phase(c, iter));
    prof_end(c, PROF_STEP);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    LEM_TRY(prof_resolve(c));
    return LEMHIP_OK;
}

extern "C" int lemhip_get_tuning(lemhip_ctx *c, lem_tuning *o)
{
    if (!c || !o) return LEMHIP_E_ARG;
    lem_dag_shape_dims(c->dag_shape, &o->dag_tile_w, &o->dag_tile_h);
    o->dag_batch = c->dag_batch;
    o->plateau_rounds_per_launch = c->plat_k;
    for (int k = 0; k < 16; k++) o->tuned_ms[k] = c->tune_ms[k];
    o->fp32_integer_fa = c->fx_enabled ? 1 : 0;
    o->use_graphs = c->use_graphs ? 1 : 0;
    o->fp32_tree_sediment = c->cs_rounds;
    return LEMHIP_OK;
}

extern "C" int lemhip_set_tuning(lemhip_ctx *c, const lem_tuning *t)
{
    if (!c || !t) return LEMHIP_E_ARG;
    int shape = -1;
    for (int k = 0; k < lem_dag_shape_count(); k++) {
        int tw, th;
        lem_dag_shape_dims(k, &tw, &th);
        if (tw == t->dag_tile_w && th == t->dag_tile_h) shape = k;
    }
    if (shape < 0) { c->err = "unsupported tile shape"; return LEMHIP_E_ARG; }
    if (t->dag_batch < 1 || t->dag_batch > 256) { c->err = "dag_batch out of range"; return LEMHIP_E_ARG; }
    int k = t->plateau_rounds_per_launch;
    if (k != 4 && k != 8 && k != 16) { c->err = "plateau rounds per launch must be 4, 8 or 16"; return LEMHIP_E_ARG; }
    if (t->fp32_integer_fa != 0 && t->fp32_integer_fa != 1) { c->err = "fp32_integer_fa must be 0 or 1"; return LEMHIP_E_ARG; }
    if (t->use_graphs != 0 && t->use_graphs != 1) { c->err = "use_graphs must be 0 or 1"; return LEMHIP_E_ARG; }
    if (t->fp32_tree_sediment < 0 || t->fp32_tree_sediment > 4) {
        c->err = "fp32_tree_sediment must be 0 to 4";
        return LEMHIP_E_ARG;
    }
    c->use_graphs = t->use_graphs != 0;
    // an explicit choice replaces a pending autotune
    c->tune_dag_pending = false;
    c->tune_plat_pending = false;
    c->dag_shape = shape;
    c->dag_batch = t->dag_batch;
    c->plat_k = k;
    c->fx_enabled = t->fp32_integer_fa != 0;
    c->cs_rounds = t->fp32_tree_sediment;
    return LEMHIP_OK;
}

extern "C" int lemhip_autotune(lemhip_ctx *c)
{
    if (!c) return LEMHIP_E_ARG;
    c->tune_dag_pending = true;
    c->tune_plat_pending = true;
    return LEMHIP_OK;
}

extern "C" int lemhip_step(lemhip_ctx *c, int iter)
{
    if (!c) return LEMHIP_E_ARG;
    return step_impl(c, nullptr, iter);
}

extern "C" int lemhip_step_slots(lemhip_ctx *c, const int *iters)
{
    if (!c || !iters) return LEMHIP_E_ARG;
    return step_impl(c, iters, 0);
}

// Several contexts stepped at once, each from its own host thread on its own
// stream: their kernel sequences overlap on the device, so one context's
// latency-bound stretches (the late rounds of the dependency passes, which
// keep few compute units busy) run beside another's bandwidth-bound kernels.
extern "C" int lemhip_step_group(lemhip_ctx *const *ctxs, int n, const int *iters, const int *const *slot_iters)
{
    if (!ctxs || n <= 0 || n > 64) return LEMHIP_E_ARG;
    for (int k = 0; k < n; k++) {
        if (!ctxs[k]) return LEMHIP_E_ARG;
        if (ctxs[k]->atlas ? !(slot_iters && slot_iters[k]) : !iters) return LEMHIP_E_ARG;
        for (int q = 0; q < k; q++)
            if (ctxs[q] == ctxs[k]) return LEMHIP_E_ARG;   // one thread per context
    }
    std::vector<int> rc(n, LEMHIP_OK);
    auto one = [&](int k) {
        lemhip_ctx *c = ctxs[k];
        rc[k] = c->atlas ? step_impl(c, slot_iters[k], 0) : step_impl(c, nullptr, iters[k]);
    };
    if (n == 1) {
        one(0);
    } else {
        std::vector<std::thread> th;
        th.reserve(n - 1);
        for (int k = 1; k < n; k++) th.emplace_back(one, k);
        one(0);
        for (auto &t : th) t.join();
    }
    for (int k = 0; k < n; k++)
        if (rc[k] != LEMHIP_OK) return rc[k];
    return LEMHIP_OK;
}

extern "C" int lemhip_profile_enable(lemhip_ctx *c, int on)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (on && !c->prof_ready) {
        for (int k = 0; k < PROF_COUNT; k++)
            for (int q = 0; q < PROF_MAXP; q++) {
                LEM_CHECK(c, hipEventCreate(&c->prof[k].a[q]));
                LEM_CHECK(c, hipEventCreate(&c->prof[k].b[q]));
            }
        c->prof_ready = true;
    }
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    LEM_TRY(prof_resolve(c));
    for (int k = 0; k < PROF_COUNT; k++) {
        c->prof[k].ms = 0.0;
        c->prof[k].launches = 0;
        c->prof[k].events = 0;
        c->prof[k].n = 0;
    }
    c->prof_on = on != 0 && c->prof_ready;
    return LEMHIP_OK;
}

extern "C" int lemhip_profile_read(lemhip_ctx *c, int n, double *ms, long long *launches)
{
    if (!c || !ms || !launches || n < 0) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    LEM_TRY(prof_resolve(c));
    for (int k = 0; k < n && k < PROF_COUNT; k++) {
        ms[k] = c->prof[k].ms;
        launches[k] = c->prof[k].launches;
    }
    return LEMHIP_OK;
}

extern "C" int lemhip_run(lemhip_ctx *c, int iter0, int nsteps)
{
    if (!c || nsteps < 0) return LEMHIP_E_ARG;
    for (int k = 0; k < nsteps; k++) LEM_TRY(lemhip_step(c, iter0 + k));
    return LEMHIP_OK;
}

static int set_climate_rect(lemhip_ctx *c, const DemDesc &D, double rain, double temp)
{
    c->tcache_ok = false;
    long n = (long)D.w * D.h;
    if (c->prec == LEMHIP_FP64)
        k_set_climate<double><<<grid1d(n), 256, 0, c->s>>>(c->W, D, rain, temp, c->dem, c->mask,
                                                           (double *)c->rain, (double *)c->temp);
    else
        k_set_climate<float><<<grid1d(n), 256, 0, c->s>>>(c->W, D, rain, temp, c->dem, c->mask,
                                                          (float *)c->rain, (float *)c->temp);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

extern "C" int lemhip_set_climate(lemhip_ctx *c, double rain, double temp)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    for (int j = 0; j < LEM_MAX_SLOTS; j++)
        if (c->hdesc[j].w > 0) LEM_TRY(set_climate_rect(c, c->hdesc[j], rain, temp));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

// no host wait: the new fields are written by a kernel on the context's
// stream, ahead of whatever the next call runs there
extern "C" int lemhip_slot_set_climate(lemhip_ctx *c, int slot, double rain, double temp)
{
    if (!c) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    LEM_CHECK(c, hipSetDevice(c->dev));
    LEM_TRY(set_climate_rect(c, D, rain, temp));
    return LEMHIP_OK;
}

extern "C" int lemhip_copy_to_last_climate(lemhip_ctx *c)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (c->prec == LEMHIP_FP64)
        k_copy_last<double><<<grid1d(c->N), 256, 0, c->s>>>(c->N, (double *)c->rain, (double *)c->temp,
                                                            (double *)c->last_rain, (double *)c->last_temp);
    else
        k_copy_last<float><<<grid1d(c->N), 256, 0, c->s>>>(c->N, (float *)c->rain, (float *)c->temp,
                                                           (float *)c->last_rain, (float *)c->last_temp);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_slot_summary(lemhip_ctx *c, int slot, lem_summary *o)
{
    if (!c || !o) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(slot_rect(c, slot, 2, &D));
    LEM_TRY(lem_read_slots(c));
    const DemScal &h = c->hdsc[slot];
    o->FA_max = h.FA_max;
    o->totE = h.totE; o->totI = h.totI; o->totG = h.totG; o->totD = h.totD;
    o->totweatherC = h.totwC; o->totweatherP = h.totwP; o->totbio = h.totbio;
    o->outletcellidx = (h.outlet / c->W - D.y0) * D.w + (h.outlet % c->W - D.x0);
    o->sinkcount = h.sinkcount;
    o->sinks_pass1 = h.sinks_pass1;
    // the remaining counters describe the whole grid's last step
    o->nlabels = c->h_nlab;
    o->nedges = c->h_nedges;
    o->plateau_rounds1 = c->h_rounds1;
    o->plateau_rounds2 = c->h_rounds2;
    o->fa_uncomputed = c->h_fa_unc;
    o->sed_uncomputed = c->h_sed_unc;
    o->totwater = h.totwater;
    o->phantom_cells = c->hphantom[slot];
    o->soilT_negative = h.soilT_negative;
    o->sed_open = c->h_sed_open;
    return LEMHIP_OK;
}

extern "C" int lemhip_summary(lemhip_ctx *c, lem_summary *o)
{
    if (!c || !o) return LEMHIP_E_ARG;
    if (c->atlas) { c->err = "atlas context: use lemhip_slot_summary"; return LEMHIP_E_STATE; }
    return lemhip_slot_summary(c, 0, o);
}

// ------------------------------------------------------ checkpoint / restart

// A snapshot holds what carries one DEM from a step to the next: the
// surface, soil thickness, mask, the sticky Slopes (the reference's restart,
// retriveProcessMatrices ioGDAL.cpp:662-804, reads height, stone, fines,
// soilM, soilB, soilT, nut and totbio back but not Slopes, so its resumed
// run's aveslope differs from an unbroken one), the soil, vegetation and
// climate fields and the outlet.  Fields are in the DEM's own row-major
// layout with the context's element types (bulk fields T, Slopes as 8
// planes of T, mask one byte), so restoring into a context of the same
// precision form is exact.
struct SnapHeader {
    char magic[8];
    int version, precision, w, h, outlet, nfields;
    double cellsize;
    long long bytes;
};
static const int kSnapFields[] = {LEM_F_DEM,   LEM_F_SOILT, LEM_F_MASK,      LEM_F_SLOPES,   LEM_F_SOILM,
                                  LEM_F_STONE, LEM_F_FINES, LEM_F_NUT,       LEM_F_SOILB,    LEM_F_TOTB,
                                  LEM_F_RAIN,  LEM_F_TEMP,  LEM_F_LAST_RAIN, LEM_F_LAST_TEMP};
static const int kSnapN = (int)(sizeof(kSnapFields) / sizeof(kSnapFields[0]));

static size_t snap_bytes(const lemhip_ctx *c, const DemDesc &D)
{
    size_t n = (size_t)D.w * D.h, b = sizeof(SnapHeader);
    for (int q = 0; q < kSnapN; q++) {
        int kind = -1;
        (void)field_ptr(const_cast<lemhip_ctx *>(c), kSnapFields[q], &kind);
        b += (kind == 4 ? 8 : 1) * n * kind_elem(c, kind);
    }
    return b;
}

static int snap_slot(lemhip_ctx *c, int slot, int need, DemDesc *D)
{
    if (!c->atlas && slot != 0) { c->err = "single-DEM context: slot must be 0"; return LEMHIP_E_ARG; }
    return slot_rect(c, slot, need, D);
}

extern "C" int lemhip_snapshot_size(lemhip_ctx *c, int slot, size_t *bytes)
{
    if (!c || !bytes) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(snap_slot(c, slot, 1, &D));
    *bytes = snap_bytes(c, D);
    return LEMHIP_OK;
}

extern "C" int lemhip_snapshot(lemhip_ctx *c, int slot, void *buf, size_t bytes)
{
    if (!c || !buf) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(snap_slot(c, slot, 2, &D));
    const size_t need = snap_bytes(c, D);
    if (bytes < need) { c->err = "snapshot buffer too small"; return LEMHIP_E_ARG; }
    LEM_CHECK(c, hipSetDevice(c->dev));
    SnapHeader h = {};
    memcpy(h.magic, "LEMSNAP", 8);
    h.version = 1;
    h.precision = c->prec;
    h.w = D.w;
    h.h = D.h;
    LEM_TRY(lem_read_slots(c));
    const int o = c->hdsc[slot].outlet;
    h.outlet = (o / c->W - D.y0) * D.w + (o % c->W - D.x0);
    h.nfields = kSnapN;
    h.cellsize = c->p.cellsize;
    h.bytes = (long long)need;
    memcpy(buf, &h, sizeof h);
    char *out = (char *)buf + sizeof h;
    const size_t n = (size_t)D.w * D.h;
    for (int q = 0; q < kSnapN; q++) {
        int kind = -1;
        void *src = field_ptr(c, kSnapFields[q], &kind);
        const size_t es = kind_elem(c, kind);
        const int planes = kind == 4 ? 8 : 1;
        LEM_TRY(stage_reserve(c, planes * n * es));
        for (int pl = 0; pl < planes; pl++) LEM_TRY(rect_copy(c, c->stage, src, pl, es, D, true));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        memcpy(out, c->stage, planes * n * es);
        out += planes * n * es;
    }
    return LEMHIP_OK;
}

extern "C" int lemhip_restore(lemhip_ctx *c, int slot, const void *buf, size_t bytes)
{
    if (!c || !buf || bytes < sizeof(SnapHeader)) return LEMHIP_E_ARG;
    DemDesc D;
    LEM_TRY(snap_slot(c, slot, 1, &D));
    SnapHeader h;
    memcpy(&h, buf, sizeof h);
    if (memcmp(h.magic, "LEMSNAP", 8) != 0 || h.version != 1) { c->err = "not a snapshot"; return LEMHIP_E_ARG; }
    if (h.precision != c->prec) { c->err = "snapshot of the other precision form"; return LEMHIP_E_ARG; }
    if (h.w != D.w || h.h != D.h) { c->err = "snapshot of a DEM of another size"; return LEMHIP_E_ARG; }
    if (h.nfields != kSnapN || (size_t)h.bytes != snap_bytes(c, D) || bytes < (size_t)h.bytes) {
        c->err = "snapshot size mismatch";
        return LEMHIP_E_ARG;
    }
    if (h.outlet < 0 || h.outlet >= D.w * D.h) { c->err = "snapshot outlet out of range"; return LEMHIP_E_ARG; }
    LEM_CHECK(c, hipSetDevice(c->dev));
    const size_t n = (size_t)D.w * D.h;
    // the rectangle starts from the snapshot's surface (slot id, defaults of
    // the per-step fields, per-slot scalars), then every field is replaced
    std::vector<double> dem(n);
    memcpy(dem.data(), (const char *)buf + sizeof h, n * 8);
    LEM_TRY(init_slot(c, slot, dem.data(), nullptr, 0));
    const char *in = (const char *)buf + sizeof h;
    for (int q = 0; q < kSnapN; q++) {
        int kind = -1;
        void *dst = field_ptr(c, kSnapFields[q], &kind);
        const size_t es = kind_elem(c, kind);
        const int planes = kind == 4 ? 8 : 1;
        LEM_TRY(stage_reserve(c, planes * n * es));
        memcpy(c->stage, in, planes * n * es);
        for (int pl = 0; pl < planes; pl++) LEM_TRY(rect_copy(c, c->stage, dst, pl, es, D, false));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        in += planes * n * es;
    }
    c->tcache_ok = false;
    c->hint[8] = (D.y0 + h.outlet / D.w) * c->W + D.x0 + h.outlet % D.w;
    LEM_CHECK(c, hipMemcpyAsync(&c->dsc[slot].outlet, &c->hint[8], sizeof(int), hipMemcpyHostToDevice, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

// ------------------------------------------------------------ test hooks

__global__ void k_math_eval(int fn, int n, const double *x, const double *y, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0;
    switch (fn) {
    case 0: r = lm_exp(x[i]); break;
    case 1: r = lm_log(x[i]); break;
    case 2: r = lm_log10(x[i]); break;
    case 3: r = lm_pow(x[i], y[i]); break;
    case 4: r = lm_tanh(x[i]); break;
    case 5: r = lm_expm1(x[i]); break;
    default: break;
    }
    out[i] = r;
}

extern "C" int lemhip_math_eval(int fn, int n, const double *x, const double *y, double *out)
{
    if (n <= 0 || !x || !y || !out) return LEMHIP_E_ARG;
    double *d = nullptr;
    if (hipMalloc(&d, 3 * (size_t)n * 8) != hipSuccess) return LEMHIP_E_NOMEM;
    int rc = LEMHIP_OK;
    if (hipMemcpy(d, x, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, y, n * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = LEMHIP_E_HIP;
    if (rc == LEMHIP_OK) {
        k_math_eval<<<(n + 255) / 256, 256>>>(fn, n, d, d + n, d + 2 * n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d + 2 * n, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = LEMHIP_E_HIP;
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int lemhip_mem_info(int device_id, double *free_bytes, double *total_bytes)
{
    if (hipSetDevice(device_id) != hipSuccess) return LEMHIP_E_HIP;
    size_t fr = 0, to = 0;
    if (hipMemGetInfo(&fr, &to) != hipSuccess) return LEMHIP_E_HIP;
    if (free_bytes) *free_bytes = (double)fr;
    if (total_bytes) *total_bytes = (double)to;
    return LEMHIP_OK;
}

// what alloc_ctx allocates per cell (bulk element t = 8 or 4 bytes)
extern "C" int lemhip_bytes_per_cell(int precision, int atlas, double *fixed_per_cell, double *grow_per_cell)
{
    if (precision != LEMHIP_FP64 && precision != LEMHIP_FP32_KAHAN) return LEMHIP_E_ARG;
    const double t = precision == LEMHIP_FP64 ? 8 : 4;
    double b = 0;
    b += 8 + 8 + 1;          // dem, soilT, mask
    b += 8 * t;              // Slopes
    b += 19 * t;             // soil, vegetation, climate and per-step bulk fields
    b += 1 + 1 + 8;          // fd, SFD, fa
    b += 4 + 4 + 4 + 4;      // plateau distances (two), candidates, labels
    b += 2 * t;              // calcite terms
    b += 4 + 8 + 1 + 1;      // scan / status words, raised surface, donor bits, label slots
    if (atlas) b += 1;       // slot id
    if (precision == LEMHIP_FP32_KAHAN) b += 64;   // sediment records, two 32 B buffers
    if (fixed_per_cell) *fixed_per_cell = b;
    // edge lists: 64 B per slot (hash keys and values, pair and weight list,
    // two compaction buffers), 8 slots per label rounded up to a power of
    // two (at most 16 per label); per-label arrays 21 B with 25% headroom;
    // for at most one sink label per 16 cells
    if (grow_per_cell) *grow_per_cell = (16 * 64.0 + 1.25 * 21.0) / 16.0;
    return LEMHIP_OK;
}

// ------------------------------------------------ allocation-free math hook

// Device staging of lemhip_math_eval_noalloc: x, y and the result, at most
// LEMHIP_NOALLOC_MAX doubles each (3 * 65536 * 8 B = 1.5 MiB on every device
// the library runs on, held until the process ends).  Host code never uses
// the array's name as an address: each device's copy is looked up once with
// hipGetSymbolAddress and kept in noalloc_ptr.
static __device__ __align__(256) double scratch[3 * LEMHIP_NOALLOC_MAX];
#define NOALLOC_DEVICES 64
static std::mutex noalloc_mu;
static double *noalloc_ptr[NOALLOC_DEVICES];

// the current device's copy of `scratch` (set up on the first call there)
static int noalloc_scratch(double **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_DEVICES) return LEMHIP_E_HIP;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LEMHIP_E_HIP;
        noalloc_ptr[dev] = (double *)p;
    }
    *out = noalloc_ptr[dev];
    return LEMHIP_OK;
}

extern "C" int lemhip_math_eval_noalloc(int fn, int n, const double *x, const double *y, double *out)
{
    if (!x || !y || !out || n <= 0) return LEMHIP_E_ARG;
    if (n > LEMHIP_NOALLOC_MAX) return LEMHIP_E_SIZE;
    double *d = nullptr;
    LEM_TRY(noalloc_scratch(&d));
    // every element of the staging is written before it is read: nothing
    // carried over from an earlier call is used
    if (hipMemcpy(d, x, (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, y, (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess)
        return LEMHIP_E_HIP;
    k_math_eval<<<(n + 255) / 256, 256>>>(fn, n, d, d + n, d + 2 * n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, d + 2 * n, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return LEMHIP_E_HIP;
    return LEMHIP_OK;
}

extern "C" int lemhip_device_info(int device_id, char *name, int name_len, int *cu_count, double *mem_gb)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) return LEMHIP_E_HIP;
    if (name && name_len > 0) {
        strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (mem_gb) *mem_gb = prop.totalGlobalMem / 1e9;
    return LEMHIP_OK;
}
