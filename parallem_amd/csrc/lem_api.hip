// lem_api.hip — the extern "C" boundary (include/lemhip.h): context
// lifecycle, the reference's host-side initialisation performed on the
// device, host-layout upload/download, and the timestep phases.
#include "lem_ctx.h"
#include <stdlib.h>
#include <string.h>
#include <new>
#include <vector>

static int grid1d(long n, int block = 256, int cap = 8192)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ------------------------------------------------------------ init kernels

template <typename T>
__global__ void k_init_state(int W, int H, lem_params p, const double *__restrict__ dem,
                             const double *__restrict__ rain_in, int soilT_mode, uint8_t *mask, double *soilT,
                             T *rain, T *temp, T *last_rain, T *last_temp, T *soilM, T *stone, T *fines, T *nut,
                             T *soilB, T *TotB, T *w, T *slopeptr, T *ero, T *inci, T *geli, T *depo, T *dz, T *wC,
                             T *wP, uint8_t *fd, uint8_t *SFD, T *slopes, long N)
{
    double rainmax = p.rain * 2;                     // Data.cpp:246-250
    double rainmin = p.rain;
    double increment = (rainmax - rainmin) / H;
    double minheight = 467.228;
    double rescale = (10000 / (p.cellsize * p.cellsize));
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
        int row = (int)(x / W);
        double z = dem[x];
        bool m = !(z == -9999);                      // memory.cu:229-237
        mask[x] = m ? 1 : 0;
        double sT = (z > 0) ? ((z - 400) / 1400) * 5 : 0.0; // memory.cu:207-208
        if (soilT_mode == 1 && sT < 0.0) sT = 0.0;
        double r, t;
        if (m) {
            r = rain_in ? rain_in[x] : rainmin + (row * increment);
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
        for (int d = 0; d < 8; d++) slopes[d * N + x] = 0;   // Data.cpp:181-185
    }
}

// outlet = first strict minimum over mask cells, lowest starting at 1000 (Data.cpp:38-66)
__global__ void k_outlet_first(long N, const double *dem, const uint8_t *mask, const unsigned long long *lowest_enc,
                               int *idx)
{
    double lowest = ord_dec(*lowest_enc);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        if (mask[i] && lowest < 1000.0 && dem[i] == lowest) atomicMin(idx, (int)i);
}

__global__ void k_outlet_set(LemScalars *sc, const int *idx, const unsigned long long *lowest_enc)
{
    double lowest = ord_dec(*lowest_enc);
    sc->outlet = (lowest < 1000.0) ? *idx : 0;
}

__global__ void k_lowest_init(long N, const double *z, const uint8_t *mask, unsigned long long *out)
{
    __shared__ unsigned long long sh[256];
    unsigned long long best = ord_enc(1000.0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double v = z[i];
        if (mask[i] && v < 1000.0) {
            unsigned long long e = ord_enc(v);
            if (e < best) best = e;
        }
    }
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && sh[threadIdx.x + s] < sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(out, sh[0]);
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

// SetClimateMatrices(iteration >= 0) (Data.cpp:231-271)
template <typename T>
__global__ void k_set_climate(int W, int H, double rain_s, double temp_s, const double *dem, const uint8_t *mask,
                              T *rain, T *temp)
{
    long N = (long)W * H;
    double rainmax = rain_s * 2, rainmin = rain_s;
    double increment = (rainmax - rainmin) / H;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
        int row = (int)(x / W);
        if (mask[x]) {
            rain[x] = (T)(rainmin + (row * increment));
            temp[x] = (T)(temp_s - ((dem[x] - 467.228) * 0.0065));
        } else {
            rain[x] = (T)-9999.;
            temp[x] = (T)-9999.;
        }
    }
}

// ------------------------------------------------------------- lifecycle

static void free_all(lemhip_ctx *c)
{
    void *ptrs[] = {c->dem, c->soilT, c->mask, c->slopes, c->soilM, c->stone, c->fines, c->nut, c->soilB, c->TotB,
                    c->rain, c->temp, c->last_rain, c->last_temp, c->fd, c->SFD, c->w, c->fa, c->slopeptr, c->ero,
                    c->inci, c->geli, c->depo, c->dz, c->wC, c->wP, c->sp0, c->sp1, c->cand, c->ws, c->scan_tmp,
                    c->z2, c->pl_front0, c->pl_front1, c->pl_U, c->pl_stamp, c->pl_Uv, c->pl_ctl, c->cub_tmp, c->hkeys, c->hvals, c->elo, c->ehi, c->ew, c->rmm, c->seen, c->status, c->comp, c->cmax,
                    c->dmask, c->tile_flags, c->dag_list[0], c->dag_list[1], c->dag_list[2], c->dag_wake, c->dag_ctl, c->partials, c->scal};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->hint) (void)hipHostFree(c->hint);
    if (c->prof_ready)
        for (int k = 0; k < PROF_COUNT; k++)
            for (int q = 0; q < PROF_MAXP; q++) {
                (void)hipEventDestroy(c->prof[k].a[q]);
                (void)hipEventDestroy(c->prof[k].b[q]);
            }
    if (c->s) (void)hipStreamDestroy(c->s);
}

extern "C" int lemhip_create(const lem_params *p, int precision, int device_id, lemhip_ctx **out)
{
    if (!p || !out) return LEMHIP_E_ARG;
    *out = nullptr;
    if (p->width <= 0 || p->height <= 0 || p->cellsize <= 0) return LEMHIP_E_ARG;
    if (precision != LEMHIP_FP64 && precision != LEMHIP_FP32_KAHAN) return LEMHIP_E_ARG;
    if ((long)p->width * p->height >= 0x7FFFFFFFL) return LEMHIP_E_ARG;  // labels and pointers are int32
    lemhip_ctx *c = new (std::nothrow) lemhip_ctx();
    if (!c) return LEMHIP_E_NOMEM;
    c->p = *p;
    c->prec = precision;
    c->dev = device_id;
    c->W = p->width;
    c->H = p->height;
    c->N = (long)p->width * p->height;
    c->tsz = precision == LEMHIP_FP64 ? 8 : 4;
    *out = c;
    if (hipSetDevice(device_id) != hipSuccess) { c->err = "hipSetDevice failed"; return LEMHIP_E_HIP; }
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
    LEM_TRY(A((void **)&c->pl_front0, N * 4));
    LEM_TRY(A((void **)&c->pl_front1, N * 4));
    LEM_TRY(A((void **)&c->pl_U, N * 4));
    LEM_TRY(A((void **)&c->pl_Uv, N * 4));
    LEM_TRY(A((void **)&c->pl_stamp, N * 4));
    LEM_TRY(A(&c->pl_ctl, 256));
    LEM_CHECK(c, hipMemset(c->pl_stamp, 0, N * 4));
    LEM_TRY(A((void **)&c->cand, N * 4));
    LEM_TRY(A((void **)&c->ws, N * 4));
    LEM_TRY(A((void **)&c->scan_tmp, N * 4));
    LEM_TRY(A((void **)&c->z2, N * 8));
    LEM_TRY(A((void **)&c->dmask, N));
    long ntiles = ((c->W + 63) / 64) * (long)((c->H + 15) / 16);
    LEM_TRY(A((void **)&c->tile_flags, ntiles * 4));
    for (int q = 0; q < 3; q++) LEM_TRY(A((void **)&c->dag_list[q], ntiles * 4));
    LEM_TRY(A((void **)&c->dag_wake, ntiles * 4));
    LEM_TRY(A(&c->dag_ctl, 256));
    LEM_CHECK(c, hipMemset(c->dag_wake, 0, ntiles * 4));
    LEM_TRY(A((void **)&c->partials, 4096 * 8));
    LEM_TRY(A((void **)&c->scal, sizeof(LemScalars)));
    LEM_CHECK(c, hipMemset(c->scal, 0, sizeof(LemScalars)));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hscal, sizeof(LemScalars), hipHostMallocDefault));
    LEM_CHECK(c, hipHostMalloc((void **)&c->hint, 64 * sizeof(int), hipHostMallocDefault));
    memset(c->hscal, 0, sizeof(LemScalars));
    return LEMHIP_OK;
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

template <typename T>
static int init_state_t(lemhip_ctx *c, const double *dem, const double *rain, int soilT_mode)
{
    long N = c->N;
    LEM_CHECK(c, hipMemcpyAsync(c->dem, dem, N * 8, hipMemcpyHostToDevice, c->s));
    double *rain_d = nullptr;
    if (rain) {
        rain_d = c->z2;
        LEM_CHECK(c, hipMemcpyAsync(rain_d, rain, N * 8, hipMemcpyHostToDevice, c->s));
    }
    k_init_state<T><<<grid1d(N), 256, 0, c->s>>>(
        c->W, c->H, c->p, c->dem, rain_d, soilT_mode, c->mask, c->soilT, (T *)c->rain, (T *)c->temp,
        (T *)c->last_rain, (T *)c->last_temp, (T *)c->soilM, (T *)c->stone, (T *)c->fines, (T *)c->nut,
        (T *)c->soilB, (T *)c->TotB, (T *)c->w, (T *)c->slopeptr, (T *)c->ero, (T *)c->inci, (T *)c->geli,
        (T *)c->depo, (T *)c->dz, (T *)c->wC, (T *)c->wP, c->fd, c->SFD, (T *)c->slopes, N);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipMemsetAsync(c->fa, 0, N * 8, c->s));
    unsigned long long *low = reinterpret_cast<unsigned long long *>(c->partials);
    int *idx = reinterpret_cast<int *>(c->partials + 1);
    unsigned long long init = ord_enc(1000.0);
    int big = 0x7FFFFFFF;
    LEM_CHECK(c, hipMemcpyAsync(low, &init, 8, hipMemcpyHostToDevice, c->s));
    LEM_CHECK(c, hipMemcpyAsync(idx, &big, 4, hipMemcpyHostToDevice, c->s));
    k_lowest_init<<<grid1d(N, 256, 2048), 256, 0, c->s>>>(N, c->dem, c->mask, low);
    k_outlet_first<<<grid1d(N), 256, 0, c->s>>>(N, c->dem, c->mask, low, idx);
    k_outlet_set<<<1, 1, 0, c->s>>>(c->scal, idx, low);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_init_state(lemhip_ctx *c, const double *dem, const double *rain, int soilT_mode)
{
    if (!c || !dem) return LEMHIP_E_ARG;
    if (soilT_mode != 0 && soilT_mode != 1) { c->err = "soilT_mode must be 0 or 1"; return LEMHIP_E_ARG; }
    return c->prec == LEMHIP_FP64 ? init_state_t<double>(c, dem, rain, soilT_mode)
                                  : init_state_t<float>(c, dem, rain, soilT_mode);
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

extern "C" int lemhip_download(lemhip_ctx *c, int field, void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    int kind = -1;
    void *src = field_ptr(c, field, &kind);
    if (!src) { c->err = "unknown field"; return LEMHIP_E_ARG; }
    long N = c->N;
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (kind == 1 || (kind == 0 && c->prec == LEMHIP_FP64) || (kind == 5 && c->prec == LEMHIP_FP64)) {
        LEM_CHECK(c, hipMemcpyAsync(host, src, N * 8, hipMemcpyDeviceToHost, c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        return LEMHIP_OK;
    }
    size_t bytes = (kind == 4) ? 8 * N * c->tsz : (kind == 5 ? N * 8 : (kind == 0 ? N * c->tsz : N));
    LEM_TRY(stage_reserve(c, bytes));
    LEM_CHECK(c, hipMemcpyAsync(c->stage, src, bytes, hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    if (kind == 0) {
        const float *s = (const float *)c->stage;
        double *d = (double *)host;
        for (long i = 0; i < N; i++) d[i] = s[i];
    } else if (kind == 5) {
        const float *s = (const float *)c->stage;
        double *d = (double *)host;
        for (long i = 0; i < N; i++) d[i] = (double)s[2 * i] + (double)s[2 * i + 1];
    } else if (kind == 2 || kind == 3) {
        const uint8_t *s = (const uint8_t *)c->stage;
        int *d = (int *)host;
        if (kind == 3) {
            for (long i = 0; i < N; i++) d[i] = s[i];
        } else {
            // nodata cells carry -9999 in the reference's fd/SFD (newflow.cu:74-78)
            std::vector<uint8_t> m(N);
            LEM_CHECK(c, hipMemcpy(m.data(), c->mask, N, hipMemcpyDeviceToHost));
            for (long i = 0; i < N; i++) d[i] = m[i] ? (int)s[i] : -9999;
        }
    } else if (kind == 4) {
        double *d = (double *)host;
        if (c->prec == LEMHIP_FP64) {
            const double *s = (const double *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < N; i++) d[8 * i + k] = s[k * N + i];
        } else {
            const float *s = (const float *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < N; i++) d[8 * i + k] = s[k * N + i];
        }
    }
    return LEMHIP_OK;
}

extern "C" int lemhip_upload(lemhip_ctx *c, int field, const void *host)
{
    if (!c || !host) return LEMHIP_E_ARG;
    int kind = -1;
    void *dst = field_ptr(c, field, &kind);
    if (!dst) { c->err = "unknown field"; return LEMHIP_E_ARG; }
    long N = c->N;
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (kind == 1 || (kind == 0 && c->prec == LEMHIP_FP64) || (kind == 5 && c->prec == LEMHIP_FP64)) {
        LEM_CHECK(c, hipMemcpyAsync(dst, host, N * 8, hipMemcpyHostToDevice, c->s));
        LEM_CHECK(c, hipStreamSynchronize(c->s));
        return LEMHIP_OK;
    }
    size_t bytes = (kind == 4) ? 8 * N * c->tsz : (kind == 5 ? N * 8 : (kind == 0 ? N * c->tsz : N));
    LEM_TRY(stage_reserve(c, bytes));
    if (kind == 0) {
        float *d = (float *)c->stage;
        const double *s = (const double *)host;
        for (long i = 0; i < N; i++) d[i] = (float)s[i];
    } else if (kind == 5) {
        float *d = (float *)c->stage;
        const double *s = (const double *)host;
        for (long i = 0; i < N; i++) {
            float hi = (float)s[i];
            d[2 * i] = hi;
            d[2 * i + 1] = (float)(s[i] - (double)hi);
        }
    } else if (kind == 2 || kind == 3) {
        uint8_t *d = (uint8_t *)c->stage;
        const int *s = (const int *)host;
        for (long i = 0; i < N; i++) d[i] = (uint8_t)(s[i] < 0 ? 0 : s[i]);
    } else if (kind == 4) {
        const double *s = (const double *)host;
        if (c->prec == LEMHIP_FP64) {
            double *d = (double *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < N; i++) d[k * N + i] = s[8 * i + k];
        } else {
            float *d = (float *)c->stage;
            for (int k = 0; k < 8; k++)
                for (long i = 0; i < N; i++) d[k * N + i] = (float)s[8 * i + k];
        }
    }
    LEM_CHECK(c, hipMemcpyAsync(dst, c->stage, bytes, hipMemcpyHostToDevice, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

extern "C" int lemhip_get_outlet(lemhip_ctx *c, int *outlet)
{
    if (!c || !outlet) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipMemcpyAsync(&c->hscal->outlet, &c->scal->outlet, sizeof(int), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    *outlet = c->hscal->outlet;
    return LEMHIP_OK;
}

extern "C" int lemhip_set_outlet(lemhip_ctx *c, int outlet)
{
    if (!c || outlet < 0 || outlet >= c->N) return LEMHIP_E_ARG;
    c->hscal->outlet = outlet;
    LEM_CHECK(c, hipMemcpyAsync(&c->scal->outlet, &c->hscal->outlet, sizeof(int), hipMemcpyHostToDevice, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

// --------------------------------------------------------------- phases

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

extern "C" int lemhip_step(lemhip_ctx *c, int iter)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    prof_begin(c, PROF_STEP);
    LEM_TRY(lem_fd_phase(c, iter));
    LEM_TRY(lem_runoff_phase(c));
    LEM_TRY(lem_fa_phase(c, iter));
    LEM_TRY(lem_mod_phase(c, iter));
    prof_end(c, PROF_STEP);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    LEM_TRY(prof_resolve(c));
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

extern "C" int lemhip_set_climate(lemhip_ctx *c, double rain, double temp)
{
    if (!c) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipSetDevice(c->dev));
    if (c->prec == LEMHIP_FP64)
        k_set_climate<double><<<grid1d(c->N), 256, 0, c->s>>>(c->W, c->H, rain, temp, c->dem, c->mask,
                                                              (double *)c->rain, (double *)c->temp);
    else
        k_set_climate<float><<<grid1d(c->N), 256, 0, c->s>>>(c->W, c->H, rain, temp, c->dem, c->mask,
                                                             (float *)c->rain, (float *)c->temp);
    LEM_LAUNCHED(c);
    LEM_CHECK(c, hipStreamSynchronize(c->s));
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

extern "C" int lemhip_summary(lemhip_ctx *c, lem_summary *o)
{
    if (!c || !o) return LEMHIP_E_ARG;
    LEM_CHECK(c, hipMemcpyAsync(c->hscal, c->scal, sizeof(LemScalars), hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    const LemScalars *h = c->hscal;
    o->FA_max = h->FA_max;
    o->totE = h->totE; o->totI = h->totI; o->totG = h->totG; o->totD = h->totD;
    o->totweatherC = h->totwC; o->totweatherP = h->totwP; o->totbio = h->totbio;
    o->outletcellidx = h->outlet;
    o->sinkcount = c->h_sinkcount;
    o->sinks_pass1 = c->h_sinks1;
    o->nlabels = c->h_nlab;
    o->nedges = c->h_nedges;
    o->plateau_rounds1 = c->h_rounds1;
    o->plateau_rounds2 = c->h_rounds2;
    o->fa_uncomputed = c->h_fa_unc;
    o->sed_uncomputed = c->h_sed_unc;
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
