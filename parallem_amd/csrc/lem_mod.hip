// lem_mod.hip — runoff weights (runoffweight.cu:21-138), the elementwise
// erosion/weathering/update kernels (eroincidep.cu, updates.cu) fused per
// cell in the reference's order, the sediment pass (lem_dag.hip) between
// them, and deterministic reductions for the per-step totals.
#include "lem_ctx.h"
#include <stddef.h>

template <typename T> int lem_fa_dag(lemhip_ctx *c);
template <typename T> int lem_sed_dag(lemhip_ctx *c);

// ---------------------------------------------------------------- math

// FP64: the shared portable routines (bit-identical with the host build);
// FP32 form: the device library's single-precision functions.
__device__ __forceinline__ double mexp(double x) { return lm_exp(x); }
__device__ __forceinline__ double mlog(double x) { return lm_log(x); }
__device__ __forceinline__ double mlog10(double x) { return lm_log10(x); }
__device__ __forceinline__ double mpow(double x, double y) { return lm_pow(x, y); }
__device__ __forceinline__ double mtanh(double x) { return lm_tanh(x); }
__device__ __forceinline__ float mexp(float x) { return expf(x); }
__device__ __forceinline__ float mlog(float x) { return logf(x); }
__device__ __forceinline__ float mlog10(float x) { return log10f(x); }
__device__ __forceinline__ float mpow(float x, float y) { return y == 2.0f ? x * x : powf(x, y); }
__device__ __forceinline__ float mtanh(float x) { return tanhf(x); }

template <typename T> struct FaS;
template <> struct FaS<double> { using type = double; };
template <> struct FaS<float> { using type = float2; };
// flow accumulation as one double: the FP32 form's float pair (hi, lo) holds
// ~48 significant bits, and every quantity MOD derives from it (overland flow
// and its cap, the incision threshold, the gelifluction cut-off, surface flow
// against evapotranspiration) is formed from the full pair in double, so a
// threshold decides the same branch as in the FP64 form
__device__ __forceinline__ double fa_full(double v) { return v; }
__device__ __forceinline__ double fa_full(float2 v) { return (double)v.x + (double)v.y; }
__device__ __forceinline__ void fa_set(double *p, double v) { *p = v; }
__device__ __forceinline__ void fa_set(float2 *p, float v) { *p = make_float2(v, 0.f); }

static int grid1d(long n, int block = 256, int cap = 8192)
{
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ---------------------------------------------------------------- runoff

// calcrunoffweight (runoffweight.cu:21-58); also seeds fa = w on mask cells,
// the value a cell keeps if it never completes (mfd_simple.cu:72)
// bad: raised when a mask cell's weight is negative or not finite (the FP32
// form's integer accumulation needs weights in [0, inf))
template <typename T>
__global__ void k_runoff(long N, const uint8_t *__restrict__ mask, const T *__restrict__ rain,
                         const T *__restrict__ stone, const T *__restrict__ TotB, T *__restrict__ soilM,
                         T *__restrict__ w, typename FaS<T>::type *__restrict__ fa, int *bad)
{
    bool b = false;
    for (long self = blockIdx.x * (long)blockDim.x + threadIdx.x; self < N; self += (long)gridDim.x * blockDim.x) {
        const uint8_t m = mask[self];   // every input in flight before the mask test
        const T st0 = stone[self], tb0 = TotB[self], sm0 = soilM[self], rn = rain[self];
        if (m == 0) {
            w[self] = (T)-9999;
            fa_set(&fa[self], (T)0);
            continue;
        }
        T runoff_coeff = ((T)1.0 - (T)0.0085 * (st0)) * mexp((T)-0.0025 * (tb0) / (T)76.5)
                         * mtanh((T)0.6 + (sm0) / (T)0.2);
        T myppt = rn / (T)1000;
        T wv = myppt * runoff_coeff;
        w[self] = wv;
        T sm = sm0;
        sm += myppt * ((T)1. - runoff_coeff);
        if (sm > (T)0.5) sm = (T)0.5;
        soilM[self] = sm;
        fa_set(&fa[self], wv);
        b |= !(wv >= (T)0) || isinf(wv);
    }
    if (__any(b) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

// ------------------------------------------------ MOD part 1 (pre-sediment)

// aveslope (eroincidep.cu:8-37), diff_erosion (:39-87), conc_erosion
// (:89-152), gelifluction (:154-198); depo = 1e-5 seed (mfd_accum.cu:390-394)
template <typename T>
__global__ void __launch_bounds__(256) k_mod1(long N, long plane, lem_params p, const uint8_t *__restrict__ mask,
                                              const T *__restrict__ slopes, const typename FaS<T>::type *__restrict__ fa,
                                              const T *__restrict__ fines, const T *__restrict__ nut,
                                              const T *__restrict__ stone, const T *__restrict__ TotB,
                                              const double *__restrict__ soilT, const T *__restrict__ temp,
                                              T *__restrict__ slopeptr, T *__restrict__ ero, T *__restrict__ inci,
                                              T *__restrict__ geli, T *__restrict__ depo)
{
    const double rescale_c = 1 / (10000 / (p.cellsize * p.cellsize));
    for (long self = blockIdx.x * (long)blockDim.x + threadIdx.x; self < N; self += (long)gridDim.x * blockDim.x) {
        // every input of the cell in flight before the mask test
        const uint8_t m = mask[self];
        T s8[8];
#pragma unroll
        for (int j = 0; j < 8; j++) s8[j] = slopes[j * plane + self];
        const double sTd = soilT[self];
        const typename FaS<T>::type fav = fa[self];
        const T tb0 = TotB[self], fi0 = fines[self], nu0 = nut[self], st0 = stone[self], tc = temp[self];
        if (m == 0) {
            slopeptr[self] = 0; ero[self] = 0; inci[self] = 0; geli[self] = 0; depo[self] = 0;
            continue;
        }
        T total = 0;
        int nonzero = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            T sj = s8[j];
            total += sj;
            if (sj > 0) nonzero++;
        }
        T sp = 0;
        if ((total > 0) && (nonzero > 0)) sp = total / (T)nonzero;
        slopeptr[self] = sp;
        T sT = (T)sTd;
        const double fad = fa_full(fav);
        T ex = mexp((T)-0.07 * (tb0) / (T)76.5);
        // diffuse
        T erodibility = ((T)2.77e-6 * mpow(fi0, (T)1.14) * ((T)12. - nu0))
                        + ((T)0.043 * ((T)p.soil_struct - (T)2.)) + ((T)0.033 * ((T)p.profile_permeability - (T)3.));
        if (erodibility < (T)0.0) erodibility = 0;
        double OFd = ((fad) * 1000) / 25;
        if (OFd > p.Qthreshold) OFd = p.Qthreshold;
        const T OF = (T)OFd;
        T slope = sp;
        if (slope == (T)0.0) slope = (T)0.0001;
        if (slope > (T)5.0) slope = (T)5.0;
        T answer = ((erodibility * OF * OF * mpow(slope, (T)1.67) *
                     (mpow(((T)110. - (st0)), (T)0.877) / (T)61.7023))) * ex;
        if (answer > sT) answer = sT;
        ero[self] = answer;
        // concentrated
        T cslope = sp;
        if (cslope > (T)0.5) cslope = (T)0.5;
        T max_incision = (T)p.max_incision;
        const double ftd = ((fad) * rescale_c) - p.Qthreshold;
        T f_temp = (T)ftd;
        T iv;
        if (ftd < 0.0) { f_temp = 0; iv = 0; }
        else {
            if (cslope == (T)0.0) iv = (T)p.Flow_A * mpow(f_temp, (T)p.Flow_B) * (T)0.01 * ex;
            else iv = (T)p.Flow_A * mpow(f_temp, (T)p.Flow_B) * (cslope) * ex;
        }
        if (iv > max_incision) iv = max_incision;
        if (iv > sT) {
            T softrockinc = sT;
            T bedrockinc = ((T)1.0 - (softrockinc / iv)) * (T)0.0001 * mpow(f_temp, (T)p.Flow_B) * (cslope) * ex;
            // FP32 form: a discharge past the float range makes the power
            // inf, and a zero slope must still give a zero term (the FP64
            // form's power stays finite there)
            if (sizeof(T) == 4 && cslope == (T)0) bedrockinc = 0;
            iv = softrockinc + bedrockinc;
            if (iv > max_incision) iv = max_incision;
        }
        inci[self] = iv;
        // gelifluction
        T kappa = (T)p.kappa;
        if (tc >= (T)-20. && tc <= (T)8.87)
            kappa = (T)0.15711 - (T)0.011835 * tc - (T)0.001195 * tc * tc + (T)0.000038 * tc * tc * tc
                    + (T)0.000002 * tc * tc * tc * tc;
        double scd = 1 - (fad / p.Gflowlim);
        if (fad > p.Gflowlim) scd = 0.0;
        const T scaler = (T)scd;
        T slopethreshold = sp;
        if (slopethreshold > (T)p.Gslope) slopethreshold = (T)p.Gslope;
        T g = kappa * slopethreshold * scaler;
        if ((answer + g) > sT) g = sT - answer;
        if (g < (T)0.0) g = 0;
        geli[self] = g;
        depo[self] = (T)0.00001;
    }
}

// the exact sub-expression order of the reference's diff_erosion with the
// FP64 routines: (((k*OF)*OF)*pow(s,1.67))*(pow(110-stone,0.877)/61.7023)*exp
// is what the template above evaluates for T = double, with ex computed once
// (exp(-0.07*TotB/76.5) is the same value at each of its call sites).

// ------------------------------------------------ MOD part 2 (post-sediment)

// surface_change (eroincidep.cu:200-217), weathering type 1 (:220-327),
// Update_Surface (updates.cu:23-46), update_nutrient_content (:48-76) and
// grow_vegetation (:78-139): each reads only its own cell, so one pass keeps
// the reference's per-cell order.
// The calcite equilibrium term of weathering type 1 (eroincidep.cu:220-327)
// depends only on a cell's temperature and on which side of 0.1 its soil
// thickness lies: both variants are computed here, with the same expressions
// as the body written in one piece, whenever the temperatures change
// (initialisation, climate updates, uploads), and MOD part 2 reads them.
template <typename T>
__global__ void __launch_bounds__(256) k_calcite_terms(long N, const uint8_t *__restrict__ mask,
                                                       const T *__restrict__ temp, T *__restrict__ ca_hi,
                                                       T *__restrict__ ca_lo)
{
    for (long self = blockIdx.x * (long)blockDim.x + threadIdx.x; self < N; self += (long)gridDim.x * blockDim.x) {
        if (mask[self] == 0) continue;
        T tc = temp[self];
        T t_abs = tc + (T)273.16;
        T t2 = t_abs * t_abs;
        T ltemp = mlog10(t_abs);
        T DHA = (T)-0.4883 + (T)8.074e-4 * tc;
        T DHB = (T)-0.3241 + (T)1.600e-4 * tc;
        T sqI = (T)0.31622776601683794; // sqrt(0.1), correctly rounded
        T denom1 = ((T)-4. * DHA * sqI) / ((T)1. + (T)5.e-8 * DHB * sqI);
        T gammCa_d = mpow((T)10., denom1);
        T gammHCO3_d = mpow((T)10., (-DHA * sqI) / ((T)1. + (T)5.4e-8 * DHB * sqI));
        T K1_d = mpow((T)10., (T)-356.3094 - (T)0.06091964 * t_abs + (T)21834.37 / t_abs + (T)126.8339 * ltemp - (T)1684915 / t2);
        T K2_d = mpow((T)10., (T)-107.8871 - (T)0.03252849 * t_abs + (T)5151.79 / t_abs + (T)38.92561 * ltemp - (T)563713.9 / t2);
        T KC_d = mpow((T)10., (T)-171.9065 - (T)0.077993 * t_abs + (T)2839.319 / t_abs + (T)71.595 * ltemp);
        T KH_d = mpow((T)10., (T)108.3865 + (T)0.01985076 * t_abs - (T)6919.53 / t_abs - (T)40.45154 * ltemp + (T)669365.0 / t2);
        // PCO2 for soil thicker than 0.1, and for the rest
        T PCO2hi = mpow((T)10., (T)-2.5 + ((tc - (T)15.) / (T)10.));
        T PCO2lo = mpow((T)10., (T)-3.5 + ((tc - (T)15.) / (T)10.));
        ca_hi[self] = mpow(PCO2hi * ((K1_d * KC_d * KH_d) / ((T)4. * K2_d * gammCa_d * gammHCO3_d * gammHCO3_d)), ((T)1. / (T)3.));
        ca_lo[self] = mpow(PCO2lo * ((K1_d * KC_d * KH_d) / ((T)4. * K2_d * gammCa_d * gammHCO3_d * gammHCO3_d)), ((T)1. / (T)3.));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_mod2(long N, lem_params p, const uint8_t *__restrict__ mask,
                                              const typename FaS<T>::type *__restrict__ fa,
                                              const T *__restrict__ ero, const T *__restrict__ inci,
                                              const T *__restrict__ geli, const T *__restrict__ depo,
                                              const T *__restrict__ temp, const T *__restrict__ rain,
                                              const T *__restrict__ last_temp, double *__restrict__ soilT,
                                              T *__restrict__ stone, T *__restrict__ fines, T *__restrict__ soilM,
                                              T *__restrict__ soilB, T *__restrict__ nut, T *__restrict__ TotB,
                                              double *__restrict__ dem, T *__restrict__ dzo, T *__restrict__ wCo,
                                              T *__restrict__ wPo, const uint8_t *__restrict__ demid,
                                              const DemScal *__restrict__ ds, const T *__restrict__ ca_hi,
                                              const T *__restrict__ ca_lo)
{
    const double rescale_w = 10000 * (10000 / (p.cellsize * p.cellsize));
    const double rescale_v = (10000 / (p.cellsize * p.cellsize));
    const double base0 = ds[0].base_level;
    for (long self = blockIdx.x * (long)blockDim.x + threadIdx.x; self < N; self += (long)gridDim.x * blockDim.x) {
        T tc = temp[self];
        if (mask[self] == 0) {
            dzo[self] = 0; wCo[self] = 0; wPo[self] = 0; TotB[self] = 0;
            continue;
        }
        T dz = (depo[self] - ero[self] - inci[self] - geli[self]);
        dzo[self] = dz;
        // weathering
        T rn = rain[self];
        // surface flow and evapotranspiration in double in both forms (FP64:
        // the reference's own expressions; FP32: the bulk inputs widened)
        double PE_d;
        if (sizeof(T) == 8) PE_d = (double)((T)1.44 * mpow(tc + (T)6., (T)2));
        else PE_d = 1.44 * (((double)tc + 6.) * ((double)tc + 6.));
        T P0_d = (T)7.7e-5, P1_d = (T)2.3;
        const double surface_flow = fa_full(fa[self]) * 1000;
        const double rnd = (double)rn;
        double evap_trans;
        if (rnd + surface_flow <= PE_d) evap_trans = rnd + surface_flow;
        else evap_trans = rnd + surface_flow - PE_d;
        double sT = soilT[self];
        T Ca2eq = (sT > 0.1) ? ca_hi[self] : ca_lo[self];
        double wCd = 0.0148148148 * (double)Ca2eq * (surface_flow / rescale_w);
        if (wCd <= 0.0) wCd = 0;
        if (wCd > 0.001) wCd = 0.001;
        const T wC = (T)wCd;
        T wP = P0_d * mexp(-(P1_d) * (T)sT);
        if (sizeof(T) == 8) wP = (T)(7.7e-5 * lm_exp(-(2.3) * sT));
        wCo[self] = wC;
        wPo[self] = wP;
        sT = sT + (double)wC * 0.02 + (double)wP;
        T st = stone[self], fi = fines[self], sm = soilM[self];
        if (sT <= 0.0) {
            sT = 0.0; st = 0; fi = 0; sm = (T)0.02;
        } else {
            T ratio = (T)((double)ero[self] / sT);
            if (wP > (T)0) st += st * (mtanh(ratio) * -mtanh(wC / wP));
            else st += st * mtanh(ratio);
            if (st < (T)0.01) st = (T)0.01;
            else if (st > (T)99.9) st = (T)99.9;
            fi = (T)100. - st;
            sm -= (T)((1.e-3 * evap_trans) / sT);
            if (sm < (T)0.02) sm = (T)0.02;
            else if (sm > (T)0.5) sm = (T)0.5;
        }
        soilT[self] = sT;
        stone[self] = st;
        fines[self] = fi;
        soilM[self] = sm;
        // update surface (FP64 in every form)
        const double base_level = demid ? ds[demid[self]].base_level : base0;
        double z = dem[self];
        z = z + (double)dz - (double)wC - (double)wP;
        if (z <= base_level) z = base_level + 0.01;
        dem[self] = z;
        // nutrients
        T tb = TotB[self], sb = soilB[self], nu = nut[self];
        if (sT <= 0.0) nu = 0;
        T leaf_fall = tb / ((T)72.8 * mlog(tb + (T)1.4));
        T decomposition = sb * mexp((T)0.12 * tc - (T)5.75);
        sb = sb + leaf_fall - decomposition;
        if (sT == 0.0 && sb > (T)0.0) nu = (T)100.0;
        else if (sT == 0.0) nu = 0;
        else nu = (T)(((double)sb / (1300. * sT)) * 100.);
        if (nu > (T)100) nu = (T)100.;
        soilB[self] = sb;
        nut[self] = nu;
        // vegetation
        T delta_temp = tc - last_temp[self];
        T tempeffect = (T)100 + ((T)100 / (tc + (T)15));
        T total_cc = ((T)3000.0 + ((T)5.0 * (rn - tempeffect))) / (T)rescale_v;
        if (delta_temp > (T)0) {
            T r_plus = delta_temp * ((T)1.0 + mtanh((tc - (T)p.temp_inflex) / (T)p.temp_sens));
            tb += ((r_plus * tb) * ((T)1. - (tb / total_cc)));
        }
        if (delta_temp < (T)0) {
            T r_minus = delta_temp * ((T)1.0 - mtanh((tc - (T)p.temp_inflex) / (T)p.temp_sens));
            tb += ((r_minus * tb) * (tb / total_cc));
        }
        if (delta_temp == (T)0) tb = total_cc;
        if (tb < (T)1) tb = (T)1.0;
        TotB[self] = tb;
    }
}

// ------------------------------------------------ per-slot reductions
//
// Grid (slot_kbmax, ns): block (k, j) covers rows y0+k, y0+k+kb, ... of slot
// s0+j with kb = slot_kb(h); blocks k >= kb have nothing to do.  Sums go
// through per-block partials combined in a fixed order, so totals are
// reproducible; min/max/first-index are exact under any order.

template <typename F>
__device__ __forceinline__ void slot_rows(const DemDesc &D, int W, F f)
{
    int kb = slot_kb(D.h);
    for (int r = D.y0 + blockIdx.x; r < D.y0 + D.h; r += kb)
        for (int c = D.x0 + threadIdx.x; c < D.x0 + D.w; c += blockDim.x) f((long)r * W + c);
}

// The same cells in the same order per thread, with the loads of U column
// steps in flight before the first of them is used (one wait per U steps
// instead of one per step)
template <int U, typename Ld, typename F>
__device__ __forceinline__ void slot_rows_u(const DemDesc &D, int W, Ld load, F f)
{
    using V = decltype(load(0L));
    const int kb = slot_kb(D.h), c1 = D.x0 + D.w;
    for (int r = D.y0 + blockIdx.x; r < D.y0 + D.h; r += kb) {
        const long row = (long)r * W;
        for (int c0 = D.x0 + threadIdx.x; c0 < c1; c0 += U * blockDim.x) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = c0 + u * blockDim.x;
                v[u] = c < c1 ? load(row + c) : V{};
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (c0 + u * (int)blockDim.x < c1) f(v[u]);
        }
    }
}

__device__ __forceinline__ bool slot_block_idle(const DemDesc &D) { return D.w == 0 || (int)blockIdx.x >= slot_kb(D.h); }

__device__ __forceinline__ double block_sum_f64(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
        if (threadIdx.x < t) sh[threadIdx.x] += sh[threadIdx.x + t];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// up to three fields summed in one pass; partial[(j*3 + q)*SLOT_KB + k];
// with a mask, cells outside it are skipped
template <typename T>
__global__ void __launch_bounds__(256) k_slot_sum3(const T *__restrict__ a0, const T *__restrict__ a1,
                                                   const T *__restrict__ a2, const uint8_t *__restrict__ mask,
                                                   int W, const DemDesc *dd, double *partial)
{
    __shared__ double sh[256];
    int j = blockIdx.y;
    DemDesc D = dd[j];
    if (slot_block_idle(D)) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    struct V3 {
        uint32_t m;
        T v0, v1, v2;
    };
    slot_rows_u<4>(
        D, W,
        [&](long i) {
            return V3{mask ? (uint32_t)mask[i] : 1u, a0[i], a1 ? a1[i] : (T)0, a2 ? a2[i] : (T)0};
        },
        [&](const V3 &v) {
            if (!v.m) return;
            s0 += (double)v.v0;
            if (a1) s1 += (double)v.v1;
            if (a2) s2 += (double)v.v2;
        });
    double *P = partial + (long)j * 3 * SLOT_KB + blockIdx.x;
    s0 = block_sum_f64(s0, sh);
    if (threadIdx.x == 0) P[0] = s0;
    if (a1) {
        s1 = block_sum_f64(s1, sh);
        if (threadIdx.x == 0) P[SLOT_KB] = s1;
    }
    if (a2) {
        s2 = block_sum_f64(s2, sh);
        if (threadIdx.x == 0) P[2 * SLOT_KB] = s2;
    }
}

__global__ void __launch_bounds__(256) k_slot_sum_final(const double *partial, int na, const DemDesc *dd, DemScal *ds,
                                                        int off0, int off1, int off2)
{
    __shared__ double sh[256];
    int j = blockIdx.x;
    DemDesc D = dd[j];
    if (D.w == 0) return;
    int kb = slot_kb(D.h);
    for (int q = 0; q < na; q++) {
        const double *P = partial + ((long)j * 3 + q) * SLOT_KB;
        double acc = 0.0;
        for (int k = threadIdx.x; k < kb; k += blockDim.x) acc += P[k];
        acc = block_sum_f64(acc, sh);
        int off = q == 0 ? off0 : (q == 1 ? off1 : off2);
        if (threadIdx.x == 0) *reinterpret_cast<double *>(reinterpret_cast<char *>(&ds[j]) + off) = acc;
    }
}

template <typename T>
static int slot_sum(lemhip_ctx *c, const void *a0, int off0, const void *a1 = nullptr, int off1 = 0,
                    const void *a2 = nullptr, int off2 = 0, const uint8_t *mask = nullptr)
{
    dim3 g(c->slot_kbmax, c->nslot);
    k_slot_sum3<T><<<g, 256, 0, c->s>>>(reinterpret_cast<const T *>(a0), reinterpret_cast<const T *>(a1),
                                        reinterpret_cast<const T *>(a2), mask, c->W, c->ddesc, c->partials);
    int na = a2 ? 3 : (a1 ? 2 : 1);
    k_slot_sum_final<<<c->nslot, 256, 0, c->s>>>(c->partials, na, c->ddesc, c->dsc, off0, off1, off2);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

__device__ __forceinline__ unsigned long long block_min_u64(unsigned long long v, unsigned long long *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
        if (threadIdx.x < t && sh[threadIdx.x + t] < sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + t];
        __syncthreads();
    }
    return sh[0];
}

__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v, unsigned long long *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
        if (threadIdx.x < t && sh[threadIdx.x + t] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + t];
        __syncthreads();
    }
    return sh[0];
}

struct ZM {
    double z;
    uint32_t m;
};

// lowest mask cell with the reference's start value 1000 and strict '<' (NaN never wins)
__global__ void __launch_bounds__(256) k_slot_lowest(const double *__restrict__ z, const uint8_t *__restrict__ mask,
                                                     int W, const DemDesc *dd, int s0, DemScal *ds)
{
    __shared__ unsigned long long sh[256];
    int s = s0 + blockIdx.y;
    DemDesc D = dd[s];
    if (slot_block_idle(D)) return;
    unsigned long long best = ord_enc(1000.0);
    if (D.w > 0)
        slot_rows_u<4>(
            D, W, [&](long i) { return ZM{z[i], (uint32_t)mask[i]}; },
            [&](const ZM &q) {
                if (q.m && q.z < 1000.0) {
                    unsigned long long e = ord_enc(q.z);
                    if (e < best) best = e;
                }
            });
    best = block_min_u64(best, sh);
    if (threadIdx.x == 0 && D.w > 0) atomicMin(&ds[s].lowest_enc, best);
}

__global__ void __launch_bounds__(256) k_slot_first_lowest(const double *__restrict__ z,
                                                           const uint8_t *__restrict__ mask, int W, const DemDesc *dd,
                                                           int s0, DemScal *ds)
{
    int s = s0 + blockIdx.y;
    DemDesc D = dd[s];
    if (slot_block_idle(D)) return;
    double lowest = ord_dec(ds[s].lowest_enc);
    if (!(lowest < 1000.0)) return;
    slot_rows(D, W, [&](long i) {
        if (mask[i] && z[i] == lowest) atomicMin(&ds[s].argidx, (int)i);
    });
}

__global__ void k_slots_lowest_init(int s0, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    ds[s0 + j].lowest_enc = ord_enc(1000.0);
    ds[s0 + j].argidx = 0x7FFFFFFF;
}

// Data.cpp:38-66: first strict minimum, the DEM's cell 0 when none is below 1000
__global__ void k_slots_outlet_from_lowest(int W, const DemDesc *dd, int s0, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    int s = s0 + j;
    DemDesc D = dd[s];
    if (D.w == 0) return;
    double lowest = ord_dec(ds[s].lowest_enc);
    ds[s].outlet = (lowest < 1000.0) ? ds[s].argidx : D.y0 * W + D.x0;
}

int lem_slot_lowest(lemhip_ctx *c, const double *z, int s0, int ns)
{
    k_slots_lowest_init<<<(ns + 255) / 256, 256, 0, c->s>>>(s0, ns, c->dsc);
    dim3 g(c->slot_kbmax, ns);
    k_slot_lowest<<<g, 256, 0, c->s>>>(z, c->mask, c->W, c->ddesc, s0, c->dsc);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

int lem_slot_first_lowest(lemhip_ctx *c, const double *z, int s0, int ns)
{
    dim3 g(c->slot_kbmax, ns);
    k_slot_first_lowest<<<g, 256, 0, c->s>>>(z, c->mask, c->W, c->ddesc, s0, c->dsc);
    k_slots_outlet_from_lowest<<<(ns + 255) / 256, 256, 0, c->s>>>(c->W, c->ddesc, s0, ns, c->dsc);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

// max of fa with the reference's start value 0 (NaN never wins)
template <typename FS>
__global__ void __launch_bounds__(256) k_slot_famax(const FS *__restrict__ fa, int W, const DemDesc *dd, DemScal *ds)
{
    __shared__ unsigned long long sh[256];
    int s = blockIdx.y;
    DemDesc D = dd[s];
    if (slot_block_idle(D)) return;
    unsigned long long best = ord_enc(0.0);
    if (D.w > 0)
        slot_rows_u<4>(
            D, W, [&](long i) { return fa[i]; },
            [&](const FS &q) {
                double v = fa_full(q);
                if (v > ord_dec(best)) best = ord_enc(v);
            });
    best = block_max_u64(best, sh);
    if (threadIdx.x == 0 && D.w > 0) atomicMax(&ds[s].famax_enc, best);
}

// first cell holding the maximum (mfd_simple.cu:339-349: strict '>' scan from 0)
template <typename FS>
__global__ void __launch_bounds__(256) k_slot_fa_first(const FS *__restrict__ fa, int W, const DemDesc *dd, DemScal *ds)
{
    int s = blockIdx.y;
    DemDesc D = dd[s];
    if (slot_block_idle(D) || !ds[s].first_iter) return;
    double m = ord_dec(ds[s].famax_enc);
    if (!(m > 0.0)) return;
    slot_rows(D, W, [&](long i) {
        if (fa_full(fa[i]) == m) atomicMin(&ds[s].argidx, (int)i);
    });
}

__global__ void k_slots_famax_init(int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    ds[j].famax_enc = ord_enc(0.0);
    ds[j].argidx = 0x7FFFFFFF;
}

__global__ void k_slots_famax_set(int W, const DemDesc *dd, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    double m = ord_dec(ds[j].famax_enc);
    ds[j].FA_max = m;
    if (ds[j].first_iter) ds[j].outlet = (m > 0.0) ? ds[j].argidx : dd[j].y0 * W + dd[j].x0;
}

// updates.cu:151 and :165-174: base level from the outlet before the update,
// then the outlet is pinned 0.00024 below its pre-update value
__global__ void k_slots_base(const double *dem, const DemDesc *dd, int ns, DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    ds[j].base_level = dem[ds[j].outlet] - 0.00012;
}

__global__ void k_slots_post(double *dem, const DemDesc *dd, int ns, const DemScal *ds)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns || dd[j].w == 0) return;
    int o = ds[j].outlet;
    dem[o] = ds[j].base_level;
    dem[o] = dem[o] - 0.00012;
}

int lem_read_slots(lemhip_ctx *c)
{
    LEM_CHECK(c, hipMemcpyAsync(c->hdsc, c->dsc, sizeof(DemScal) * c->nslot, hipMemcpyDeviceToHost, c->s));
    LEM_CHECK(c, hipStreamSynchronize(c->s));
    return LEMHIP_OK;
}

// ---------------------------------------------------------------- phases

template <typename T>
static int runoff_t(lemhip_ctx *c)
{
    using FS = typename FaS<T>::type;
    prof_begin(c, PROF_RUNOFF);
    LEM_CHECK(c, hipMemsetAsync(&c->scal->w_bad, 0, sizeof(int), c->s));
    k_runoff<T><<<grid1d(c->N), 256, 0, c->s>>>(c->N, c->mask, (const T *)c->rain, (const T *)c->stone,
                                                 (const T *)c->TotB, (T *)c->soilM, (T *)c->w, (FS *)c->fa,
                                                 &c->scal->w_bad);
    prof_end(c, PROF_RUNOFF);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

#define DS_OFF(f) ((int)offsetof(DemScal, f))

// calcwater (runoffweight.cu:356-377): total runoff weight over mask cells
template <typename T>
static int calcwater_t(lemhip_ctx *c)
{
    return slot_sum<T>(c, c->w, DS_OFF(totwater), nullptr, 0, nullptr, 0, c->mask);
}

int lem_calcwater(lemhip_ctx *c)
{
    return c->prec == LEMHIP_FP64 ? calcwater_t<double>(c) : calcwater_t<float>(c);
}

// computeRunOffWeights then calcwater (lem.cpp:88-89)
int lem_runoff_phase(lemhip_ctx *c)
{
    LEM_TRY(c->prec == LEMHIP_FP64 ? runoff_t<double>(c) : runoff_t<float>(c));
    return lem_calcwater(c);
}


template <typename T>
static int fa_t(lemhip_ctx *c)
{
    using FS = typename FaS<T>::type;
    // the engine's launch choices are timed on the pass that uses it: here
    // unless flow accumulation runs in integers (then on sediment routing)
    const bool integer_fa = sizeof(T) == 4 && c->p.flowdirectiontype == 0 && c->fx_enabled;
    if (c->tune_dag_pending && !integer_fa) LEM_TRY(lem_tune_dag(c));
    prof_begin(c, PROF_FA);
    // FP32 form, single flow directions: exact integer subtree sums
    // (lem_fa32.hip); otherwise, and for inputs those cannot represent, the
    // dependency engine
    int fallback = 1;
    if (integer_fa) LEM_TRY(lem_fa32_sfd(c, &fallback));
    if (fallback) LEM_TRY(lem_fa_dag<T>(c));
    prof_end(c, PROF_FA, c->dag_rounds_last);
    // FA_max and the first-iteration outlet rule (mfd_simple.cu:332-357)
    int ns = c->nslot;
    dim3 g(c->slot_kbmax, ns);
    k_slots_famax_init<<<(ns + 255) / 256, 256, 0, c->s>>>(ns, c->dsc);
    k_slot_famax<FS><<<g, 256, 0, c->s>>>((const FS *)c->fa, c->W, c->ddesc, c->dsc);
    k_slot_fa_first<FS><<<g, 256, 0, c->s>>>((const FS *)c->fa, c->W, c->ddesc, c->dsc);
    k_slots_famax_set<<<(ns + 255) / 256, 256, 0, c->s>>>(c->W, c->ddesc, ns, c->dsc);
    LEM_LAUNCHED(c);
    return LEMHIP_OK;
}

int lem_fa_phase(lemhip_ctx *c, int iter)
{
    (void)iter;
    return c->prec == LEMHIP_FP64 ? fa_t<double>(c) : fa_t<float>(c);
}

// Runs `launches` (kernel launches on the context stream only: no copies,
// fills or host waits) directly, or, with graphs on, replays them from a
// graph captured the first time and again whenever the key (the launch
// shapes: band of rows, slot count, reduction split) changes
template <class F>
static int run_segment(lemhip_ctx *c, GraphSeg &g, long key, F launches)
{
    if (!c->use_graphs) return launches();
    if (g.key != key || !g.exec) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
        g.key = -1;
        LEM_CHECK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        const int rc = launches();
        c->capturing = false;
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->s, &gr);
        if (rc != LEMHIP_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess || !gr) {
            c->err = std::string("graph capture: ") + hipGetErrorString(e);
            return LEMHIP_E_HIP;
        }
        g.graph = gr;
        LEM_CHECK(c, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        g.key = key;
    }
    LEM_CHECK(c, hipGraphLaunch(g.exec, c->s));
    return LEMHIP_OK;
}

template <typename T>
static int mod_t(lemhip_ctx *c)
{
    using FS = typename FaS<T>::type;
    long N = c->N;
    int ns = c->nslot;
    int g = grid1d(N, 256, 16384);
    const long key = ((N * 256 + ns) * 2048) + c->slot_kbmax;
    // live timing: the kernels alone when launched directly; the whole
    // segment (kernel and totals) when replayed from a graph
    if (c->use_graphs) prof_begin(c, PROF_MOD1);
    LEM_TRY(run_segment(c, c->gseg[0], key, [&]() -> int {
        if (!c->use_graphs) prof_begin(c, PROF_MOD1);
        k_mod1<T><<<g, 256, 0, c->s>>>(N, c->NP, c->p, c->mask, (const T *)c->slopes, (const FS *)c->fa,
                                       (const T *)c->fines, (const T *)c->nut, (const T *)c->stone,
                                       (const T *)c->TotB, c->soilT, (const T *)c->temp, (T *)c->slopeptr,
                                       (T *)c->ero, (T *)c->inci, (T *)c->geli, (T *)c->depo);
        if (!c->use_graphs) prof_end(c, PROF_MOD1);
        LEM_LAUNCHED(c);
        return slot_sum<T>(c, c->ero, DS_OFF(totE), c->inci, DS_OFF(totI), c->geli, DS_OFF(totG));
    }));
    if (c->use_graphs) prof_end(c, PROF_MOD1);
    if (c->tune_dag_pending) {
        // still pending only in the FP32 form with integer flow accumulation
        if (sizeof(T) == 4 && !(c->p.flowdirectiontype == 0 && c->cs_rounds > 0)) LEM_TRY(lem_tune_sed(c));
        c->tune_dag_pending = false;
    }
    prof_begin(c, PROF_SED);
    {
        int fallback = 1;
        if (sizeof(T) == 4 && c->p.flowdirectiontype == 0 && c->cs_rounds > 0) LEM_TRY(lem_sed32(c, &fallback));
        if (fallback) LEM_TRY(lem_sed_dag<T>(c));
    }
    prof_end(c, PROF_SED, c->dag_rounds_last);
    // the low-soil variant sits one allocated grid after the other, whatever
    // band of rows this step covers (an atlas steps only its used rows, and
    // that band grows and shrinks as DEMs come and go)
    T *ca_hi = reinterpret_cast<T *>(c->tcache), *ca_lo = ca_hi + c->NP;
    if (!c->tcache_ok) {
        k_calcite_terms<T><<<g, 256, 0, c->s>>>(N, c->mask, (const T *)c->temp, ca_hi, ca_lo);
        LEM_LAUNCHED(c);
        c->tcache_ok = true;
    }
    if (c->use_graphs) prof_begin(c, PROF_MOD2);
    LEM_TRY(run_segment(c, c->gseg[1], key, [&]() -> int {
        LEM_TRY(slot_sum<T>(c, c->depo, DS_OFF(totD)));
        k_slots_base<<<(ns + 255) / 256, 256, 0, c->s>>>(c->dem, c->ddesc, ns, c->dsc);
        if (!c->use_graphs) prof_begin(c, PROF_MOD2);
        k_mod2<T><<<g, 256, 0, c->s>>>(N, c->p, c->mask, (const FS *)c->fa, (const T *)c->ero, (const T *)c->inci,
                                       (const T *)c->geli, (const T *)c->depo, (const T *)c->temp,
                                       (const T *)c->rain, (const T *)c->last_temp, c->soilT, (T *)c->stone,
                                       (T *)c->fines, (T *)c->soilM, (T *)c->soilB, (T *)c->nut, (T *)c->TotB,
                                       c->dem, (T *)c->dz, (T *)c->wC, (T *)c->wP, c->demid, c->dsc, ca_hi, ca_lo);
        if (!c->use_graphs) prof_end(c, PROF_MOD2);
        k_slots_post<<<(ns + 255) / 256, 256, 0, c->s>>>(c->dem, c->ddesc, ns, c->dsc);
        LEM_LAUNCHED(c);
        return slot_sum<T>(c, c->wC, DS_OFF(totwC), c->wP, DS_OFF(totwP), c->TotB, DS_OFF(totbio));
    }));
    if (c->use_graphs) prof_end(c, PROF_MOD2);
    return LEMHIP_OK;
}

int lem_mod_phase(lemhip_ctx *c, int iter)
{
    (void)iter;
    return c->prec == LEMHIP_FP64 ? mod_t<double>(c) : mod_t<float>(c);
}
