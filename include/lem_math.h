/*
 * lem_math.h — portable FP64 elementary functions for the landscape-evolution
 * timestep, compiled identically for the gfx950 kernels (hipcc) and for host
 * code (gcc/g++).
 *
 * Why this exists: the reference evaluates exp/log/log10/pow/tanh with the
 * CUDA device math library (call sites: MOD_kernels/runoffweight.cu:43,
 * MOD_kernels/eroincidep.cu:64,79-81,137-148,241-313, MOD_kernels/updates.cu:63-64,
 * 112,119).  That library is not available here, so every FP64 transcendental
 * on the path goes through the routines below.  They use only IEEE-754
 * binary64 +, -, *, / (correctly rounded on both x86-64 SSE2 and AMDGPU) and
 * exact bit manipulation, and every translation unit that includes this file
 * must be compiled with -ffp-contract=off, so a host build and a device build
 * return bit-identical results for every input.
 *
 * Methods (all standard, written from the mathematics):
 *  - exp:   Cody-Waite reduction x = k*ln2 + r with a two-part ln2
 *           (|r| <= ln2/2), Taylor polynomial of degree 13 for e^r - 1
 *           (truncation < 5e-18 relative), scaling by 2^k through the exponent
 *           field.
 *  - log:   x = m*2^e with m in [sqrt(1/2), sqrt(2)); f = m-1 (exact),
 *           s = f/(2+f), log(1+f) = f - (f^2/2 - s*(f^2/2 + R(s^2))) where
 *           R(z) = sum_{k>=1} 2 z^k/(2k+1) (the atanh series), 11 terms.
 *  - log10: log(x) * (1/ln 10).
 *  - pow:   special cases, y == 2 -> x*x, base 10 -> exp of 10's exponent
 *           with a two-product split of z*ln10; otherwise l = log(x) is
 *           refined by c = (x - e^l)/e^l (x - e^l is exact, the two being
 *           within a factor of two), y*l is carried as an exact two-product
 *           p + err, and the result is e^p * (1 + err + y*c).
 *  - expm1: Taylor for |x| <= ln2/2, else 2^k(1+em1(r)) - 1.
 *  - tanh:  t = expm1(2|x|)/(expm1(2|x|)+2), sign restored; 1 for |x| > 22.
 * Accuracy against a correctly rounded libm is a few ulp; the unit tests
 * state the bound they check.
 */
#ifndef LEM_MATH_H
#define LEM_MATH_H

#ifdef __HIPCC__
#define LM_FN static __host__ __device__ inline
#else
#define LM_FN static inline
#endif

LM_FN double lm_from_bits(unsigned long long u) { double d; __builtin_memcpy(&d, &u, 8); return d; }
LM_FN unsigned long long lm_to_bits(double d) { unsigned long long u; __builtin_memcpy(&u, &d, 8); return u; }

LM_FN double lm_inf(void) { return lm_from_bits(0x7FF0000000000000ULL); }
LM_FN double lm_nan(void) { return lm_from_bits(0x7FF8000000000000ULL); }
LM_FN double lm_fabs(double x) { return lm_from_bits(lm_to_bits(x) & 0x7FFFFFFFFFFFFFFFULL); }

/* 2^k for k in [-1022, 1023] */
LM_FN double lm_pow2i(int k) { return lm_from_bits((unsigned long long)(k + 1023) << 52); }

/* y * 2^k for k in [-2100, 2100], in at most three exact-or-final multiplications */
LM_FN double lm_ldexp(double y, int k)
{
    if (k > 1023) { y = y * lm_pow2i(1023); k -= 1023; if (k > 1023) k = 1023; }
    else if (k < -1022) {
        y = y * lm_pow2i(-1022); k += 1022;
        if (k < -1022) { y = y * lm_pow2i(-1022); k += 1022; if (k < -1022) k = -1022; }
    }
    return y * lm_pow2i(k);
}

#define LM_LN2_HI   6.93147180369123816490e-01  /* ln2 with the low 32 mantissa bits zero */
#define LM_LN2_LO   1.90821492927058770002e-10  /* ln2 - LM_LN2_HI */
#define LM_INV_LN2  1.44269504088896338700e+00
#define LM_INV_LN10 0.43429448190325176

/* e^r - 1 for |r| <= ~0.35: r + r^2 * (1/2! + r/3! + ... + r^11/13!) */
LM_FN double lm_em1_core(double r)
{
    double q = 1.6059043836821613e-10;
    q = q * r + 2.08767569878681e-09;
    q = q * r + 2.505210838544172e-08;
    q = q * r + 2.755731922398589e-07;
    q = q * r + 2.7557319223985893e-06;
    q = q * r + 2.48015873015873e-05;
    q = q * r + 0.0001984126984126984;
    q = q * r + 0.001388888888888889;
    q = q * r + 0.008333333333333333;
    q = q * r + 0.041666666666666664;
    q = q * r + 0.16666666666666666;
    q = q * r + 0.5;
    return r + (r * r) * q;
}

/* round-half-away-from-zero of x / ln2, as an int; callers bound |x| */
LM_FN int lm_nearest_k(double x)
{
    double t = x * LM_INV_LN2;
    return (int)(t < 0.0 ? t - 0.5 : t + 0.5);
}

LM_FN double lm_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return lm_inf();
    if (x < -745.1332191019412) return 0.0;
    int k = lm_nearest_k(x);
    double hi = x - (double)k * LM_LN2_HI;
    double lo = (double)k * LM_LN2_LO;
    double r = hi - lo;
    double y = 1.0 + lm_em1_core(r);
    return lm_ldexp(y, k);
}

LM_FN double lm_expm1(double x)
{
    if (x != x) return x;
    if (x > 709.0) return lm_exp(x);
    if (x < -40.0) return -1.0;
    if (x <= 0.34657359027997264 && x >= -0.34657359027997264) return lm_em1_core(x);
    int k = lm_nearest_k(x);
    double hi = x - (double)k * LM_LN2_HI;
    double lo = (double)k * LM_LN2_LO;
    double r = hi - lo;
    double e = lm_em1_core(r);
    if (k > 0 && k <= 53) return lm_ldexp(e, k) + (lm_pow2i(k) - 1.0);
    return lm_ldexp(1.0 + e, k) - 1.0;
}

LM_FN double lm_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return lm_nan();
    if (x == 0.0) return -lm_inf();
    if (x == lm_inf()) return x;
    int e = 0;
    unsigned long long u = lm_to_bits(x);
    if (u < (1ULL << 52)) { x = x * 18014398509481984.0; u = lm_to_bits(x); e = -54; } /* subnormal: * 2^54 */
    e += (int)(u >> 52) - 1023;
    double m = lm_from_bits((u & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0.08695652173913043;
    R = R * z + 0.09523809523809523;
    R = R * z + 0.10526315789473684;
    R = R * z + 0.11764705882352941;
    R = R * z + 0.13333333333333333;
    R = R * z + 0.15384615384615385;
    R = R * z + 0.18181818181818182;
    R = R * z + 0.2222222222222222;
    R = R * z + 0.2857142857142857;
    R = R * z + 0.4;
    R = R * z + 0.6666666666666666;
    R = R * z;
    double hfsq = 0.5 * f * f;
    double dk = (double)e;
    return dk * LM_LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LM_LN2_LO)) - f);
}

LM_FN double lm_log10(double x) { return lm_log(x) * LM_INV_LN10; }

/* exact split of a double into 26-bit halves (Dekker), no fused operations */
LM_FN void lm_split(double a, double *hi, double *lo)
{
    double c = 134217729.0 * a; /* 2^27 + 1 */
    double h = c - (c - a);
    *hi = h;
    *lo = a - h;
}

/* 10^z with z*ln10 carried as an unevaluated sum p + q */
LM_FN double lm_exp10(double z)
{
    const double LN10 = 2.302585092994046;
    const double LN10_LO = -2.1707562233822494e-16; /* ln10 - LN10 */
    if (z != z) return z;
    double p = z * LN10;
    double zh, zl, lh, ll;
    lm_split(z, &zh, &zl);
    lm_split(LN10, &lh, &ll);
    double err = ((zh * lh - p) + zh * ll + zl * lh) + zl * ll; /* z*LN10 - p, exactly */
    double q = err + z * LN10_LO;
    double r = lm_exp(p);
    return r + r * q;
}

LM_FN double lm_pow(double x, double y)
{
    if (y == 2.0) return x * x;
    if (x != x || y != y) return lm_nan();
    if (y == 0.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : lm_inf();
    if (x < 0.0) return lm_nan();
    if (x == lm_inf()) return y > 0.0 ? lm_inf() : 0.0;
    if (x == 10.0) return lm_exp10(y);
    double l = lm_log(x);
    double e = lm_exp(l);
    double c = (e > 0.0 && e < lm_inf()) ? (x - e) / e : 0.0;
    double p = y * l;
    double yh, yl, lh, ll;
    lm_split(y, &yh, &yl);
    lm_split(l, &lh, &ll);
    double err = ((yh * lh - p) + yh * ll + yl * lh) + yl * ll; /* y*l - p, exactly */
    double q = err + y * c;
    double r = lm_exp(p);
    return r + r * q;
}

LM_FN double lm_tanh(double x)
{
    if (x != x) return x;
    double ax = lm_fabs(x);
    double t;
    if (ax > 22.0) t = 1.0;
    else if (ax < 3.725290298461914e-09) t = ax; /* 2^-28 */
    else { double e = lm_expm1(2.0 * ax); t = e / (e + 2.0); }
    return x < 0.0 ? -t : t;
}

#endif /* LEM_MATH_H */
