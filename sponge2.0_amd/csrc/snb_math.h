// snb_math.h -- the float arithmetic of the nonbonded path, host+device.
//
// Every function here uses only IEEE-defined operations (add, sub, mul,
// correctly rounded div, fma, floor, rint, integer bit manipulation), so the
// same source evaluated on gfx950 and on a CPU gives the same bits.  The
// library is built with -ffp-contract=off; an fma appears only where it is
// written.  Per-pair / per-point terms are rounded to int64 fixed point
// (scale 2^32) before they are summed, which makes every accumulator
// independent of summation order.
#pragma once
#include <stdint.h>

#include "snb_erfcx_table.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SNB_HD __host__ __device__ __forceinline__
#else
#define SNB_HD static inline
#endif

#define SNB_FIXED_SCALE 4294967296.0          // 2^32
#define SNB_FIXED_INV 2.3283064365386963e-10  // 2^-32

SNB_HD float snb_f_from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }
SNB_HD uint32_t snb_bits_from_f(float f) { return __builtin_bit_cast(uint32_t, f); }

// round(f * 2^32), ties to even, exact for |f| < 2^19: adding 1.5*2^52 puts
// the integer in the low mantissa bits of a double in [2^52, 2^53).
SNB_HD int64_t snb_to_fixed(float f)
{
    double d = __builtin_fma((double)f, SNB_FIXED_SCALE, 6755399441055744.0);
    return (int64_t)__builtin_bit_cast(uint64_t, d) - (int64_t)0x4338000000000000LL;
}

SNB_HD float snb_fixed_to_float(int64_t v) { return (float)((double)v * SNB_FIXED_INV); }
SNB_HD double snb_fixed_to_double(int64_t v) { return (double)v * SNB_FIXED_INV; }

// 1/sqrt(x), x > 0 normal: seed 0x5f375a86 - (bits >> 1) (Lomont, "Fast
// inverse square root", 2003; relative error below 3.5e-2), three Newton
// steps y <- y (1.5 - (x/2) y^2): the error squares each step (x 1.5),
// 1.8e-3, 5e-6, then 4e-11 before the float rounding of the steps.
SNB_HD float snb_rsqrtf(float x)
{
    float y = snb_f_from_bits(0x5f375a86u - (snb_bits_from_f(x) >> 1));
    float h = 0.5f * x;
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    return y;
}

// round(v) to int64, ties to even, for a float v that already carries the
// 2^32 fixed-point scale (|v| < 2^51): v + 1.5*2^52 in double rounds to an
// integer, which is the same as snb_to_fixed(v / 2^32).
SNB_HD int64_t snb_fixed_from_scaled(float v)
{
    double d = (double)v + 6755399441055744.0;
    return (int64_t)__builtin_bit_cast(uint64_t, d) - (int64_t)0x4338000000000000LL;
}

// ---------------------------------------------------------------------------
// Screened-Coulomb force shape h(x) = erfc(x) + (2/sqrt(pi)) x e^{-x^2}, the
// factor of q_i q_j / r^3 in the direct-space force (Lennard_Jones_force.cpp
// :87-95 with x = beta r).  Tabulated as SNB_COUL_N cubic Hermite pieces on
// x in [0, X], X = beta * cutoff * (1 + 2^-7): node k at x_k = k / scale,
// scale = (float)(SNB_COUL_N / X); piece k holds {c0, c1, c2, c3} of
// h(x_k + t / scale) = c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1], from h and
// h'(x) = -(4/sqrt(pi)) x^2 e^{-x^2} at both ends, evaluated in double by the
// functions below and rounded once to float.  Interpolation error
// max|h''''| dx^4 / 384 < 1e-9 at 256 pieces over x <= 4.
// ---------------------------------------------------------------------------
#define SNB_COUL_N 256

// Piece index and local coordinate of x >= 0; x a little past X (the rounding
// of r = r2 * rsqrt(r2) at the cutoff) stays on the last piece.
SNB_HD int snb_coul_piece(float x, float scale, float* t)
{
    float u = x * scale;
    int k = (int)u;
    k = k < SNB_COUL_N - 1 ? k : SNB_COUL_N - 1;
    *t = u - (float)k;
    return k;
}

SNB_HD float snb_coul_eval(float t, float c0, float c1, float c2, float c3)
{
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c3, t, c2), t, c1), t, c0);
}

// ---------------------------------------------------------------------------
// FP64 pieces (energy steps and the exclusion correction)
// ---------------------------------------------------------------------------
SNB_HD double snb_f64_from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }

// exp(y), y in [-700, 0]: Cody-Waite y = n ln2 + r with ln2 split into a
// 33-bit head and a tail, Taylor polynomial of degree 13 for e^r on
// |r| <= ln2/2 (truncation error < 5e-18), 2^n in the exponent field.
SNB_HD double snb_exp_d(double y)
{
    y = __builtin_fmax(y, -700.0);
    double n = __builtin_rint(y * 1.4426950408889634);
    double r = __builtin_fma(-n, 0.6931471806019545, y);
    r = __builtin_fma(-n, -4.2009150726810846e-11, r);
    double p = 1.0 / 6227020800.0;               // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);  // 1/12!
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    int64_t ni = (int64_t)n;
    return p * snb_f64_from_bits((uint64_t)(ni + 1023) << 52);
}

// erfc(x) and g(x) = (2/sqrt(pi)) x e^{-x^2} for x >= 0, at a fixed cost:
// erfc(x) = e^{-x^2} erfcx(x), erfcx (the scaled complementary error
// function) from piecewise Chebyshev-interpolation polynomials of degree 12
// on half-unit intervals of [0, 6) (snb_erfcx_table.h, relative error below
// 1e-15), evaluated by Horner's rule in t = (x - centre) / 0.25; erfc(x) = 0
// from x = 6 on (erfc(6) = 2e-17).  erf = 1 - erfc and the exclusion
// correction's erf(x) - g(x) follow from these two without a series.
SNB_HD double snb_erfc_parts_d(double x, double* g_out)
{
    const double e = snb_exp_d(-(x * x));
    *g_out = 1.1283791670955126 * x * e;
    if (!(x < 6.0)) return 0.0;
    int k = (int)(x * 2.0);
    k = k < 0 ? 0 : k;
    const double t = (x - ((double)k + 0.5) * 0.5) * 4.0;
    const double* c = snb_erfcx_coef[k];
    double p = c[SNB_ERFCX_DEG];
    for (int j = SNB_ERFCX_DEG - 1; j >= 0; j--) p = __builtin_fma(p, t, c[j]);
    return e * p;
}

// The SNB_COUL_N pieces of h (4 floats each, layout above) for beta and
// cutoff; returns scale.  h(x) = 1 - (erf(x) - (2/sqrt(pi)) x e^{-x^2}), the
// second term straight from the series, so there is no cancellation.
SNB_HD float snb_coul_table(float beta, float cutoff, float* tab)
{
    const double X = (double)beta * (double)cutoff * (1.0 + 1.0 / 128.0);
    if (!(X > 0.0))
    {
        // beta = 0: h = 1 everywhere, every x lands on piece 0 at t = 0
        for (int k = 0; k < SNB_COUL_N; k++)
        {
            tab[4 * k] = 1.0f;
            tab[4 * k + 1] = tab[4 * k + 2] = tab[4 * k + 3] = 0.0f;
        }
        return 0.0f;
    }
    const float scale = (float)((double)SNB_COUL_N / X);
    const double dx = 1.0 / (double)scale;
    const double four_over_sqrt_pi = 2.2567583341910252;
    double h0 = 0.0, d0 = 0.0;
    for (int k = 0; k <= SNB_COUL_N; k++)
    {
        const double x = (double)k * dx;
        double g;
        const double h1 = snb_erfc_parts_d(x, &g) + g;
        const double d1 = ((-four_over_sqrt_pi * x) * x) * snb_exp_d(-(x * x)) * dx;
        if (k > 0)
        {
            float* c = tab + 4 * (k - 1);
            c[0] = (float)h0;
            c[1] = (float)d0;
            c[2] = (float)((3.0 * (h1 - h0) - 2.0 * d0) - d1);
            c[3] = (float)((2.0 * (h0 - h1) + d0) + d1);
        }
        h0 = h1;
        d0 = d1;
    }
    return scale;
}

// round(d * 2^32) to nearest-even for |d| < 2^19.
SNB_HD int64_t snb_to_fixed_d(double d)
{
    double v = __builtin_fma(d, SNB_FIXED_SCALE, 6755399441055744.0);
    return (int64_t)__builtin_bit_cast(uint64_t, v) - (int64_t)0x4338000000000000LL;
}

// ---------------------------------------------------------------------------
// Random numbers for the Langevin thermostat
// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplications with the
// published multipliers and Weyl key increments.
SNB_HD void snb_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++)
    {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// uniform in (0, 1): the top 24 bits plus half a step, exact in float
SNB_HD float snb_uniform01(uint32_t x)
{
    return (float)(x >> 8) * 5.9604644775390625e-08f + 2.98023223876953125e-08f;
}

// ln(u) for u in (0, 1]: u = m 2^e with m in [sqrt(1/2), sqrt(2)), then
// ln m = 2 atanh(s) with s = (m - 1)/(m + 1), |s| < 0.1716, as the odd series
// 2 s (1 + s^2/3 + s^4/5 + s^6/7 + s^8/9) (truncation below 1e-9).
SNB_HD float snb_log_unit(float u)
{
    uint32_t b = snb_bits_from_f(u);
    int e = (int)(b >> 23) - 127;
    float m = snb_f_from_bits((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f)
    {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 1.0f / 9.0f;
    p = __builtin_fmaf(p, s2, 1.0f / 7.0f);
    p = __builtin_fmaf(p, s2, 0.2f);
    p = __builtin_fmaf(p, s2, 1.0f / 3.0f);
    p = __builtin_fmaf(p, s2, 1.0f);
    return __builtin_fmaf((float)e, 0.693147182f, (2.0f * s) * p);
}

// sin and cos of 2 pi u for u in (0, 1): x = u - rint(u) in [-1/2, 1/2], a
// quarter-turn count j = rint(4x) and a remainder angle a = 2 pi (x - j/4)
// in [-pi/4, pi/4], Taylor polynomials to a^9 / a^8 (error below 3e-9), then
// the quarter turns.
SNB_HD void snb_sincos_2pi(float u, float* s_out, float* c_out)
{
    const float x = u - __builtin_rintf(u);
    const float jf = __builtin_rintf(4.0f * x);
    const float a = (x - 0.25f * jf) * 6.28318548f;
    const float a2 = a * a;
    float sp = 1.0f / 362880.0f;
    sp = __builtin_fmaf(sp, a2, -1.0f / 5040.0f);
    sp = __builtin_fmaf(sp, a2, 1.0f / 120.0f);
    sp = __builtin_fmaf(sp, a2, -1.0f / 6.0f);
    const float sa = __builtin_fmaf(a * a2, sp, a);
    float cp = 1.0f / 40320.0f;
    cp = __builtin_fmaf(cp, a2, -1.0f / 720.0f);
    cp = __builtin_fmaf(cp, a2, 1.0f / 24.0f);
    cp = __builtin_fmaf(cp, a2, -0.5f);
    const float ca = __builtin_fmaf(a2, cp, 1.0f);
    const int j = (int)jf & 3;  // rotations by j quarter turns (j in -2..2)
    float s = sa, c = ca;
    if (j == 1)
    {
        s = ca;
        c = -sa;
    }
    else if (j == 2)
    {
        s = -sa;
        c = -ca;
    }
    else if (j == 3)
    {
        s = -ca;
        c = sa;
    }
    *s_out = s;
    *c_out = c;
}

// Two standard normal numbers from two 32-bit words (Box-Muller).
SNB_HD void snb_box_muller(uint32_t x0, uint32_t x1, float* n0, float* n1)
{
    const float r = __builtin_sqrtf(-2.0f * snb_log_unit(snb_uniform01(x0)));
    float s, c;
    snb_sincos_2pi(snb_uniform01(x1), &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

// Triangular LJ type-pair index (Lennard_Jones_force.h:18-27).
SNB_HD int snb_lj_pair_index(int a, int b)
{
    int hi = a > b ? a : b, lo = a > b ? b : a;
    return ((hi * (hi + 1)) >> 1) + lo;
}

// Cubic B-spline weight w_k(t) and derivative for stencil offset k, with the
// reference's coefficient tables and evaluation order (PM_force.cpp:18-24,
// :776-805, :859-880).
SNB_HD float snb_bspline_w(int k, float t)
{
    const float ma = k == 0 ? 1.0f / 6.0f : (k == 1 ? -0.5f : (k == 2 ? 0.5f : -1.0f / 6.0f));
    const float mb = k == 0 ? 0.0f : (k == 1 ? 0.5f : (k == 2 ? -1.0f : 0.5f));
    const float mc = k == 0 ? 0.0f : (k == 1 ? 0.5f : (k == 2 ? 0.0f : -0.5f));
    const float md = k == 0 ? 0.0f : (k == 1 ? 1.0f / 6.0f : (k == 2 ? 4.0f / 6.0f : 1.0f / 6.0f));
    float t2 = t * t;
    return ((ma * t) * t2 + mb * t2) + mc * t + md;
}

SNB_HD float snb_bspline_dw(int k, float t)
{
    const float da = k == 0 ? 0.5f : (k == 1 ? -1.5f : (k == 2 ? 1.5f : -0.5f));
    const float db = k == 0 ? 0.0f : (k == 1 ? 1.0f : (k == 2 ? -2.0f : 1.0f));
    const float dc = k == 0 ? 0.0f : (k == 1 ? 0.5f : (k == 2 ? 0.0f : -0.5f));
    float t2 = t * t;
    return (da * t2 + db * t) + dc;
}
