// Relaxations of sin, cos, odd integer powers and non-integer powers for the
// expression-program bound, with the tangent tables they read.
//
// The tables depend only on the function, so a plan builds them once on the
// host (the same IEEE operations as on the device: gearhip_math.h, no
// contraction) and keeps them in device memory beside the tape; the bound
// kernel reads one entry per endpoint.
//
// Reference (auto_LiRPA/operators):
//   sin  trigonometric.py BoundSin: tables _ensure_precomputed (:49-69) from
//        BoundSShaped.precompute_relaxation (s_shaped.py:125-202, x_limit
//        pi/2) and the bisection tangents get_*_bound (trigonometric.py:
//        235-321); bound_relax (:71-85) = s_shaped.py:557-574 with
//        branch_input_domain (trigonometric.py:110-153), bound_relax_impl
//        (s_shaped.py:325-549), bound_relax_branch + get_bound_tb
//        (trigonometric.py:323-506), roundoff margin 8 eps (1 + max |x|);
//   cos  BoundCos (:509-527): sin on the interval shifted by pi/2;
//   x^p  odd p: BoundPow (s_shaped.py:652-770, 913-916) on the interval
//        with its upper end raised to lower + 1e-8 (:795-797); non-integer
//        p > 0: bound_relax_noninteger (:974-1037) on the interval cut to
//        x >= 0 (domain_aware_ops).
#pragma once

#include <stdint.h>

#include "../../include/gearhip_math.h"

#if defined(__HIPCC__)
#define GH_SS __host__ __device__ inline
#else
#define GH_SS inline
#endif

namespace gh {
namespace ss {

constexpr double PI = 3.141592653589793;
constexpr double EPS64 = 2.220446049250313080847e-16;
constexpr int SIN_TABLE = 1001;   // linspace(0, 2 pi, 1001)
constexpr int SIN_D = 162;        // int((pi/2) / 0.01) + 5
constexpr int POW_D = 50005;      // int(500 / 0.01) + 5

// sin table block (doubles): d_lower, d_upper (SIN_D each), then d and
// cycles of the four endpoint tables (SIN_TABLE each)
constexpr int SIN_DL = 0, SIN_DU = SIN_D, SIN_LL = 2 * SIN_D;
constexpr int SIN_BLOCK = 2 * SIN_D + 8 * SIN_TABLE;
enum { T_LL_D = 0, T_LL_C, T_LU_D, T_LU_C, T_UL_D, T_UL_C, T_UU_D, T_UU_C };
// odd power block: d_lower, d_upper (POW_D each)
constexpr int POW_BLOCK = 2 * POW_D;

// double -> int64 as the reference's tensors convert (out of range: INT64_MIN)
GH_SS int64_t to_long(double q)
{
    if (!(q > -9.2233720368547758e18 && q < 9.2233720368547758e18)) return INT64_MIN;
    return (int64_t)q;
}
GH_SS double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
GH_SS int64_t idx1(double q)  // max(0, long(q)) + 1
{
    int64_t i = to_long(q);
    return (i > 0 ? i : 0) + 1;
}

// ---------------------------------------------------------------- tables
// tangent line at t evaluated against sin at x (check_bound, :224-233)
GH_SS double sin_margin(double t, double x)
{
    double d = gh_cos(t);
    return gh_sin(x) - (d * (x - t) + gh_sin(t));
}

// endpoint tangents: kind 0 lower-left, 1 upper-left, 2 lower-right,
// 3 upper-right (get_lower_left_bound ... get_upper_right_bound)
GH_SS void sin_endpoint(int kind, double x, double *d, double *cyc)
{
    double cycles, t, c2;
    if (kind >= 2) {  // right ends reuse the left searches on a reflected point
        if (kind == 2) {
            cycles = floor((x + 0.5 * PI) / (2 * PI)) * (2 * PI);
            sin_endpoint(0, PI - (x - cycles), &t, &c2);
            *d = 3 * PI - t;
        } else {
            cycles = floor((x - 0.5 * PI) / (2 * PI)) * (2 * PI);
            sin_endpoint(1, 3 * PI - (x - cycles), &t, &c2);
            *d = 5 * PI - t;
        }
        *cyc = cycles - 2 * PI;
        return;
    }
    const bool lower = kind == 0;
    cycles = floor((x + (lower ? 0.5 * PI : -0.5 * PI)) / (2 * PI)) * (2 * PI);
    x = x - cycles;
    const bool use = lower ? x >= PI : x >= 2.0 * PI;
    double left = lower ? PI : 2.0 * PI, right = lower ? 1.5 * PI : 2.5 * PI, last = right;
    for (int s = 0; s < 20; ++s) {
        double mid = (left + right) / 2.;
        bool pos = sin_margin(mid, x) > 0;
        if (pos == lower) {
            right = mid;
            last = mid;
        } else {
            left = mid;
        }
    }
    *d = use ? x : last;
    *cyc = cycles;
}

inline void build_sin_tables(double *T)
{
    for (int i = 0; i < SIN_D; ++i) {
        double up = 0.01 * (double)i, l = -1.0, r = 0.0;
        auto ok_lower = [](double u, double d) { return gh_cos(d) * (u - d) + gh_sin(d) <= gh_sin(u); };
        auto ok_upper = [](double lo, double d) { return gh_cos(d) * (lo - d) + gh_sin(d) >= gh_sin(lo); };
        while (!ok_lower(up, l)) l = l * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (ok_lower(up, m)) l = m;
            else r = m;
        }
        T[SIN_DL + i] = l;
        double lo = -0.01 * (double)i;
        l = 0.0;
        r = 1.0;
        while (!ok_upper(lo, r)) r = r * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (ok_upper(lo, m)) r = m;
            else l = m;
        }
        T[SIN_DU + i] = r;
    }
    // linspace(0, 2 pi, 1001): the first half from the start, the second
    // from the end, one fma each
    const double step = (2 * PI - 0.0) / (SIN_TABLE - 1);
    for (int i = 0; i < SIN_TABLE; ++i) {
        double x = i < SIN_TABLE / 2 ? fma(step, (double)i, 0.0)
                                     : fma(-step, (double)(SIN_TABLE - 1 - i), 2 * PI);
        for (int k = 0; k < 4; ++k)
            sin_endpoint(k, x, &T[SIN_LL + (2 * k) * SIN_TABLE + i],
                         &T[SIN_LL + (2 * k + 1) * SIN_TABLE + i]);
    }
}

GH_SS double fpow(int p, double x) { return gh_powi(x, p); }
GH_SS double dfpow(int p, double x) { return (double)p * gh_powi(x, p - 1); }

inline void build_pow_tables(int p, double *T)
{
    for (int i = 0; i < POW_D; ++i) {
        // an upper end >= 0: the tangent point whose line stays above f there
        double up = 0.01 * (double)i, l = -1.0, r = 0.0;
        auto ok_upper = [p](double lo, double d) { return dfpow(p, d) * (lo - d) + fpow(p, d) >= fpow(p, lo); };
        auto ok_lower = [p](double u, double d) { return dfpow(p, d) * (u - d) + fpow(p, d) <= fpow(p, u); };
        while (!ok_upper(up, l)) l = l * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (ok_upper(up, m)) l = m;
            else r = m;
        }
        T[POW_D + i] = l;  // d_upper
        // a lower end <= 0: the tangent point whose line stays below f there
        double lo = -0.01 * (double)i;
        l = 0.0;
        r = 1.0;
        while (!ok_lower(lo, r)) r = r * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (ok_lower(lo, m)) r = m;
            else l = m;
        }
        T[i] = r;  // d_lower
    }
}

// ----------------------------------------------------------- relaxations
// planes R = (lw, lb, uw, ub); add_linear_relaxation: w = k, b = -x0 k + y0
GH_SS void put(double &w, double &b, double k, double x0, double y0)
{
    w = k;
    b = -x0 * k + y0;
}

GH_SS double n_crossing(double start, double end, double s)
{
    double cycles = floor((end - start) / (2 * PI));
    double dist = floor((s - start) / (2 * PI));
    double real_s = s - dist * 2 * PI;
    double real_end = end - cycles * 2 * PI;
    return (double)(real_s >= start) * (double)(real_s <= real_end) + cycles;
}

// one of the two endpoint tangents or the tangent at the midpoint
// (get_bound_tb, trigonometric.py:356-413)
GH_SS double tb_pick(double mid, double lower, double upper, double tl, double tr, double cl,
                     double cr, double lbc, double ubc, double end)
{
    double ul = tl + cl + lbc, le = end + cl + lbc;
    double ur = tr + cr + ubc, re = end + cr + ubc;
    bool lm = mid < ul && ul <= upper;
    bool lr = mid >= ul && mid < le;
    bool ib = mid >= le && mid < re;
    bool rm = mid >= ur && ur >= lower;
    bool rr = !lr && (mid >= re && mid < ur);
    double t = (double)lm * tl + (double)lr * (mid - cl - lbc) + (double)ib * end +
               (double)rm * tr + (double)rr * (mid - cr - ubc);
    return (lm || lr || ib || rm || rr) ? t : tl;
}

// the relaxation on an interval that is not one S-shaped piece:
// bound_relax_branch (trigonometric.py:469-506)
GH_SS void sin_branch(const double *T, double lb, double ub, double R[4])
{
    ub = fmax(ub, lb + 1e-8);
    double sub = gh_sin(ub), slb = gh_sin(lb);
    double gap = gh_sin((ub + lb) / 2) - (sub + slb) / 2.;
    double k1 = (sub - slb) / fmax(ub - lb, 1e-10);
    double b1 = slb - k1 * lb;
    // points where cos x equals the secant slope (get_intersection, theta = pi/2)
    double cr1 = gh_asin(k1) - 0.5 * PI;
    double cr2 = PI - cr1 - 2 * (0.5 * PI);
    bool use = n_crossing(lb, ub, cr1) + n_crossing(lb, ub, cr2) == 1;
    // get_bound_tb: the table tangents of the folded endpoints
    const double step = 2 * PI / (SIN_TABLE - 1);
    double lbc = floor(lb / (2 * PI)) * (2 * PI);
    double lf = clampd(lb - lbc, 0, 2 * PI);
    double ubc = floor(ub / (2 * PI)) * (2 * PI);
    double uf = clampd(ub - ubc, 0, 2 * PI);
    int64_t il = to_long(lf / step);
    int64_t iu = to_long(ceil(uf / step));
    if (iu > SIN_TABLE - 1) iu = SIN_TABLE - 1;
    const double *E = T + SIN_LL;
    auto tab = [&](int k, int64_t i) { return E[k * SIN_TABLE + i]; };
    double mid = (lb + ub) / 2;
    double tlo = tb_pick(mid, lb, ub, tab(T_LL_D, il), tab(T_UL_D, iu), tab(T_LL_C, il),
                         tab(T_UL_C, iu), lbc, ubc, 1.5 * PI);
    double tup = tb_pick(mid, lb, ub, tab(T_LU_D, il), tab(T_UU_D, iu), tab(T_LU_C, il),
                         tab(T_UU_C, iu), lbc, ubc, 2.5 * PI);
    double dl = gh_cos(tlo);
    double bl = gh_sin(tlo) -
                dl * (tlo + (tlo <= 1.5 * PI ? tab(T_LL_C, il) + lbc : tab(T_UL_C, iu) + ubc));
    double du = gh_cos(tup);
    double bu = gh_sin(tup) -
                du * (tup + (tup <= 2.5 * PI ? tab(T_LU_C, il) + lbc : tab(T_UU_C, iu) + ubc));
    bool lul = gap >= 0 && use, uul = gap < 0 && use;
    R[0] = lul ? k1 : dl;
    R[1] = lul ? b1 : bl;
    R[2] = uul ? k1 : du;
    R[3] = uul ? b1 : bu;
}

// sin over [l, u]; returns the undefined sides
GH_SS int relax_sin(const double *T, double l, double u, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    // branch_input_domain: which quarter period each end lies in
    double lf = l - floor(l / (2 * PI)) * (2 * PI);
    double uf = u - floor(u / (2 * PI)) * (2 * PI);
    auto quad = [](double x) {
        return (x >= 0 && x < PI / 2) ? 1 : (x >= PI / 2 && x < PI) ? 2
             : (x >= PI && x < 3 * PI / 2) ? 3 : (x >= 3 * PI / 2 && x < 2 * PI) ? 4 : 0;
    };
    int ql = quad(lf), qu = quad(uf);
    bool sl = (u - l <= PI) && (((ql == 2 || ql == 3) && (qu == 2 || qu == 3)) ||
                                ((ql == 1 || ql == 4) && (qu == 1 || qu == 4)));
    if (!sl) {
        sin_branch(T, l, u, R);
    } else {
        bool mneg = (ql == 3 || ql == 4) && (qu == 3 || qu == 4);
        bool mpos = (ql == 1 || ql == 2) && (qu == 1 || qu == 2);
        bool mboth = !(mneg || mpos);
        bool cc = -gh_sin(l) >= 0;
        double lw = 0, lb = 0, uw = 0, ub = 0;
        double yl = gh_sin(l), yu = gh_sin(u);
        double width = u - l;
        double kd = (yu - yl) / fmax(width, 1e-8);
        if (width < 1e-8) kd = gh_cos(l);
        if (mneg) put(uw, ub, kd, l, yl);
        if (mpos) put(lw, lb, kd, l, yl);
        // generate_d_lower_upper (trigonometric.py:155-205)
        double k = floor(u / (2 * PI));
        double uc = u - k * (2 * PI);
        double dlo = (T[SIN_DL + idx1(clampd(uc, 0, PI / 2) / 0.01)] + k * 2 * PI) *
                     (double)(uc >= 0 && uc <= PI / 2);
        double dup = (PI - T[SIN_DU + idx1((PI - clampd(uc, PI, 3 * PI / 2)) / -0.01)] + k * 2 * PI) *
                     (double)(uc >= PI && uc <= 3 * PI / 2);
        k = floor(l / (2 * PI));
        double lc = l - k * (2 * PI);
        dup += (T[SIN_DU + idx1((clampd(lc, 3 * PI / 2, 2 * PI) - 2 * PI) / -0.01)] + (k + 1) * 2 * PI) *
               (double)(lc >= 3 * PI / 2 && lc <= 2 * PI);
        dlo += (PI - T[SIN_DL + idx1((PI - clampd(lc, PI / 2, 3 * PI)) / 0.01)] + k * 2 * PI) *
               (double)(lc >= PI / 2 && lc <= PI);
        bool mdl = (cc ? kd < gh_cos(l) : kd > gh_cos(u)) && mboth;
        bool mdu = (cc ? kd < gh_cos(u) : kd > gh_cos(l)) && mboth;
        double m = (l + u) / 2, ym = gh_sin(m), km = gh_cos(m);
        if (mneg) put(lw, lb, km, m, ym);
        if (mpos) put(uw, ub, km, m, ym);
        if (mdl) put(lw, lb, kd, l, yl);
        if (mboth && !mdl) put(lw, lb, gh_cos(dlo), dlo, gh_sin(dlo));
        if (mdu) put(uw, ub, kd, l, yl);
        if (mboth && !mdu) put(uw, ub, gh_cos(dup), dup, gh_sin(dup));
        // the smoothness rule of the S-shaped pieces (s_shaped.py:513-539)
        if ((yl < yu && dlo >= m) || (yl >= yu && dlo < m)) put(lw, lb, km, m, ym);
        if ((yl < yu && dup < m) || (yl >= yu && dup >= m)) put(uw, ub, km, m, ym);
        R[0] = lw; R[1] = lb; R[2] = uw; R[3] = ub;
    }
    double ro = 8 * EPS64 * (1 + fmax(fabs(l), fabs(u)));
    R[1] = R[1] - ro;
    R[3] = R[3] + ro;
    return 0;
}

GH_SS int relax_cos(const double *T, double l, double u, double R[4])
{
    const double h = 0.5 * PI;
    if (!isfinite(l) || !isfinite(u)) return 3;
    int ud = relax_sin(T, l + h, u + h, R);
    R[1] = R[1] + R[0] * h;
    R[3] = R[3] + R[2] * h;
    return ud;
}

// odd p >= 3 with its table block D (d_lower at 0, d_upper at POW_D)
GH_SS int relax_pow_odd(const double *D, int p, double l, double u, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    u = fmax(u, l + 1e-8);
    const double c2 = (double)(p * (p - 1));
    double d2l = c2 * gh_powi(l, p - 2), d2u = c2 * gh_powi(u, p - 2);
    bool mneg = d2l >= 0 && d2u >= 0, mpos = d2l < 0 && d2u < 0, mboth = !(mneg || mpos);
    bool cc = d2l >= 0;
    double lw = 0, lb = 0, uw = 0, ub = 0;
    double yl = fpow(p, l), yu = fpow(p, u);
    double width = u - l;
    double kd = (yu - yl) / fmax(width, 1e-8);
    if (width < 1e-8) kd = dfpow(p, l);
    if (mneg) put(uw, ub, kd, l, yl);
    if (mpos) put(lw, lb, kd, l, yl);
    int64_t iu = idx1(u / 0.01), il = idx1(-l / 0.01);
    double dup = iu < POW_D ? D[POW_D + iu] : l;
    double dlo = il < POW_D ? D[il] : u;
    bool mdl = (cc ? kd < dfpow(p, l) : kd > dfpow(p, u)) && mboth;
    bool mdu = (cc ? kd < dfpow(p, u) : kd > dfpow(p, l)) && mboth;
    double m = (l + u) / 2, ym = fpow(p, m), km = dfpow(p, m);
    if (mneg) put(lw, lb, km, m, ym);
    if (mpos) put(uw, ub, km, m, ym);
    if (mdl) put(lw, lb, kd, l, yl);
    if (mboth && !mdl) put(lw, lb, dfpow(p, dlo), dlo, fpow(p, dlo));
    if (mdu) put(uw, ub, kd, l, yl);
    if (mboth && !mdu) put(uw, ub, dfpow(p, dup), dup, fpow(p, dup));
    if (mboth && dlo < m) put(lw, lb, km, m, ym);   // inverse S shape
    if (mboth && dup >= m) put(uw, ub, km, m, ym);
    R[0] = lw; R[1] = lb; R[2] = uw; R[3] = ub;
    return 0;
}

GH_SS int relax_pow_frac(double p, double l, double u, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    u = fmax(u, l + 1e-8);
    if (l < 0) l = 0.0;
    if (u < l) return 3;
    double lt = (p > 0 && p < 1) ? fmax(l, 1e-9) : l;
    double fl = gh_powf(l, p), fu = gh_powf(u, p);
    double ks = (fu - fl) / fmax(u - l, 1e-8);
    double m = (lt + u) / 2, kt, fm;
    if (p > 1 && m == 0) {
        kt = 0.0;
        fm = 0.0;
    } else {
        kt = p * gh_powf(m, p - 1);
        fm = gh_powf(m, p);
    }
    if (p > 0 && p < 1) {
        put(R[0], R[1], ks, l, fl);
        put(R[2], R[3], kt, m, fm);
    } else {
        put(R[0], R[1], kt, m, fm);
        put(R[2], R[3], ks, l, fl);
    }
    return 0;
}

}  // namespace ss
}  // namespace gh
