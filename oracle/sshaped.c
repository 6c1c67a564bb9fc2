/*
 * CPU oracle (test infrastructure only) -- plain C restatement of the
 * reference's relaxations of sin, cos, odd integer powers and non-integer
 * powers, with the tangent tables they read.  Only tests, the smoke check
 * and the benchmark's CPU-baseline leg load this code (inside
 * liboracle.so); it is the checker, never the product path.
 *
 * Arithmetic follows the reference's tensor expressions operation by
 * operation (IEEE double, no contraction); torch's sin / cos / pow are
 * replaced by the deterministic functions of gearhip_math.h, so results
 * agree with the reference's to a few units in the last place except where
 * a comparison sits within that distance of a tie.
 *
 * Sin (auto_LiRPA/operators/trigonometric.py, BoundSin, on top of
 * s_shaped.py BoundSShaped):
 *   - d tables: BoundSShaped.precompute_relaxation (s_shaped.py:125-202)
 *     with x_limit = pi/2: 162 tangent points each, 100-step bisection;
 *   - endpoint tables over linspace(0, 2 pi, 1001): the 20-step bisection
 *     tangents get_lower_left_bound / get_upper_left_bound /
 *     get_lower_right_bound / get_upper_right_bound (trigonometric.py:
 *     240-321), built by _ensure_precomputed (:49-69);
 *   - bound_relax (:71-85) = BoundSShaped.bound_relax (s_shaped.py:557-574):
 *     branch_input_domain (trigonometric.py:110-153), bound_relax_impl
 *     (s_shaped.py:325-549, vanilla CROWN branch with the smoothness
 *     enhancement), bound_relax_branch (trigonometric.py:469-506) with
 *     get_bound_tb (:323-440) and get_intersection / n_crossing (:207-222);
 *     then the roundoff margin 8 eps (1 + max(|l|, |u|)).
 *   The reference's golden bounds (solver/tests/data_64/s_shape_test_data)
 *   were written before two of its later changes: the right-endpoint table
 *   index rounded up (get_bound_tb :339-343) and the roundoff margin.
 *   GHO_SIN_LEGACY_INDEX selects the earlier index rounding so the tests can
 *   pin the tables and every branch against that file exactly.
 * Cos: BoundCos.bound_relax (trigonometric.py:509-527): sin's relaxation on
 *   the interval shifted by pi/2, the biases shifted back.
 * Odd powers (s_shaped.py BoundPow, exponent p odd >= 3): tables
 *   BoundPow.precompute_relaxation (:694-770, x_limit 500: 50,005 tangent
 *   points each), generate_d_lower_upper (:652-664), retrieve_from_precompute
 *   (:223-248), branch_input_domain (:666-691), then bound_relax_impl with
 *   inverse_s_shape set (:913-916, 500-512); the interval's upper end is
 *   first raised to lower + 1e-8 (:795-797).
 * Non-integer powers p > 0: bound_relax_noninteger (:974-1037) on the
 *   interval restricted to x >= 0 (domain_aware_ops, :640-650 and
 *   base.py:258-286).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sshaped.h"
#include "gearhip_math.h"

#define PI 3.141592653589793
#define EPS64 2.220446049250313080847e-16
#define SIN_TABLE 1001
#define SIN_D 162         /* int((pi/2) / 0.01) + 5 */
#define POW_D 50005       /* int(500 / 0.01) + 5 */

/* torch's double -> int64 conversion (x86 cvttsd2si: out of range -> INT64_MIN) */
static int64_t to_long(double q)
{
    if (!(q > -9.2233720368547758e18 && q < 9.2233720368547758e18)) return INT64_MIN;
    return (int64_t)q;
}
static double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

/* ------------------------------------------------------------ sin tables */
static double sin_dl[SIN_D], sin_du[SIN_D];
static double tab_x[SIN_TABLE];
static double xll_d[SIN_TABLE], xll_c[SIN_TABLE], xlu_d[SIN_TABLE], xlu_c[SIN_TABLE];
static double xul_d[SIN_TABLE], xul_c[SIN_TABLE], xuu_d[SIN_TABLE], xuu_c[SIN_TABLE];
static int sin_ready;

/* check_lower / check_upper of BoundSShaped.precompute_relaxation (s_shaped.py:144-156) */
static int sin_check_lower(double upper, double d)
{
    double k = gh_cos(d);
    return k * (upper - d) + gh_sin(d) <= gh_sin(upper);
}
static int sin_check_upper(double lower, double d)
{
    double k = gh_cos(d);
    return k * (lower - d) + gh_sin(d) >= gh_sin(lower);
}

/* BoundSin.check_bound (trigonometric.py:224-233) */
static double check_bound(double t, double x)
{
    double d = gh_cos(t);
    double val = d * (x - t) + gh_sin(t);
    return gh_sin(x) - val;
}

/* get_lower_left_bound (trigonometric.py:235-259) */
static void lower_left(double xl, double *d, double *cyc)
{
    double cycles = floor((xl + 0.5 * PI) / (2 * PI)) * (2 * PI);
    xl = xl - cycles;
    int use = xl >= PI;
    double left = PI, right = 1.5 * PI, last = right;
    for (int s = 0; s < 20; ++s) {
        double mid = (left + right) / 2.;
        if (check_bound(mid, xl) > 0) { right = mid; last = mid; }
        else left = mid;
    }
    *d = use ? xl : last;
    *cyc = cycles;
}

/* get_upper_left_bound (trigonometric.py:261-285) */
static void upper_left(double xl, double *d, double *cyc)
{
    double cycles = floor((xl - 0.5 * PI) / (2 * PI)) * (2 * PI);
    xl = xl - cycles;
    int use = xl >= 2.0 * PI;
    double left = 2.0 * PI, right = 2.5 * PI, last = right;
    for (int s = 0; s < 20; ++s) {
        double mid = (left + right) / 2.;
        if (check_bound(mid, xl) > 0) left = mid;
        else { right = mid; last = mid; }
    }
    *d = use ? xl : last;
    *cyc = cycles;
}

/* get_lower_right_bound / get_upper_right_bound (trigonometric.py:287-321) */
static void lower_right(double xu, double *d, double *cyc)
{
    double cycles = floor((xu + 0.5 * PI) / (2 * PI)) * (2 * PI);
    xu = xu - cycles;
    double t, c;
    lower_left(PI - xu, &t, &c);
    *d = 3 * PI - t;
    *cyc = cycles - 2 * PI;
}
static void upper_right(double xu, double *d, double *cyc)
{
    double cycles = floor((xu - 0.5 * PI) / (2 * PI)) * (2 * PI);
    xu = xu - cycles;
    double t, c;
    upper_left(3 * PI - xu, &t, &c);
    *d = 5 * PI - t;
    *cyc = cycles - 2 * PI;
}

static void sin_build(void)
{
    if (sin_ready) return;
    for (int i = 0; i < SIN_D; ++i) {
        double up = 0.01 * (double)i, l = -1.0, r = 0.0;
        while (!sin_check_lower(up, l)) l = l * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (sin_check_lower(up, m)) l = m;
            else r = m;
        }
        sin_dl[i] = l;
        double lo = -0.01 * (double)i;
        l = 0.0;
        r = 1.0;
        while (!sin_check_upper(lo, r)) r = r * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (sin_check_upper(lo, m)) r = m;
            else l = m;
        }
        sin_du[i] = r;
    }
    /* torch.linspace(0, 2 pi, 1001): step (end - start) / (steps - 1), the
     * first half from the start, the second from the end, each one fma */
    double step = (2 * PI - 0.0) / (SIN_TABLE - 1);
    for (int i = 0; i < SIN_TABLE; ++i)
        tab_x[i] = i < SIN_TABLE / 2 ? fma(step, (double)i, 0.0)
                                     : fma(-step, (double)(SIN_TABLE - 1 - i), 2 * PI);
    for (int i = 0; i < SIN_TABLE; ++i) {
        lower_left(tab_x[i], &xll_d[i], &xll_c[i]);
        upper_left(tab_x[i], &xlu_d[i], &xlu_c[i]);
        lower_right(tab_x[i], &xul_d[i], &xul_c[i]);
        upper_right(tab_x[i], &xuu_d[i], &xuu_c[i]);
    }
    sin_ready = 1;
}

/* ------------------------------------------------------ sin relaxation */
/* BoundSin.generate_d_lower_upper (trigonometric.py:155-205) */
static void sin_gen_d(double lower, double upper, double *dlo, double *dup)
{
    double k = floor(upper / (2 * PI));
    double uc = upper - k * (2 * PI);
    double c1 = (uc >= 0 && uc <= PI / 2) ? 1.0 : 0.0;
    double ucn = clampd(uc, 0, PI / 2);
    int64_t idx = to_long(ucn / 0.01);
    idx = (idx > 0 ? idx : 0) + 1;
    double d_lower = (sin_dl[idx] + k * 2 * PI) * c1;
    double c2 = (uc >= PI && uc <= 3 * PI / 2) ? 1.0 : 0.0;
    ucn = clampd(uc, PI, 3 * PI / 2);
    idx = to_long((PI - ucn) / -0.01);
    idx = (idx > 0 ? idx : 0) + 1;
    double d_upper = (PI - sin_du[idx] + k * 2 * PI) * c2;
    k = floor(lower / (2 * PI));
    double lc = lower - k * (2 * PI);
    double c3 = (lc >= 3 * PI / 2 && lc <= 2 * PI) ? 1.0 : 0.0;
    double lcn = clampd(lc, 3 * PI / 2, 2 * PI);
    idx = to_long((lcn - 2 * PI) / -0.01);
    idx = (idx > 0 ? idx : 0) + 1;
    d_upper += (sin_du[idx] + (k + 1) * 2 * PI) * c3;
    double c4 = (lc >= PI / 2 && lc <= PI) ? 1.0 : 0.0;
    lcn = clampd(lc, PI / 2, 3 * PI);
    idx = to_long((PI - lcn) / 0.01);
    idx = (idx > 0 ? idx : 0) + 1;
    d_lower += (PI - sin_dl[idx] + k * 2 * PI) * c4;
    *dlo = d_lower;
    *dup = d_upper;
}

/* n_crossing / get_intersection (trigonometric.py:207-222) */
static double n_crossing(double start, double end, double s)
{
    double cycles = floor((end - start) / (2 * PI));
    double dist = floor((s - start) / (2 * PI));
    double real_s = s - dist * 2 * PI;
    double real_end = end - cycles * 2 * PI;
    return (double)(real_s >= start) * (double)(real_s <= real_end) + cycles;
}
static double get_intersection(double start, double end, double c, double theta)
{
    double c1 = gh_asin(c) - theta;
    double c2 = PI - c1 - 2 * theta;
    return n_crossing(start, end, c1) + n_crossing(start, end, c2);
}

/* the tangent chosen between the two endpoint tangents (get_bound_tb :356-413) */
static double tb_pick(double mid, double lower, double upper, double tl, double tr, double cl,
                      double cr, double lbc, double ubc, double end)
{
    double ul = tl + cl + lbc, le = end + cl + lbc;
    double ur = tr + cr + ubc, re = end + cr + ubc;
    int lm = mid < ul && ul <= upper;
    int lr = mid >= ul && mid < le;
    int ib = mid >= le && mid < re;
    int rm = mid >= ur && ur >= lower;
    int rr = !lr && (mid >= re && mid < ur);
    double t = (double)lm * tl + (double)lr * (mid - cl - lbc) + (double)ib * end +
               (double)rm * tr + (double)rr * (mid - cr - ubc);
    return (lm || lr || ib || rm || rr) ? t : tl;
}

/* get_bound_tb (trigonometric.py:323-440), vanilla CROWN branch */
static void get_bound_tb(double lb, double ub, int legacy, double *dl, double *bl, double *du,
                         double *bu)
{
    double lower = lb, upper = ub;
    double step = 2 * PI / (SIN_TABLE - 1);
    double lbc = floor(lb / (2 * PI)) * (2 * PI);
    lb = clampd(lb - lbc, 0, 2 * PI);
    double ubc = floor(ub / (2 * PI)) * (2 * PI);
    ub = clampd(ub - ubc, 0, 2 * PI);
    int64_t il = to_long(lb / step);
    int64_t iu = legacy ? to_long(ub / step) : to_long(ceil(ub / step));
    if (iu > SIN_TABLE - 1) iu = SIN_TABLE - 1;
    double mid = (lower + upper) / 2;
    double tlo = tb_pick(mid, lower, upper, xll_d[il], xul_d[iu], xll_c[il], xul_c[iu], lbc, ubc,
                         1.5 * PI);
    double tup = tb_pick(mid, lower, upper, xlu_d[il], xuu_d[iu], xlu_c[il], xuu_c[iu], lbc, ubc,
                         2.5 * PI);
    double d = gh_cos(tlo);
    *dl = d;
    *bl = gh_sin(tlo) - d * (tlo + (tlo <= 1.5 * PI ? xll_c[il] + lbc : xul_c[iu] + ubc));
    d = gh_cos(tup);
    *du = d;
    *bu = gh_sin(tup) - d * (tup + (tup <= 2.5 * PI ? xlu_c[il] + lbc : xuu_c[iu] + ubc));
}

/* bound_relax_branch (trigonometric.py:469-506) */
static void sin_branch(double lb, double ub, int legacy, double R[4])
{
    ub = fmax(ub, lb + 1e-8);
    double sub = gh_sin(ub), slb = gh_sin(lb);
    double mid = (sub + slb) / 2.;
    double smid = gh_sin((ub + lb) / 2);
    double gap = smid - mid;
    double k1 = (sub - slb) / fmax(ub - lb, 1e-10);
    double b1 = slb - k1 * lb;
    double gc = get_intersection(lb, ub, k1, 0.5 * PI);
    int use = gc == 1;
    int uul = gap < 0 && use, lul = gap >= 0 && use;
    double dl, bl, du, bu;
    get_bound_tb(lb, ub, legacy, &dl, &bl, &du, &bu);
    R[0] = lul ? k1 : dl;
    R[1] = lul ? b1 : bl;
    R[2] = uul ? k1 : du;
    R[3] = uul ? b1 : bu;
}

/* add_linear_relaxation (activation_base.py:49-82): w = k, b = -x0 k + y0 */
static void put(double *w, double *b, double k, double x0, double y0)
{
    *w = k;
    *b = -x0 * k + y0;
}

int gho_relax_sin(double l, double u, int flags, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    sin_build();
    int legacy = (flags & GHO_SIN_LEGACY_INDEX) != 0;
    double lw = 0, lb = 0, uw = 0, ub = 0;
    /* branch_input_domain (trigonometric.py:110-153) */
    double lbc = l - floor(l / (2 * PI)) * (2 * PI);
    double ubc = u - floor(u / (2 * PI)) * (2 * PI);
#define QUAD(x, a, b) ((x) >= (a) && (x) < (b))
    int l1 = QUAD(lbc, 0, PI / 2), l2 = QUAD(lbc, PI / 2, PI), l3 = QUAD(lbc, PI, 3 * PI / 2),
        l4 = QUAD(lbc, 3 * PI / 2, 2 * PI);
    int u1 = QUAD(ubc, 0, PI / 2), u2 = QUAD(ubc, PI / 2, PI), u3 = QUAD(ubc, PI, 3 * PI / 2),
        u4 = QUAD(ubc, 3 * PI / 2, 2 * PI);
#undef QUAD
    int sl = (u - l <= PI) && (((l2 || l3) && (u2 || u3)) || ((l1 || l4) && (u1 || u4)));
    int mneg = (l3 || l4) && (u3 || u4) && sl;
    int mpos = (l1 || l2) && (u1 || u2) && sl;
    int mboth = sl != (mneg || mpos);
    int cc = -gh_sin(l) >= 0;
    /* bound_relax_impl (s_shaped.py:325-549), vanilla branch */
    double yl = gh_sin(l), yu = gh_sin(u);
    double width = u - l;
    double kd = (yu - yl) / fmax(width, 1e-8);
    if (width < 1e-8) kd = gh_cos(l);
    if (mneg) put(&uw, &ub, kd, l, yl);
    if (mpos) put(&lw, &lb, kd, l, yl);
    double dlo, dup;
    sin_gen_d(l, u, &dlo, &dup);
    int mdl = (cc ? kd < gh_cos(l) : kd > gh_cos(u)) && mboth;
    int mdu = (cc ? kd < gh_cos(u) : kd > gh_cos(l)) && mboth;
    double m = (l + u) / 2, ym = gh_sin(m), km = gh_cos(m);
    if (mneg) put(&lw, &lb, km, m, ym);
    if (mpos) put(&uw, &ub, km, m, ym);
    if (mdl) put(&lw, &lb, kd, l, yl);
    if (mboth && !mdl) put(&lw, &lb, gh_cos(dlo), dlo, gh_sin(dlo));
    if (mdu) put(&uw, &ub, kd, l, yl);
    if (mboth && !mdu) put(&uw, &ub, gh_cos(dup), dup, gh_sin(dup));
    /* smoothness enhancement on the sigmoid-like pieces (:513-539) */
    if (sl && yl < yu && dlo >= m) put(&lw, &lb, km, m, ym);
    if (sl && yl < yu && dup < m) put(&uw, &ub, km, m, ym);
    if (sl && yl >= yu && dlo < m) put(&lw, &lb, km, m, ym);
    if (sl && yl >= yu && dup >= m) put(&uw, &ub, km, m, ym);
    if (!sl) {
        double Rb[4];
        sin_branch(l, u, legacy, Rb);
        lw = Rb[0]; lb = Rb[1]; uw = Rb[2]; ub = Rb[3];
    }
    /* roundoff margin (trigonometric.py:78-85) */
    double ro = 8 * EPS64 * (1 + fmax(fabs(l), fabs(u)));
    R[0] = lw;
    R[1] = lb - ro;
    R[2] = uw;
    R[3] = ub + ro;
    return 0;
}

int gho_relax_cos(double l, double u, int flags, double R[4])
{
    const double half_pi = 0.5 * PI;
    if (!isfinite(l) || !isfinite(u)) return 3;
    int ud = gho_relax_sin(l + half_pi, u + half_pi, flags, R);
    R[1] = R[1] + R[0] * half_pi;
    R[3] = R[3] + R[2] * half_pi;
    return ud;
}

/* ------------------------------------------------------ odd powers */
typedef struct { int p; double *dl, *du; } pow_tab;
static pow_tab pow_tabs[4];

static double fpow(int p, double x) { return gh_powi(x, p); }
static double dfpow(int p, double x) { return (double)p * gh_powi(x, p - 1); }

/* BoundPow.precompute_relaxation (s_shaped.py:694-770) */
static const pow_tab *pow_build(int p)
{
    for (int k = 0; k < 4; ++k)
        if (pow_tabs[k].p == p) return &pow_tabs[k];
    int slot = -1;
    for (int k = 0; k < 4; ++k)
        if (pow_tabs[k].p == 0) { slot = k; break; }
    if (slot < 0) return NULL;
    double *dl = (double *)malloc(sizeof(double) * POW_D);
    double *du = (double *)malloc(sizeof(double) * POW_D);
    for (int i = 0; i < POW_D; ++i) {
        /* given an upper end >= 0, the tangent point whose line is above f there */
        double up = 0.01 * (double)i, l = -1.0, r = 0.0;
#define CHECK_UPPER(lo, d) (dfpow(p, d) * ((lo) - (d)) + fpow(p, d) >= fpow(p, lo))
#define CHECK_LOWER(up_, d) (dfpow(p, d) * ((up_) - (d)) + fpow(p, d) <= fpow(p, up_))
        while (!CHECK_UPPER(up, l)) l = l * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (CHECK_UPPER(up, m)) l = m;
            else r = m;
        }
        du[i] = l;
        /* given a lower end <= 0, the tangent point whose line is below f there */
        double lo = -0.01 * (double)i;
        l = 0.0;
        r = 1.0;
        while (!CHECK_LOWER(lo, r)) r = r * 2;
        for (int it = 0; it < 100; ++it) {
            double m = (l + r) / 2;
            if (CHECK_LOWER(lo, m)) r = m;
            else l = m;
        }
        dl[i] = r;
#undef CHECK_UPPER
#undef CHECK_LOWER
    }
    pow_tabs[slot].dl = dl;
    pow_tabs[slot].du = du;
    pow_tabs[slot].p = p;
    return &pow_tabs[slot];
}

/* retrieve_from_precompute (s_shaped.py:223-248) */
static double retrieve(const double *tab, double bound, double def)
{
    int64_t idx = to_long(bound / 0.01);
    idx = (idx > 0 ? idx : 0) + 1;
    return idx < POW_D ? tab[idx] : def;
}

int gho_relax_pow_odd(int p, double l, double u, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    const pow_tab *T = pow_build(p);
    if (!T) return 3;
    u = fmax(u, l + 1e-8);
    double lw = 0, lb = 0, uw = 0, ub = 0;
    /* branch_input_domain (s_shaped.py:666-691): one inflection at 0 */
    double d2l = (double)(p * (p - 1)) * gh_powi(l, p - 2);
    double d2u = (double)(p * (p - 1)) * gh_powi(u, p - 2);
    int mneg = d2l >= 0 && d2u >= 0;
    int mpos = d2l < 0 && d2u < 0;
    int mboth = !(mneg || mpos);
    int cc = d2l >= 0;
    double yl = fpow(p, l), yu = fpow(p, u);
    double width = u - l;
    double kd = (yu - yl) / fmax(width, 1e-8);
    if (width < 1e-8) kd = dfpow(p, l);
    if (mneg) put(&uw, &ub, kd, l, yl);
    if (mpos) put(&lw, &lb, kd, l, yl);
    /* generate_d_lower_upper, odd exponent (s_shaped.py:652-662) */
    double dup = retrieve(T->du, u, l);
    double dlo = retrieve(T->dl, -l, u);
    int mdl = (cc ? kd < dfpow(p, l) : kd > dfpow(p, u)) && mboth;
    int mdu = (cc ? kd < dfpow(p, u) : kd > dfpow(p, l)) && mboth;
    double m = (l + u) / 2, ym = fpow(p, m), km = dfpow(p, m);
    if (mneg) put(&lw, &lb, km, m, ym);
    if (mpos) put(&uw, &ub, km, m, ym);
    if (mdl) put(&lw, &lb, kd, l, yl);
    if (mboth && !mdl) put(&lw, &lb, dfpow(p, dlo), dlo, fpow(p, dlo));
    if (mdu) put(&uw, &ub, kd, l, yl);
    if (mboth && !mdu) put(&uw, &ub, dfpow(p, dup), dup, fpow(p, dup));
    /* inverse S shape (s_shaped.py:500-512) */
    if (mboth && dlo < m) put(&lw, &lb, km, m, ym);
    if (mboth && dup >= m) put(&uw, &ub, km, m, ym);
    R[0] = lw; R[1] = lb; R[2] = uw; R[3] = ub;
    return 0;
}

/* --------------------------------------------------- non-integer powers */
int gho_relax_pow_frac(double p, double l, double u, double R[4])
{
    if (!isfinite(l) || !isfinite(u)) return 3;
    u = fmax(u, l + 1e-8);                 /* s_shaped.py:805 */
    if (l < 0) l = 0.0;                    /* domain_aware_ops restriction */
    if (u < l) return 3;                   /* no point of the domain */
    double lt = (p > 0 && p < 1) ? fmax(l, 1e-9) : l;
    double fl = gh_powf(l, p), fu = gh_powf(u, p);
    double ks = (fu - fl) / fmax(u - l, 1e-8);
    double m = (lt + u) / 2;
    double kt, fm;
    if (p > 1 && m == 0) { kt = 0.0; fm = 0.0; }
    else { kt = p * gh_powf(m, p - 1); fm = gh_powf(m, p); }
    if (p > 0 && p < 1) {
        put(&R[0], &R[1], ks, l, fl);
        put(&R[2], &R[3], kt, m, fm);
    } else {
        put(&R[0], &R[1], kt, m, fm);
        put(&R[2], &R[3], ks, l, fl);
    }
    return 0;
}

/* The tables, for their tests: which 0 sin d_lower / d_upper (162 each),
 * 1 sin endpoint tables (x, then d and cycles of xl_lower, xl_upper,
 * xu_lower, xu_upper: 9 x 1001), 2 odd power p d_lower / d_upper (50,005
 * each).  Returns the number of doubles written (out may be NULL). */
int gho_sshape_tables(int which, int p, double *out)
{
    if (which == 0) {
        sin_build();
        if (out) { memcpy(out, sin_dl, sizeof sin_dl); memcpy(out + SIN_D, sin_du, sizeof sin_du); }
        return 2 * SIN_D;
    }
    if (which == 1) {
        sin_build();
        const double *t[9] = {tab_x, xll_d, xll_c, xlu_d, xlu_c, xul_d, xul_c, xuu_d, xuu_c};
        if (out)
            for (int k = 0; k < 9; ++k) memcpy(out + k * SIN_TABLE, t[k], sizeof(double) * SIN_TABLE);
        return 9 * SIN_TABLE;
    }
    if (which == 2) {
        const pow_tab *T = pow_build(p);
        if (!T) return 0;
        if (out) { memcpy(out, T->dl, sizeof(double) * POW_D); memcpy(out + POW_D, T->du, sizeof(double) * POW_D); }
        return 2 * POW_D;
    }
    return 0;
}

/* One unary relaxation over arrays, for the tests: op as gearhip_program.h,
 * c its constant; R (N, 4) = lw, lb, uw, ub; returns the OR of the
 * undefined sides. */
int gho_relax_unary(int op, double c, int N, const double *l, const double *u, double *R,
                    int flags)
{
    int any = 0;
    for (int k = 0; k < N; ++k) {
        int ud = 3;
        if (op == GH_OP_SIN) ud = gho_relax_sin(l[k], u[k], flags, R + 4 * k);
        else if (op == GH_OP_COS) ud = gho_relax_cos(l[k], u[k], flags, R + 4 * k);
        else if (op == GH_OP_POW && c == floor(c) && ((int)c) % 2 == 1)
            ud = gho_relax_pow_odd((int)c, l[k], u[k], R + 4 * k);
        else if (op == GH_OP_POW && c != floor(c))
            ud = gho_relax_pow_frac(c, l[k], u[k], R + 4 * k);
        any |= ud;
    }
    return any;
}

void gho_sshape_init(int pow_mask)
{
    sin_build();
    for (int p = 3; p < 32; p += 2)
        if (pow_mask & (1 << p)) pow_build(p);
}
