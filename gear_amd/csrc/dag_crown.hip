// Persistent FP64 CROWN bound of expression programs (gfx950).
//
// One wavefront bounds one row of one program over 64 boxes (a lane per
// box) by walking the row's tape (dag_tape.h) with wave-uniform control flow:
// tape entries are uniform loads, per-lane state is [slot][lane] in the
// wavefront's workspace.  Wavefronts stay resident and pull (row, 64-box
// chunk) items from a device counter until the items run out, so uneven
// rows and uneven batches of many programs share one launch.
//
// Per box the arithmetic follows the FP64 restatement of the reference's
// CROWN exactly, in the same order (intermediate bounds per node in
// program order, a backward pass per input of a relaxed node intersected
// with its interval, one lower pass per row; bound_general.py:1032-1222,
// backward_bound.py:880-1206), with the elementary functions of
// gearhip_math.h, so each bound is bit-identical to the host restatement.
#include <hip/hip_runtime.h>

#include "common.h"
#include "dag_tape.h"
#include "gearhip_internal.h"
#include "sshape.h"
#include "../../include/gearhip_math.h"

namespace {

using gh::DagFwd;
using gh::DagJob;
using gh::DagPass;
using gh::DagStep;
using gh::finite_abs;
using gh::widen;

constexpr double EPS64 = 2.220446049250313080847e-16;
constexpr double EPSILON = 1e-12;  // auto_LiRPA operators/base.py:32
constexpr double PI = 3.14159265358979323846;
constexpr double H1 = (double)(1 + 1) * EPS64 / 2.0 * (1.0 + 1e-6);  // gamma_1
__host__ __device__ constexpr double kround(int n) { return (double)(n + 16) * EPS64 * 4.0; }
constexpr double K1 = kround(1);
constexpr double K2 = kround(2);
constexpr int kLanes = 64;
// LDS slots per wavefront at most (64 slots = 32 KB)
constexpr int kMaxLdsSlots = 64;

__device__ __forceinline__ double add_enc(double a, double b, double s)
{
    return gh::add_enclosed(a, b, s, K1);
}
__device__ __forceinline__ double pos(double a) { return a > 0 ? a : 0.0; }
__device__ __forceinline__ double neg(double a) { return a < 0 ? a : 0.0; }
// a * d with 0 * d = 0 (multiply_by_A_signs keeps zeros exact)
__device__ __forceinline__ double mulz(double a, double d) { return a == 0 ? 0.0 : a * d; }
__device__ __forceinline__ double re1(double m) { return m * K1; }

// Per-lane slot view of the wavefront workspace.
// Slot a of the A area [ab, ab + ka) lives in LDS (the most accessed slots:
// every pass reads and writes them), every other slot in the wavefront's
// global workspace.  Slots are wave-uniform, so the choice is a uniform branch.
struct Lane {
    double *w;    // workspace base + lane
    double *lds;  // LDS area + lane
    int ab, ka;
    __device__ __forceinline__ double &operator[](int slot) const
    {
        const unsigned q = (unsigned)(slot - ab);
        if (q < (unsigned)ka) return lds[(size_t)q * kLanes];
        return w[(size_t)slot * kLanes];
    }
};

__device__ __forceinline__ bool crosses(double l, double u, double point)
{
    // trigonometric.py:447-456
    double lc = (l - point) / (2 * PI), uc = (u - point) / (2 * PI);
    double ro = 2 * EPS64 * (1 + fmax(fabs(lc), fabs(uc)));
    return (u - l >= 2 * PI) || (floor(uc + ro) > floor(lc - ro));
}

__device__ __forceinline__ void enclose_add(double &l, double &u)
{
    // interval_bound.py:81-95
    l = widen(l, finite_abs(l) * K1, -1.0);
    u = widen(u, finite_abs(u) * K1, 1.0);
}

// interval bound of one node from its operands' current bounds
__device__ void ibp(int op, bool self, double la, double ua, double lb_, double ub_, double c,
                    double &lo, double &hi)
{
    double l = 0, u = 0;
    switch (op) {
    case GH_OP_CONST: l = u = c; break;
    case GH_OP_ADD: l = la + lb_; u = ua + ub_; enclose_add(l, u); break;
    case GH_OP_SUB: l = la - ub_; u = ua - lb_; enclose_add(l, u); break;
    case GH_OP_ADDC: l = la + c; u = ua + c; enclose_add(l, u); break;
    case GH_OP_NEG: l = -ua; u = -la; break;
    case GH_OP_MULC:  // bivariate.py:519-536
        if (c > 0) { l = la * c; u = ua * c; }
        else if (c < 0) { l = ua * c; u = la * c; }
        else { l = u = 0.0; }
        break;
    case GH_OP_MUL:
        if (self) {  // bivariate.py:568-578
            double m = pos(la) - pos(-ua);
            l = m * m;
            u = fmax(la * la, ua * ua);
            if (isnan(u)) u = INFINITY;
        } else {
            double r0 = la * lb_, r1 = la * ub_, r2 = ua * lb_, r3 = ua * ub_;
            l = fmin(fmin(r0, r1), fmin(r2, r3));
            u = fmax(fmax(r0, r1), fmax(r2, r3));
            if (isnan(r0) || isnan(r1) || isnan(r2) || isnan(r3)) { l = -INFINITY; u = INFINITY; }
        }
        break;
    case GH_OP_SQR:  // activations.py:221-232
        l = ua < 0 ? ua * ua : (la > 0 ? la * la : 0.0);
        u = fmax(la * la, ua * ua);
        break;
    case GH_OP_POW: {  // s_shaped.py:1088-1130
        if (c != floor(c)) {  // non-integer: increasing on the domain x >= 0
            l = gh_powf(la > 0 ? la : 0.0, c);
            u = gh_powf(ua, c);
            break;
        }
        double pl = gh_powi(la, (int)c), pu = gh_powi(ua, (int)c);
        if (((int)c) % 2 == 0) {
            l = (la <= 0 && ua >= 0) ? 0.0 : fmin(pl, pu);
            u = fmax(pl, pu);
        } else {
            l = pl;
            u = pu;
        }
        break;
    }
    case GH_OP_EXP: l = gh_exp(la); u = gh_exp(ua); break;  // convex_concave.py:418-427
    case GH_OP_LOG:  // convex_concave.py:92-103, restricted to x >= 0
        l = gh_log(la > 0 ? la : 0.0);
        u = gh_log(ua);
        break;
    case GH_OP_RECIP:  // convex_concave.py:286-313
        l = 1.0 / ua;
        u = 1.0 / la;
        if ((la < 0 && ua > 0) || (la < 0 && ua == 0) || (la == 0 && ua == 0)) l = -INFINITY;
        if ((la < 0 && ua > 0) || (la == 0 && ua == 0) || (la == 0 && ua > 0)) u = INFINITY;
        break;
    case GH_OP_SIN: {  // trigonometric.py:445-467
        double sl = gh_sin(la), su = gh_sin(ua);
        u = crosses(la, ua, PI / 2) ? 1.0 : fmax(sl, su);
        l = crosses(la, ua, 3 * PI / 2) ? -1.0 : fmin(sl, su);
        break;
    }
    case GH_OP_COS: {  // the extrema of cos at 0 and pi (trigonometric.py:509-514)
        double sl = gh_cos(la), su = gh_cos(ua);
        u = crosses(la, ua, 0.0) ? 1.0 : fmax(sl, su);
        l = crosses(la, ua, PI) ? -1.0 : fmin(sl, su);
        break;
    }
    default: l = -INFINITY; u = INFINITY;
    }
    if (isnan(l)) l = -INFINITY;
    if (isnan(u)) u = INFINITY;
    lo = l;
    hi = u;
}

// Relaxation planes of a relaxed node from its operands' bounds.
// R: unary (lw, lb, uw, ub); MUL (alpha_l, beta_l, gamma_l, alpha_u,
// beta_u, gamma_u).  Returns the undefined sides (1 lower, 2 upper).
__device__ int relax(int op, double l, double u, double yl, double yu, double c, double il,
                     double iu, const double *T, double R[6])
{
    for (int k = 0; k < 6; ++k) R[k] = 0.0;
    switch (op) {
    case GH_OP_MUL: {  // bivariate.py:190-261, 40-75, 288-334
        double xl = l, xu = u;
        if (!isfinite(xl) || !isfinite(xu) || !isfinite(yl) || !isfinite(yu)) return 3;
        double m1 = yl * (yl >= 0 ? xl : xu) + xl * (xl >= 0 ? yl : yu) + (-xl) * yl;
        double m2 = yu * (yu >= 0 ? xl : xu) + xu * (xu >= 0 ? yl : yu) + (-xu) * yu;
        double rl = (m1 >= m2) ? 1.0 : 0.0;
        double M1 = yu * (yu >= 0 ? xu : xl) + xl * (xl >= 0 ? yu : yl) + (-yu) * xl;
        double M2 = yl * (yl >= 0 ? xu : xl) + xu * (xu >= 0 ? yu : yl) + (-yl) * xu;
        double ru = (M1 <= M2) ? 1.0 : 0.0;
        R[0] = (yl - yu) * rl + yu;
        R[1] = (xl - xu) * rl + xu;
        R[2] = (yu * xu - yl * xl) * rl - yu * xu;
        R[3] = (yu - yl) * ru + yl;
        R[4] = (xl - xu) * ru + xu;
        R[5] = (yl * xu - yu * xl) * ru - yl * xu;
        return 0;
    }
    case GH_OP_SQR: {  // activations.py:186-208
        if (!isfinite(l) || !isfinite(u)) return 2;
        double k = l + u;
        R[2] = k;
        R[3] = -l * k + l * l;
        double mid = pos(l) - pos(-u);
        R[0] = 2 * mid;
        R[1] = -mid * (2 * mid) + mid * mid;
        return 0;
    }
    case GH_OP_POW: {
        if (c != floor(c)) return gh::ss::relax_pow_frac(c, l, u, R);
        int p = (int)c;
        if (!isfinite(l) || !isfinite(u)) return (p % 2 == 0) ? 2 : 3;
        if (p % 2 == 1) return gh::ss::relax_pow_odd(T, p, l, u, R);
        u = fmax(u, l + 1e-8);  // s_shaped.py:795
        // even power: bound_relax_even_power (s_shaped.py:940-972)
        double yl_ = gh_powi(l, p), yu_ = gh_powi(u, p);
        double eps = fmax(fabs(yl_), fabs(yu_)) * 1e-6 + 1e-12;
        double k = (yu_ - yl_) / fmax(u - l, 1e-8);
        R[2] = k;
        R[3] = -l * k + (yl_ + eps);
        double mid = l >= 0 ? l : (u <= 0 ? u : 0.0);
        double dk = p * gh_powi(mid, p - 1);
        R[0] = dk;
        R[1] = -mid * dk + (gh_powi(mid, p) - eps);
        return 0;
    }
    case GH_OP_EXP: {  // convex_concave.py:569-625
        if (!isfinite(l) || !isfinite(u)) return 3;
        double du = u - l;
        double ne = -gh_expm1(l - u);
        double mp = u + gh_log(fmax(ne, EPSILON)) - gh_log(fmax(du, EPSILON));
        double m = du > 1e-6 ? mp : (l + u) / 2;
        m = fmin(fmax(m, l), u);
        double el = gh_exp(l), em = gh_exp(m), eu = gh_exp(u);
        double lk = isfinite(em) ? em : 0.0;
        double lx = isfinite(m) ? m : 0.0;
        double lbb = -lx * lk + lk;
        bool ls = isfinite(lk) && isfinite(lbb);
        R[0] = ls ? lk : 0.0;
        R[1] = ls ? lbb : 0.0;
        double d = u - l, k;
        if (d < 1) k = el * (d > 0 ? gh_expm1(fmin(d, 1.0)) / fmax(d, 1e-300) : 1.0);
        else k = (eu - el) / fmax(d, 1.0);
        double ubb = -l * k + el;
        bool us = isfinite(k) && isfinite(ubb);
        R[2] = us ? k : 0.0;
        R[3] = us ? ubb : INFINITY;
        return 0;
    }
    case GH_OP_LOG: {  // convex_concave.py:64-92
        double lr = l > 0 ? l : 0.0;
        if (!(lr > 0) || !isfinite(u)) return 3;
        double rl = gh_log(lr), ru = gh_log(u), du = u - lr;
        double ku = (ru - rl) / fmax(du, EPSILON);
        R[0] = ku;
        R[1] = -lr * ku + rl;
        double dl = fmax(ru - rl, EPSILON);
        double m = du > 1e-6 ? du / dl : (lr + u) / 2;
        m = fmin(fmax(m, lr), u);
        double k = 1.0 / m;
        R[2] = k;
        R[3] = -m * k + gh_log(m);
        return 0;
    }
    case GH_OP_RECIP: {  // convex_concave.py:315-375
        if (l <= 0 && u >= 0) return 3;
        if (!isfinite(l) || !isfinite(u)) return 3;
        double sk = -1.0 / (l * u);
        double mid = (l + u) / 2, tk = -1.0 / (mid * mid);
        if (l > 0) {
            R[2] = sk;
            R[3] = -l * sk + 1.0 / l;
            R[0] = tk;
            R[1] = -mid * tk + 1.0 / mid;
        } else {
            R[0] = sk;
            R[1] = -u * sk + 1.0 / u;
            R[2] = tk;
            R[3] = -mid * tk + 1.0 / mid;
        }
        return 0;
    }
    case GH_OP_SIN: return gh::ss::relax_sin(T, l, u, R);
    case GH_OP_COS: return gh::ss::relax_cos(T, l, u, R);
    default: return 3;
    }
}

__device__ __forceinline__ double mag_of(Lane W, int V, int v)
{
    double l = W[v], u = W[V + v];
    return (l > u) ? INFINITY : fmax(-l, u);
}

// A contribution x into slot s of node u; returns its merge charge.
__device__ __forceinline__ double push(Lane W, int s, double x, int kind, int V, int u)
{
    if (kind == gh::PUSH_FIRST) {
        W[s] = x;
        return 0.0;
    }
    double total = W[s] + x;
    W[s] = total;
    if (kind == gh::PUSH_DEFER) return 0.0;
    double m = mag_of(W, V, u);
    return isfinite(m) ? H1 * (fabs(total) * m) : (total != 0 ? INFINITY : 0.0);
}

__device__ __forceinline__ double inexact_abs(double c)
{
    int e;
    double m = frexp(c, &e);
    return (c == 0 || fabs(m) == 0.5) ? 0.0 : fabs(c);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// One backward pass; writes its lower (and upper) bound.
__device__ void crown_pass(const DagPass *passes, const DagStep *steps, const int32_t *cin,
                           int pidx, Lane W, int V, int n, double &out_l, double &out_u,
                           double &bias_l, int &undef_l)
{
    const DagPass P = passes[pidx];
    const bool want_u = P.want_u != 0;
    W[P.sstart] = 1.0;
    if (want_u) W[P.sstart + 1] = 1.0;
    double lacc = 0.0, uacc = 0.0, sxl = 0.0, sxu = 0.0;
    int und = 0;
    const DagStep *S = steps + P.step_off;
    for (int k = 0; k < P.nsteps; ++k) {
        const DagStep st = S[k];
        const int op = uni(st.op), v = uni(st.v), sv = uni(st.sv);
        double al = W[sv];
        double au = want_u ? W[sv + 1] : 0.0;
        if (op == GH_OP_INPUT) {
            // x_i's A into the root: merge charged as gamma_1 <|A_x|, m_x>
            double mx = finite_abs(fmax(-W[v], W[V + v]));
            double nl = 0.0 + al, lb_ = 0.0;
            W[sv] = nl;
            if (st.flags & gh::STEP_FIRST_INPUT) {
                sxl = fabs(nl) * mx;
            } else {
                sxl = sxl + fabs(nl) * mx;
                lb_ = -H1 * sxl;
            }
            lacc = add_enc(lacc, lb_, -1.0);
            if (want_u) {
                double nu = 0.0 + au, ub_ = 0.0;
                W[sv + 1] = nu;
                if (st.flags & gh::STEP_FIRST_INPUT) {
                    sxu = fabs(nu) * mx;
                } else {
                    sxu = sxu + fabs(nu) * mx;
                    ub_ = H1 * sxu;
                }
                uacc = add_enc(uacc, ub_, 1.0);
            }
            continue;
        }
        if (op == GH_OP_CONST) {
            // add_constant_node (backward_bound.py:1604-1625)
            double c = st.c;
            lacc = add_enc(lacc, widen(al * c, re1(fabs(al) * fabs(c)), -1.0), -1.0);
            if (want_u) uacc = add_enc(uacc, widen(au * c, re1(fabs(au) * fabs(c)), 1.0), 1.0);
            continue;
        }
        const int a = uni(st.a), b = uni(st.b), sa = uni(st.sa), sb = uni(st.sb);
        const int ka = uni(st.ka), kb = uni(st.kb);
        double lb_ = 0.0, ub_ = 0.0, el = 0.0, eu = 0.0;
        if (st.flags & gh::STEP_DEFERRED) {  // backward_bound.py:801-806
            double m = mag_of(W, V, v);
            el += isfinite(m) ? H1 * fabs(al) * m : (al != 0 ? INFINITY : 0.0);
            if (want_u) eu += isfinite(m) ? H1 * fabs(au) * m : (au != 0 ? INFINITY : 0.0);
        }
        switch (op) {
        case GH_OP_ADD:
        case GH_OP_SUB:
        case GH_OP_ADDC: {  // add_sub.py:70-238, rounding of the operands (backward_bound.py:1057-1070)
            double sg = (op == GH_OP_SUB) ? -1.0 : 1.0;
            double ma = finite_abs(fmax(fabs(W[a]), fabs(W[V + a])));
            double mb = (op == GH_OP_ADDC) ? fabs(st.c) : finite_abs(fmax(fabs(W[b]), fabs(W[V + b])));
            double err = ((0.0 + ma) + mb) * K2;
            lb_ = lb_ - finite_abs(al) * err;
            if (want_u) ub_ = ub_ + finite_abs(au) * err;
            el += push(W, sa, al, ka, V, a);
            if (want_u) eu += push(W, sa + 1, au, ka, V, a);
            if (op != GH_OP_ADDC) {
                el += push(W, sb, sg * al, kb, V, b);
                if (want_u) eu += push(W, sb + 1, sg * au, kb, V, b);
            } else {
                double c = st.c;
                lb_ = lb_ + widen(mulz(al, c), re1(fabs(al) * fabs(c)), -1.0);
                if (want_u) ub_ = ub_ + widen(mulz(au, c), re1(fabs(au) * fabs(c)), 1.0);
            }
            break;
        }
        case GH_OP_NEG:
            el += push(W, sa, -al, ka, V, a);
            if (want_u) eu += push(W, sa + 1, -au, ka, V, a);
            break;
        case GH_OP_MULC: {  // product by a constant, A67-charged when inexact
            double c = st.c, ie = inexact_abs(c), m = mag_of(W, V, a);
            el += (ie == 0) ? 0.0 : (isfinite(m) ? H1 * (fabs(al) * (ie * m)) : INFINITY);
            el += push(W, sa, al * c, ka, V, a);
            if (want_u) {
                eu += (ie == 0) ? 0.0 : (isfinite(m) ? H1 * (fabs(au) * (ie * m)) : INFINITY);
                eu += push(W, sa + 1, au * c, ka, V, a);
            }
            break;
        }
        case GH_OP_MUL: {  // bivariate.py:401-488
            const int rs = uni(st.rs);
            double R0 = W[rs], R1 = W[rs + 1], R2 = W[rs + 2], R3 = W[rs + 3], R4 = W[rs + 4],
                   R5 = W[rs + 5];
            int ud = (int)W[rs + 6];
            if (ud) und |= 1;
            double ax = mulz(pos(al), R0) + mulz(neg(al), R3);
            double ay = mulz(pos(al), R1) + mulz(neg(al), R4);
            lb_ = mulz(pos(al), R2) + mulz(neg(al), R5);
            el += push(W, sa, ax, ka, V, a);
            el += push(W, sb, ay, kb, V, b);
            if (want_u) {
                if (ud) und |= 2;
                double bx = mulz(pos(au), R3) + mulz(neg(au), R0);
                double by = mulz(pos(au), R4) + mulz(neg(au), R1);
                ub_ = mulz(pos(au), R5) + mulz(neg(au), R2);
                eu += push(W, sa + 1, bx, ka, V, a);
                eu += push(W, sb + 1, by, kb, V, b);
            }
            break;
        }
        default: {  // relaxed unary (activation_base.py:137-182)
            const int rs = uni(st.rs);
            double R0 = W[rs], R1 = W[rs + 1], R2 = W[rs + 2], R3 = W[rs + 3];
            int ud = (int)W[rs + 4];
            if ((ud & 1) && al > 0) und |= 1;
            if ((ud & 2) && al < 0) und |= 1;
            double ax = mulz(pos(al), R0) + mulz(neg(al), R2);
            lb_ = mulz(pos(al), R1) + mulz(neg(al), R3);
            el += push(W, sa, ax, ka, V, a);
            if (want_u) {
                if ((ud & 2) && au > 0) und |= 2;
                if ((ud & 1) && au < 0) und |= 2;
                double bx = mulz(pos(au), R2) + mulz(neg(au), R0);
                ub_ = mulz(pos(au), R3) + mulz(neg(au), R1);
                eu += push(W, sa + 1, bx, ka, V, a);
            }
            break;
        }
        }
        lacc = add_enc(lacc, lb_ - el, -1.0);
        if (want_u) uacc = add_enc(uacc, ub_ + eu, 1.0);
    }
    // box concretization (concretize_bounds.py:267-321, fp_enclosure.py:60-82)
    const double kn = kround(n);
    double bl = widen(lacc, finite_abs(lacc) * kn, -1.0);
    double base_l = 0.0, et_l = 0.0, ma_l = 0.0, base_u = 0.0, et_u = 0.0, ma_u = 0.0;
    for (int k = 0; k < P.ncin; ++k) {
        const int e = uni(cin[P.cin_off + k]);
        const int v = e >> 16, s = e & 0xFFFF;
        double l = W[v], u = W[V + v];
        double x0 = (u + l) / 2.0, ep = (u - l) / 2.0;
        double mx = finite_abs(fmax(-l, u));
        double Al = W[s];
        base_l = fma(Al, x0, base_l);
        et_l = fma(fabs(Al), ep, et_l);
        ma_l = fma(finite_abs(Al), mx, ma_l);
        if (want_u) {
            double Au = W[s + 1];
            base_u = fma(Au, x0, base_u);
            et_u = fma(fabs(Au), ep, et_u);
            ma_u = fma(finite_abs(Au), mx, ma_u);
        }
    }
    double rl = widen(base_l - et_l, ma_l * kn, -1.0);
    if (!isfinite(rl)) rl = -INFINITY;
    double lo = add_enc(bl, rl, -1.0);
    if (isnan(lo) || (und & 1)) lo = -INFINITY;
    out_l = lo;
    out_u = INFINITY;
    bias_l = lacc;
    undef_l = und & 1;
    if (want_u) {
        double bu = widen(uacc, finite_abs(uacc) * kn, 1.0);
        double ru = widen(base_u + et_u, ma_u * kn, 1.0);
        if (!isfinite(ru)) ru = INFINITY;
        double hi = add_enc(bu, ru, 1.0);
        if (isnan(hi) || (und & 2)) hi = INFINITY;
        out_u = hi;
    }
}

// Bound one row of one program over 64 boxes (one per lane).
__device__ void run_job(const DagJob &J, const DagFwd *fwd, const DagPass *passes,
                        const DagStep *steps, const int32_t *cin, const int32_t *gin,
                        const double *tabs, const gh_dag_batch &bt, int64_t count,
                        int64_t chunk, Lane W)
{
    const int lane = threadIdx.x;
    const int64_t box = chunk * kLanes + lane;
    const bool valid = box < count;
    const int64_t bx = valid ? box : count - 1;
    const int V = J.V, ni = J.ni, n = J.n;
    for (int k = 0; k < ni; ++k) {
        const int g = uni(gin[J.gin_off + k]);
        W[k] = bt.xL[bx * n + g];
        W[V + k] = bt.xU[bx * n + g];
    }
    for (int k = ni; k < V; ++k) {
        const DagFwd f = fwd[J.fwd_off + (k - ni)];
        const int op = uni(f.op), a = uni(f.a), b = uni(f.b), rs = uni(f.rs), pass = uni(f.pass);
        const bool binary = op == GH_OP_ADD || op == GH_OP_SUB || op == GH_OP_MUL;
        double la = W[a], ua = W[V + a];
        double lb_ = binary ? W[b] : 0.0, ub_ = binary ? W[V + b] : 0.0;
        double il, iu;
        ibp(op, a == b, la, ua, lb_, ub_, f.c, il, iu);
        W[k] = il;
        W[V + k] = iu;
        if (rs >= 0) {
            double R[6];
            int ud = relax(op, la, ua, lb_, ub_, f.c, il, iu, tabs + uni(f.tab), R);
            if (op == GH_OP_MUL) {
                for (int q = 0; q < 6; ++q) W[rs + q] = R[q];
                W[rs + 6] = (double)ud;
            } else {
                for (int q = 0; q < 4; ++q) W[rs + q] = R[q];
                W[rs + 4] = (double)ud;
            }
        }
        if (pass >= 0) {
            // CROWN bound intersected with the interval (bound_general.py:1160-1185)
            double cl, cu, bias;
            int und;
            crown_pass(passes, steps, cin, pass, W, V, n, cl, cu, bias, und);
            W[k] = fmax(cl, il);
            W[V + k] = fmin(cu, iu);
        }
    }
    double out, bias = 0.0;
    int und = 0;
    if (J.final_pass < 0) {
        out = W[J.out];
    } else {
        double cl, cu;
        crown_pass(passes, steps, cin, J.final_pass, W, V, n, cl, cu, bias, und);
        // a row with no relaxation falls back to its interval (backward_bound.py:826-877)
        out = (cl == -INFINITY) ? W[J.out] : fmax(cl, -INFINITY);
    }
    if (!valid) return;
    const size_t br = (size_t)box * J.R + J.row;
    bt.lb[br] = out;
    if (bt.lA) {
        // the row's plane over the program's inputs: zeros, then the A of
        // the inputs the row reads (a bare input: its unit vector)
        double *a = bt.lA + br * n;
        for (int i = 0; i < n; ++i) a[i] = 0.0;
        if (J.final_pass < 0) {
            a[uni(gin[J.gin_off + J.out])] = 1.0;
        } else if (!und) {
            const DagPass P = passes[J.final_pass];
            for (int k = 0; k < P.ncin; ++k) {
                const int e = uni(cin[P.cin_off + k]);
                a[uni(gin[J.gin_off + (e >> 16)])] = W[e & 0xFFFF];
            }
        }
        bt.lbias[br] = und ? -INFINITY : bias;
        if (bt.undef) bt.undef[br] = (uint8_t)und;
    }
}

// Chunk counts per job and their exclusive prefix sum; resets the queue.
__global__ void __launch_bounds__(1024)
dag_schedule_kernel(const DagJob *jobs, int64_t njobs, const gh_dag_batch *batches,
                    int64_t *job_start, int64_t *counter)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    int64_t carry = 0;
    for (int64_t base = 0; base < njobs; base += 1024) {
        const int64_t j = base + t;
        int64_t c = 0;
        if (j < njobs) {
            int64_t cnt = *batches[jobs[j].prog].count;
            c = cnt > 0 ? (cnt + kLanes - 1) / kLanes : 0;
        }
        part[t] = c;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int64_t add = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += add;
            __syncthreads();
        }
        if (j < njobs) job_start[j] = carry + part[t] - c;
        carry += part[1023];
        __syncthreads();
    }
    if (t == 0) {
        job_start[njobs] = carry;
        *counter = 0;
    }
}

__global__ void __launch_bounds__(64)
dag_crown_kernel(const DagJob *__restrict__ jobs, int64_t njobs, const DagFwd *__restrict__ fwd,
                 const DagPass *__restrict__ passes,
                 const DagStep *__restrict__ steps, const int32_t *__restrict__ cin,
                 const int32_t *__restrict__ gin, const double *__restrict__ tabs,
                 const gh_dag_batch *__restrict__ batches,
                 const int64_t *__restrict__ job_start, int64_t *counter, double *ws,
                 int64_t slots, int lds_slots)
{
    extern __shared__ double sh_a[];
    Lane W{ws + (size_t)blockIdx.x * (size_t)slots * kLanes + threadIdx.x, sh_a + threadIdx.x, 0,
           lds_slots};
    const int64_t total = job_start[njobs];
    for (;;) {
        unsigned long long item = 0;
        if (threadIdx.x == 0) item = atomicAdd(reinterpret_cast<unsigned long long *>(counter), 1ull);
        const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)item);
        const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(item >> 32));
        const int64_t it = (int64_t)(((unsigned long long)hi32 << 32) | lo32);
        if (it >= total) break;
        // the job holding item `it`: largest j with job_start[j] <= it
        int64_t lo = 0, hi = njobs - 1;
        while (lo < hi) {
            int64_t mid = (lo + hi + 1) >> 1;
            if (job_start[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const DagJob J = jobs[lo];
        const gh_dag_batch bt = batches[J.prog];
        const int64_t count = *bt.count;
        W.ab = uni(J.abase);
        run_job(J, fwd, passes, steps, cin, gin, tabs, bt, count, it - job_start[lo], W);
    }
}

size_t sched_bytes(int64_t njobs) { return (((size_t)(njobs + 2) * 8) + 255) & ~size_t(255); }

}  // namespace

extern "C" size_t gh_dag_workspace_bytes(const gh_dag_plan *plan, int waves)
{
    if (!plan || waves < 1) return 0;
    return sched_bytes(plan->njobs) + (size_t)waves * (size_t)plan->max_slots * kLanes * sizeof(double);
}

extern "C" int gh_dag_crown_f64(const gh_dag_plan *plan, const gh_dag_batch *batches, void *ws,
                                size_t ws_bytes, int waves, int lds_slots, hipStream_t stream)
{
    if (!plan || waves < 1 || lds_slots < 0 || lds_slots > kMaxLdsSlots) return GH_EINVAL;
    if (plan->njobs == 0) return GH_OK;
    if (!batches || !ws || ws_bytes < gh_dag_workspace_bytes(plan, waves)) return GH_EINVAL;
    int64_t *job_start = static_cast<int64_t *>(ws);
    int64_t *counter = job_start + plan->njobs + 1;
    double *slots = reinterpret_cast<double *>(static_cast<char *>(ws) + sched_bytes(plan->njobs));
    hipLaunchKernelGGL(dag_schedule_kernel, dim3(1), dim3(1024), 0, stream, plan->jobs,
                       plan->njobs, batches, job_start, counter);
    const size_t lds = (size_t)lds_slots * kLanes * sizeof(double);
    hipLaunchKernelGGL(dag_crown_kernel, dim3(waves), dim3(kLanes), lds, stream, plan->jobs,
                       plan->njobs, plan->fwd, plan->passes, plan->steps, plan->cin, plan->gin,
                       plan->tabs, batches, job_start, counter, slots, plan->max_slots, lds_slots);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
