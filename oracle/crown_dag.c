/*
 * CPU oracle (test infrastructure only) -- plain C restatement of the
 * reference's FP64 CROWN for scalar expression programs: nonlinear programs
 * whose rows are expression DAGs over x in a box.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may load
 * this library; it is the checker, never the product path.
 *
 * Program: nodes in topological order; nodes 0..n-1 are the inputs x_i (the
 * reference's Gather / slice of the input vector), then operations
 *   CONST c, ADD a b, SUB a b, NEG a, MULC a c (a * constant), ADDC a c
 *   (a + constant), MUL a b (both perturbed), SQR a, POW a p (an integer
 *   p >= 3 or a non-integer p > 0), EXP a, LOG a, RECIP a, SIN a, COS a
 * (division is lowered to MUL(a, RECIP(b)) as optimize_graph.py:103-122
 * does; x ** 2 to SQR as convert_sqr does, optimize_graph.py:124).
 * Row r of the specification is node out[r], C = +1, lower bounds.
 *
 * Per box (auto_LiRPA/bound_general.py:1032-1222, backward_bound.py:880-1206):
 *   - nodes are visited in index order; a node's interval is interval
 *     arithmetic from its inputs' current bounds (interval_bound.py:60-95,
 *     each operator's interval_propagate);
 *   - a node that feeds a relaxed operator (MUL, SQR, POW, EXP, LOG, RECIP,
 *     SIN) gets CROWN bounds from a backward pass over its ancestors with
 *     both sides, intersected with its interval (compare_crown_with_ibp_
 *     intermediate, bound_general.py:1183-1185);
 *   - a relaxed node's planes are built from its inputs' bounds (the
 *     operators' bound_relax, cited below);
 *   - each row gets a final lower-only pass.
 * A backward pass visits the ancestors of its start node in decreasing
 * index order (a reverse topological order) and applies, per node, the
 * operator's bound_backward, the Add/Sub operand-rounding charge
 * (backward_bound.py:1057-1070), the A67 charges of products by inexact
 * constants and of A merges (backward_bound.py:611-823) and the add_bound
 * widening of the accumulated bias (fp_enclosure.py:94-98); then box
 * concretization with its margins (concretize_bounds.py:267-321,
 * fp_enclosure.py:60-82).  A relaxation that is undefined for the pass
 * (an interval through a pole, an infinite endpoint) makes that side of the
 * pass vacuous, and the row falls back to its interval bound
 * (backward_bound.py:295-463, 826-877).
 *
 * The relaxations of sin, cos, odd integer powers and non-integer powers,
 * with their tangent tables, are in sshaped.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "crown_dag.h"
#include "gearhip_math.h"
#include "sshaped.h"

#define EPS64 2.220446049250313080847e-16
#define EPSILON 1e-12  /* auto_LiRPA operators/base.py:32 */

static double k_round(long n) { return (double)(n + 16) * EPS64 * 4.0; }
static double h_gamma(long k) { return (double)(k + 1) * EPS64 / 2.0 * (1.0 + 1e-6); }
static double finite_abs(double v) { return isfinite(v) ? fabs(v) : 0.0; }
static double widen(double b, double e, double s)
{
    if (isinf(e) && e > 0) return s * INFINITY;
    return isfinite(b) ? b + s * e : b;
}
static double add_enclosed(double a, double b, double s)
{
    double r = a + b;
    return widen(r, finite_abs(r) * k_round(1), s);
}
/* rounding_error(m, 1) */
static double rounding_error_1(double m) { return m * k_round(1); }
static double pos(double a) { return a > 0 ? a : 0.0; }
static double neg(double a) { return a < 0 ? a : 0.0; }
/* a * d with 0 * d = 0 (multiply_by_A_signs keeps zeros exact) */
static double mulz(double a, double d) { return a == 0 ? 0.0 : a * d; }

static int is_relaxed(int op)
{
    return op == GH_OP_MUL || op == GH_OP_SQR || op == GH_OP_POW || op == GH_OP_EXP ||
           op == GH_OP_LOG || op == GH_OP_RECIP || op == GH_OP_SIN || op == GH_OP_COS;
}

/* ---------------------------------------------------------------- IBP */
static void enclose_add(double *l, double *u)
{
    /* interval_bound.py:81-95 (n = #inputs - 1 = 1) */
    *l = widen(*l, finite_abs(*l) * k_round(1), -1.0);
    *u = widen(*u, finite_abs(*u) * k_round(1), 1.0);
}

static int crosses(double l, double u, double point)
{
    /* trigonometric.py:447-456 */
    double lc = (l - point) / (2 * M_PI), uc = (u - point) / (2 * M_PI);
    double ro = 2 * EPS64 * (1 + fmax(fabs(lc), fabs(uc)));
    return (u - l >= 2 * M_PI) || (floor(uc + ro) > floor(lc - ro));
}

static void ibp_node(const gh_node *nd, const double *L, const double *U, double *lo, double *hi)
{
    double la = L[nd->a], ua = U[nd->a];
    double lb_ = (nd->op == GH_OP_ADD || nd->op == GH_OP_SUB || nd->op == GH_OP_MUL) ? L[nd->b] : 0;
    double ub_ = (nd->op == GH_OP_ADD || nd->op == GH_OP_SUB || nd->op == GH_OP_MUL) ? U[nd->b] : 0;
    double c = nd->c, l = 0, u = 0;
    switch (nd->op) {
    case GH_OP_CONST: l = u = c; break;
    case GH_OP_ADD: l = la + lb_; u = ua + ub_; enclose_add(&l, &u); break;
    case GH_OP_SUB: l = la - ub_; u = ua - lb_; enclose_add(&l, &u); break;
    case GH_OP_ADDC: l = la + c; u = ua + c; enclose_add(&l, &u); break;
    case GH_OP_NEG: l = -ua; u = -la; break;
    case GH_OP_MULC: /* bivariate.py:519-536 */
        if (c > 0) { l = la * c; u = ua * c; }
        else if (c < 0) { l = ua * c; u = la * c; }
        else { l = u = 0.0; }
        break;
    case GH_OP_MUL:
        if (nd->a == nd->b) { /* bivariate.py:568-578 */
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
    case GH_OP_SQR: /* activations.py:221-232 */
        l = ua < 0 ? ua * ua : (la > 0 ? la * la : 0.0);
        u = fmax(la * la, ua * ua);
        break;
    case GH_OP_POW: {  /* s_shaped.py:1088-1130 */
        if (c != floor(c)) {  /* non-integer: increasing on the domain x >= 0 */
            l = gh_powf(la > 0 ? la : 0.0, c);
            u = gh_powf(ua, c);
            break;
        }
        double pl = gh_powi(la, (int)c), pu = gh_powi(ua, (int)c);
        if (((int)c) % 2 == 0) {
            l = (la <= 0 && ua >= 0) ? 0.0 : fmin(pl, pu);
            u = fmax(pl, pu);
        } else { l = pl; u = pu; }
        break;
    }
    case GH_OP_EXP: l = gh_exp(la); u = gh_exp(ua); break;  /* convex_concave.py:418-427 */
    case GH_OP_LOG: /* convex_concave.py:92-103: restrict to x >= 0 */
        l = gh_log(la > 0 ? la : 0.0); u = gh_log(ua); break;
    case GH_OP_RECIP: /* convex_concave.py:286-313 */
        l = 1.0 / ua; u = 1.0 / la;
        if ((la < 0 && ua > 0) || (la < 0 && ua == 0) || (la == 0 && ua == 0)) l = -INFINITY;
        if ((la < 0 && ua > 0) || (la == 0 && ua == 0) || (la == 0 && ua > 0)) u = INFINITY;
        break;
    case GH_OP_SIN: { /* trigonometric.py:445-467 */
        double sl = gh_sin(la), su = gh_sin(ua);
        u = crosses(la, ua, M_PI / 2) ? 1.0 : fmax(sl, su);
        l = crosses(la, ua, M_PI * 3 / 2) ? -1.0 : fmin(sl, su);
        break;
    }
    case GH_OP_COS: { /* the same with the extrema of cos (trigonometric.py:509-514) */
        double sl = gh_cos(la), su = gh_cos(ua);
        u = crosses(la, ua, 0.0) ? 1.0 : fmax(sl, su);
        l = crosses(la, ua, M_PI) ? -1.0 : fmin(sl, su);
        break;
    }
    default: l = -INFINITY; u = INFINITY;
    }
    if (isnan(l)) l = -INFINITY;
    if (isnan(u)) u = INFINITY;
    *lo = l;
    *hi = u;
}

/* -------------------------------------------------------- relaxations */
/* GHO_SIN_LEGACY_INDEX (sshaped.h) or 0; set by gho_dag_set_flags */
static int dag_flags;

void gho_dag_set_flags(int flags) { dag_flags = flags; }

/* R[0..5]: unary (lw, lb, uw, ub); MUL (al, bl, gl, au, bu, gu).
 * returns bit 1: lower side undefined, bit 2: upper side undefined. */
static int relax_node(const gh_node *nd, const double *L, const double *U, double *R)
{
    double l = L[nd->a], u = U[nd->a];
    memset(R, 0, sizeof(double) * 6);
    switch (nd->op) {
    case GH_OP_MUL: { /* bivariate.py:190-261, 40-75, 288-334 */
        double xl = l, xu = u, yl = L[nd->b], yu = U[nd->b];
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
    case GH_OP_SQR: { /* activations.py:186-208 */
        if (!isfinite(l) || !isfinite(u)) return 2;  /* convex tangent lower survives */
        double k = l + u;
        R[2] = k; R[3] = -l * k + l * l;
        double mid = pos(l) - pos(-u);
        R[0] = 2 * mid; R[1] = -mid * (2 * mid) + mid * mid;
        return 0;
    }
    case GH_OP_POW: {
        if (nd->c != floor(nd->c)) return gho_relax_pow_frac(nd->c, l, u, R);
        int p = (int)nd->c;
        if (!isfinite(l) || !isfinite(u)) return (p % 2 == 0) ? 2 : 3;
        u = fmax(u, l + 1e-8);  /* s_shaped.py:795 */
        if (p % 2 == 0) {  /* bound_relax_even_power, s_shaped.py:940-972 */
            double yl_ = gh_powi(l, p), yu_ = gh_powi(u, p);
            double eps = fmax(fabs(yl_), fabs(yu_)) * 1e-6 + 1e-12;
            double k = (yu_ - yl_) / fmax(u - l, 1e-8);
            R[2] = k; R[3] = -l * k + (yl_ + eps);
            double mid = l >= 0 ? l : (u <= 0 ? u : 0.0);
            double dk = p * gh_powi(mid, p - 1);
            R[0] = dk; R[1] = -mid * dk + (gh_powi(mid, p) - eps);
            return 0;
        }
        return gho_relax_pow_odd(p, l, u, R);
    }
    case GH_OP_EXP: { /* convex_concave.py:569-625 */
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
        int ls = isfinite(lk) && isfinite(lbb);
        R[0] = ls ? lk : 0.0; R[1] = ls ? lbb : 0.0;
        double d = u - l, k;
        if (d < 1) k = el * (d > 0 ? gh_expm1(fmin(d, 1.0)) / fmax(d, 1e-300) : 1.0);
        else k = (eu - el) / fmax(d, 1.0);
        double ubb = -l * k + el;
        int us = isfinite(k) && isfinite(ubb);
        R[2] = us ? k : 0.0; R[3] = us ? ubb : INFINITY;
        return 0;
    }
    case GH_OP_LOG: { /* convex_concave.py:64-92 */
        double lr = l > 0 ? l : 0.0;
        if (!(lr > 0) || !isfinite(u)) return 3;
        double rl = gh_log(lr), ru = gh_log(u), du = u - lr;
        double ku = (ru - rl) / fmax(du, EPSILON);
        R[0] = ku; R[1] = -lr * ku + rl;
        double dl = fmax(ru - rl, EPSILON);
        double m = du > 1e-6 ? du / dl : (lr + u) / 2;
        m = fmin(fmax(m, lr), u);
        double k = 1.0 / m;
        R[2] = k; R[3] = -m * k + gh_log(m);
        return 0;
    }
    case GH_OP_RECIP: { /* convex_concave.py:315-375 */
        if (l <= 0 && u >= 0) return 3;
        if (!isfinite(l) || !isfinite(u)) return 3;
        double sk = -1.0 / (l * u);
        double mid = (l + u) / 2, tk = -1.0 / (mid * mid);
        if (l > 0) {
            R[2] = sk; R[3] = -l * sk + 1.0 / l;
            R[0] = tk; R[1] = -mid * tk + 1.0 / mid;
        } else {
            R[0] = sk; R[1] = -u * sk + 1.0 / u;
            R[2] = tk; R[3] = -mid * tk + 1.0 / mid;
        }
        return 0;
    }
    case GH_OP_SIN: return gho_relax_sin(l, u, dag_flags, R);
    case GH_OP_COS: return gho_relax_cos(l, u, dag_flags, R);
    default: return 3;
    }
}

/* ------------------------------------------------------- backward pass */
typedef struct {
    double *L, *U;          /* current bounds per node */
    double *R;              /* 6 per node */
    int *udef;              /* relaxation undefined bits per node */
    double *Al, *Au;        /* A per node for this pass */
    int *hasl, *hasu;       /* A set (not None) */
    int *rem, *anc, *dfl, *dfu;
    double *Axl, *Axu;      /* root A */
    int *hasx;
} dag_ws;

static double inexact_abs(double c)
{
    int e;
    double m = frexp(c, &e);
    return (c == 0 || fabs(m) == 0.5) ? 0.0 : fabs(c);
}

static double mag_of(const dag_ws *w, int v)
{
    double l = w->L[v], u = w->U[v];
    double m = (l > u) ? INFINITY : fmax(-l, u);
    return m;
}

/* push A into input v (one side); returns the merge charge for that side */
static double push_A(dag_ws *w, int v, double a, int side, int later, int is_input_root)
{
    double *A = side ? w->Au : w->Al;
    int *has = side ? w->hasu : w->hasl;
    int *df = side ? w->dfu : w->dfl;
    double charge = 0.0;
    if (has[v] && !is_input_root) {
        double total = A[v] + a;
        int last = !later && w->rem[v] == 0;
        double m = mag_of(w, v);
        if (last) df[v] = 1;
        else if (isfinite(m)) charge = h_gamma(1) * (fabs(total) * m);
        else charge = (total != 0) ? INFINITY : 0.0;
        A[v] = total;
    } else if (has[v]) {
        A[v] = A[v] + a;
    } else {
        A[v] = a;
        has[v] = 1;
    }
    return charge;
}

/* CROWN pass from node s: lower (and upper) bound of node s over the box.
 * Returns bits: 1 lower vacuous, 2 upper vacuous. */
static void crown_pass(const gh_prog *P, dag_ws *w, const double *xL, const double *xU, int s,
                       int want_u, double *out_l, double *out_u, double *plane_A,
                       double *plane_b, int *undef)
{
    const gh_node *N = P->nodes;
    int n = P->n;
    for (int v = 0; v <= s; ++v) {
        w->anc[v] = 0; w->rem[v] = 0; w->Al[v] = w->Au[v] = 0.0;
        w->hasl[v] = w->hasu[v] = 0; w->dfl[v] = w->dfu[v] = 0;
    }
    w->anc[s] = 1;
    for (int v = s; v >= n; --v) {
        if (!w->anc[v]) continue;
        const gh_node *nd = &N[v];
        int op = nd->op;
        if (op == GH_OP_CONST) continue;
        w->anc[nd->a] = 1; w->rem[nd->a]++;
        if (op == GH_OP_ADD || op == GH_OP_SUB || op == GH_OP_MUL) { w->anc[nd->b] = 1; w->rem[nd->b]++; }
    }
    for (int i = 0; i < n; ++i) { w->Axl[i] = w->Axu[i] = 0.0; }
    int hasx_l = 0, hasx_u = 0;
    double sxl = 0.0, sxu = 0.0;  /* running sum |Ax| m over the root */
    w->Al[s] = 1.0; w->hasl[s] = 1;
    if (want_u) { w->Au[s] = 1.0; w->hasu[s] = 1; }
    double lacc = 0.0, uacc = 0.0;
    int und = 0;
    for (int v = s; v >= 0; --v) {
        if (!w->anc[v]) continue;
        const gh_node *nd = &N[v];
        int op = nd->op;
        double al = w->Al[v], au = w->Au[v];
        int hl = w->hasl[v], hu = w->hasu[v];
        if (op == GH_OP_INPUT) {
            /* Gather of x_i into the root: its merge into x's A is charged
             * as gamma_1 <|A_x|, m_x> (the root is always charged at once). */
            int i = nd->a;
            double mx = finite_abs(fmax(-xL[i], xU[i]));
            double lb_ = 0.0, ub_ = 0.0;
            if (hl) {
                double old = w->Axl[i];
                w->Axl[i] = old + al;
                if (hasx_l) { sxl = sxl - fabs(old) * mx + fabs(w->Axl[i]) * mx; lb_ = -h_gamma(1) * sxl; }
                else { sxl = fabs(w->Axl[i]) * mx; hasx_l = 1; }
            }
            if (hu) {
                double old = w->Axu[i];
                w->Axu[i] = old + au;
                if (hasx_u) { sxu = sxu - fabs(old) * mx + fabs(w->Axu[i]) * mx; ub_ = h_gamma(1) * sxu; }
                else { sxu = fabs(w->Axu[i]) * mx; hasx_u = 1; }
            }
            lacc = add_enclosed(lacc, lb_, -1.0);
            uacc = add_enclosed(uacc, ub_, 1.0);
            continue;
        }
        if (op == GH_OP_CONST) {
            /* add_constant_node (backward_bound.py:1604-1625) */
            double c = nd->c;
            if (hl) lacc = add_enclosed(lacc, widen(al * c, rounding_error_1(fabs(al) * fabs(c)), -1.0), -1.0);
            if (hu) uacc = add_enclosed(uacc, widen(au * c, rounding_error_1(fabs(au) * fabs(c)), 1.0), 1.0);
            continue;
        }
        int a = nd->a, b = nd->b;
        double lb_ = 0.0, ub_ = 0.0, el = 0.0, eu = 0.0;
        /* the node's own deferred merge (backward_bound.py:801-806) */
        if (w->dfl[v] && hl) { double m = mag_of(w, v); el += isfinite(m) ? h_gamma(1) * fabs(al) * m : (al != 0 ? INFINITY : 0.0); }
        if (w->dfu[v] && hu) { double m = mag_of(w, v); eu += isfinite(m) ? h_gamma(1) * fabs(au) * m : (au != 0 ? INFINITY : 0.0); }
        w->rem[a]--;
        if (op == GH_OP_ADD || op == GH_OP_SUB || op == GH_OP_MUL) w->rem[b]--;
        switch (op) {
        case GH_OP_ADD: case GH_OP_SUB: case GH_OP_ADDC: {
            double sb = (op == GH_OP_SUB) ? -1.0 : 1.0;
            double ma = finite_abs(fmax(fabs(w->L[a]), fabs(w->U[a])));
            double mb = (op == GH_OP_ADDC) ? fabs(nd->c) : finite_abs(fmax(fabs(w->L[b]), fabs(w->U[b])));
            double err = ((0.0 + ma) + mb) * k_round(2);
            if (hl) lb_ = lb_ - finite_abs(al) * err;
            if (hu) ub_ = ub_ + finite_abs(au) * err;
            if (hl) el += push_A(w, a, al, 0, op != GH_OP_ADDC && b == a, 0);
            if (hu) eu += push_A(w, a, au, 1, op != GH_OP_ADDC && b == a, 0);
            if (op != GH_OP_ADDC) {
                if (hl) el += push_A(w, b, sb * al, 0, 0, 0);
                if (hu) eu += push_A(w, b, sb * au, 1, 0, 0);
            } else {
                /* the constant operand, as add_constant_node adds it
                 * (backward_bound.py:1604-1625) */
                double c = nd->c;
                if (hl) lb_ = lb_ + widen(mulz(al, c), rounding_error_1(fabs(al) * fabs(c)), -1.0);
                if (hu) ub_ = ub_ + widen(mulz(au, c), rounding_error_1(fabs(au) * fabs(c)), 1.0);
            }
            break;
        }
        case GH_OP_NEG:
            if (hl) el += push_A(w, a, -al, 0, 0, 0);
            if (hu) eu += push_A(w, a, -au, 1, 0, 0);
            break;
        case GH_OP_MULC: {
            double c = nd->c, ie = inexact_abs(c), m = mag_of(w, a);
            if (hl) { el += (ie == 0) ? 0.0 : (isfinite(m) ? h_gamma(1) * (fabs(al) * (ie * m)) : INFINITY); el += push_A(w, a, al * c, 0, 0, 0); }
            if (hu) { eu += (ie == 0) ? 0.0 : (isfinite(m) ? h_gamma(1) * (fabs(au) * (ie * m)) : INFINITY); eu += push_A(w, a, au * c, 1, 0, 0); }
            break;
        }
        case GH_OP_MUL: {
            const double *Rv = w->R + 6 * (size_t)v;
            int ud = w->udef[v];
            if (hl) {
                if (ud) und |= 1;
                double ax = mulz(pos(al), Rv[0]) + mulz(neg(al), Rv[3]);
                double ay = mulz(pos(al), Rv[1]) + mulz(neg(al), Rv[4]);
                lb_ = mulz(pos(al), Rv[2]) + mulz(neg(al), Rv[5]);
                el += push_A(w, a, ax, 0, b == a, 0);
                el += push_A(w, b, ay, 0, 0, 0);
            }
            if (hu) {
                if (ud) und |= 2;
                double ax = mulz(pos(au), Rv[3]) + mulz(neg(au), Rv[0]);
                double ay = mulz(pos(au), Rv[4]) + mulz(neg(au), Rv[1]);
                ub_ = mulz(pos(au), Rv[5]) + mulz(neg(au), Rv[2]);
                eu += push_A(w, a, ax, 1, b == a, 0);
                eu += push_A(w, b, ay, 1, 0, 0);
            }
            break;
        }
        default: { /* relaxed unary: lw lb uw ub (activation_base.py:137-182) */
            const double *Rv = w->R + 6 * (size_t)v;
            int ud = w->udef[v];
            if (hl) {
                if ((ud & 1) && al > 0) und |= 1;
                if ((ud & 2) && al < 0) und |= 1;
                double ax = mulz(pos(al), Rv[0]) + mulz(neg(al), Rv[2]);
                lb_ = mulz(pos(al), Rv[1]) + mulz(neg(al), Rv[3]);
                el += push_A(w, a, ax, 0, 0, 0);
            }
            if (hu) {
                if ((ud & 2) && au > 0) und |= 2;
                if ((ud & 1) && au < 0) und |= 2;
                double ax = mulz(pos(au), Rv[2]) + mulz(neg(au), Rv[0]);
                ub_ = mulz(pos(au), Rv[3]) + mulz(neg(au), Rv[1]);
                eu += push_A(w, a, ax, 1, 0, 0);
            }
            break;
        }
        }
        lacc = add_enclosed(lacc, lb_ - el, -1.0);
        uacc = add_enclosed(uacc, ub_ + eu, 1.0);
    }
    /* concretization over the box (concretize_bounds.py:267-321) */
    double bl = widen(lacc, finite_abs(lacc) * k_round(n), -1.0);
    double bu = widen(uacc, finite_abs(uacc) * k_round(n), 1.0);
    double base_l = 0.0, et_l = 0.0, ma_l = 0.0, base_u = 0.0, et_u = 0.0, ma_u = 0.0;
    for (int i = 0; i < n; ++i) {
        double x0 = (xU[i] + xL[i]) / 2.0, ep = (xU[i] - xL[i]) / 2.0;
        double mx = finite_abs(fmax(-xL[i], xU[i]));
        base_l = fma(w->Axl[i], x0, base_l); et_l = fma(fabs(w->Axl[i]), ep, et_l);
        ma_l = fma(finite_abs(w->Axl[i]), mx, ma_l);
        base_u = fma(w->Axu[i], x0, base_u); et_u = fma(fabs(w->Axu[i]), ep, et_u);
        ma_u = fma(finite_abs(w->Axu[i]), mx, ma_u);
    }
    double rl = widen(base_l - et_l, ma_l * k_round(n), -1.0);
    if (!isfinite(rl)) rl = -INFINITY;
    double ru = widen(base_u + et_u, ma_u * k_round(n), 1.0);
    if (!isfinite(ru)) ru = INFINITY;
    double lo = add_enclosed(bl, rl, -1.0), hi = add_enclosed(bu, ru, 1.0);
    if (isnan(lo) || (und & 1)) lo = -INFINITY;
    if (isnan(hi) || (und & 2)) hi = INFINITY;
    *out_l = lo;
    *out_u = want_u ? hi : INFINITY;
    /* the lower plane A x + bias over the inputs, before concretization
     * (backward_bound.py:1093-1111); a discarded row's plane is vacuous:
     * A = 0, bias -inf (_zero_undefined_root_A, :445-463) */
    if (plane_A) {
        for (int i = 0; i < n; ++i) plane_A[i] = (und & 1) ? 0.0 : w->Axl[i];
        *plane_b = (und & 1) ? -INFINITY : lacc;
        *undef = und & 1;
    }
}

static void bound_box(const gh_prog *P, dag_ws *w, const double *xL, const double *xU,
                      const uint8_t *needs, double *lb, double *lA, double *lbias,
                      uint8_t *undef)
{
    const gh_node *N = P->nodes;
    int n = P->n, V = P->V;
    for (int i = 0; i < n; ++i) { w->L[i] = xL[i]; w->U[i] = xU[i]; }
    for (int v = n; v < V; ++v) {
        double il, iu;
        ibp_node(&N[v], w->L, w->U, &il, &iu);
        w->L[v] = il; w->U[v] = iu;
        /* the node's own planes come from its inputs' final bounds */
        w->udef[v] = 0;
        if (is_relaxed(N[v].op))
            w->udef[v] = relax_node(&N[v], w->L, w->U, w->R + 6 * (size_t)v);
        if (needs[v] && N[v].op != GH_OP_CONST) {
            double cl, cu;
            crown_pass(P, w, xL, xU, v, 1, &cl, &cu, NULL, NULL, NULL);
            /* intersect with the interval (bound_general.py:1160-1185) */
            w->L[v] = fmax(cl, il);
            w->U[v] = fmin(cu, iu);
        }
    }
    for (int r = 0; r < P->R; ++r) {
        int o = P->out[r];
        double cl, cu;
        double *pA = lA ? lA + (size_t)r * n : NULL;
        if (o < n) {
            /* a bare input: its plane is the unit row */
            lb[r] = xL[o];
            if (pA) {
                for (int i = 0; i < n; ++i) pA[i] = (i == o) ? 1.0 : 0.0;
                lbias[r] = 0.0;
                undef[r] = 0;
            }
            continue;
        }
        int und = 0;
        double b = 0.0;
        crown_pass(P, w, xL, xU, o, 0, &cl, &cu, pA, &b, &und);
        if (pA) { lbias[r] = b; undef[r] = (uint8_t)und; }
        /* undefined rows fall back to the interval (backward_bound.py:826-877) */
        lb[r] = (cl == -INFINITY) ? w->L[o] : fmax(cl, -INFINITY);
    }
}

void gho_dag_needs(const gh_prog *P, uint8_t *needs)
{
    memset(needs, 0, (size_t)P->V);
    for (int v = P->n; v < P->V; ++v) {
        const gh_node *nd = &P->nodes[v];
        if (!is_relaxed(nd->op)) continue;
        if (nd->a >= P->n) needs[nd->a] = 1;
        if (nd->op == GH_OP_MUL && nd->b >= P->n) needs[nd->b] = 1;
    }
}

int gho_dag_crown(const gh_prog *P, int B, const double *xL, const double *xU, double *lb,
                  int threads)
{
    return gho_dag_crown_planes(P, B, xL, xU, lb, NULL, NULL, NULL, threads);
}

/* As gho_dag_crown, with each row's lower plane: lA (B, R, n), lbias (B, R)
 * and the rows whose relaxation was undefined, undef (B, R). */
int gho_dag_crown_planes(const gh_prog *P, int B, const double *xL, const double *xU,
                         double *lb, double *lA, double *lbias, uint8_t *undef, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int V = P->V, n = P->n;
    uint8_t *needs = (uint8_t *)malloc((size_t)V);
    gho_dag_needs(P, needs);
    /* the tangent tables are built once, before the threads read them */
    int pow_mask = 0;
    for (int v = n; v < V; ++v) {
        const gh_node *nd = &P->nodes[v];
        if (nd->op == GH_OP_POW && nd->c == floor(nd->c) && ((int)nd->c) % 2 == 1 && nd->c < 32)
            pow_mask |= 1 << (int)nd->c;
    }
    gho_sshape_init(pow_mask);
#pragma omp parallel
    {
        dag_ws w;
        double *buf = (double *)calloc((size_t)V * 10 + 2 * (size_t)n, sizeof(double));
        int *ib = (int *)calloc((size_t)V * 8, sizeof(int));
        w.L = buf; w.U = buf + V; w.R = buf + 2 * V; w.Al = buf + 8 * V; w.Au = buf + 9 * V;
        w.Axl = buf + 10 * V; w.Axu = buf + 10 * V + n;
        w.udef = ib; w.hasl = ib + V; w.hasu = ib + 2 * V; w.rem = ib + 3 * V; w.anc = ib + 4 * V;
        w.dfl = ib + 5 * V; w.dfu = ib + 6 * V; w.hasx = ib + 7 * V;
#pragma omp for schedule(dynamic, 16)
        for (int b = 0; b < B; ++b)
            bound_box(P, &w, xL + (size_t)b * n, xU + (size_t)b * n, needs, lb + (size_t)b * P->R,
                      lA ? lA + (size_t)b * P->R * n : NULL, lbias ? lbias + (size_t)b * P->R : NULL,
                      undef ? undef + (size_t)b * P->R : NULL);
        free(buf);
        free(ib);
    }
    free(needs);
    return 0;
}

/* Interval (IBP) value of every row, for the fallback tests. */
int gho_dag_ibp(const gh_prog *P, int B, const double *xL, const double *xU, double *lo, double *hi)
{
    int V = P->V, n = P->n;
    double *L = (double *)malloc(sizeof(double) * V), *U = (double *)malloc(sizeof(double) * V);
    for (int b = 0; b < B; ++b) {
        for (int i = 0; i < n; ++i) { L[i] = xL[(size_t)b * n + i]; U[i] = xU[(size_t)b * n + i]; }
        for (int v = n; v < V; ++v) ibp_node(&P->nodes[v], L, U, &L[v], &U[v]);
        for (int r = 0; r < P->R; ++r) { lo[(size_t)b * P->R + r] = L[P->out[r]]; hi[(size_t)b * P->R + r] = U[P->out[r]]; }
    }
    free(L); free(U);
    return 0;
}

/* Evaluate the rows at points X (N, n) -> F (N, R). */
int gho_dag_eval(const gh_prog *P, int N_, const double *X, double *F)
{
    int V = P->V, n = P->n;
    double *Y = (double *)malloc(sizeof(double) * V);
    for (int k = 0; k < N_; ++k) {
        for (int i = 0; i < n; ++i) Y[i] = X[(size_t)k * n + i];
        for (int v = n; v < V; ++v) {
            const gh_node *nd = &P->nodes[v];
            double a = Y[nd->a], b = (nd->op == GH_OP_ADD || nd->op == GH_OP_SUB || nd->op == GH_OP_MUL) ? Y[nd->b] : 0, c = nd->c, y = 0;
            switch (nd->op) {
            case GH_OP_CONST: y = c; break;
            case GH_OP_ADD: y = a + b; break;
            case GH_OP_SUB: y = a - b; break;
            case GH_OP_NEG: y = -a; break;
            case GH_OP_MULC: y = a * c; break;
            case GH_OP_ADDC: y = a + c; break;
            case GH_OP_MUL: y = a * b; break;
            case GH_OP_SQR: y = a * a; break;
            case GH_OP_POW: y = pow(a, c); break;
            case GH_OP_EXP: y = exp(a); break;
            case GH_OP_LOG: y = log(a); break;
            case GH_OP_RECIP: y = 1.0 / a; break;
            case GH_OP_SIN: y = sin(a); break;
            case GH_OP_COS: y = cos(a); break;
            default: y = NAN;
            }
            Y[v] = y;
        }
        for (int r = 0; r < P->R; ++r) F[(size_t)k * P->R + r] = Y[P->out[r]];
    }
    free(Y);
    return 0;
}

/* The shared elementary functions (gearhip_math.h) over an array, for their
 * accuracy tests: fn 0 exp, 1 expm1, 2 log, 3 sin, 4 cos, 5 asin, 10 + p x^p. */
int gho_math(int fn, int N, const double *x, double *y)
{
    for (int k = 0; k < N; ++k) {
        double v = x[k], r;
        switch (fn) {
        case 0: r = gh_exp(v); break;
        case 1: r = gh_expm1(v); break;
        case 2: r = gh_log(v); break;
        case 3: r = gh_sin(v); break;
        case 4: r = gh_cos(v); break;
        case 5: r = gh_asin(v); break;
        default:
            if (fn < 10) return 1;
            r = gh_powi(v, fn - 10);
        }
        y[k] = r;
    }
    return 0;
}
