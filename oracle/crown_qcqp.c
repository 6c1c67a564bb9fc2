/*
 * CPU oracle (test infrastructure only) -- plain C restatement of the
 * reference's FP64 CROWN lower bound for a quadratically constrained
 * quadratic program, and of the branch-and-bound split / branch / prune
 * steps that drive it.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may
 * load this library; it is the checker, never the product path.
 *
 * Problem.  Rows s = 0..S-1 (row 0 the objective, rows 1.. the
 * constraints) are quadratics  f_s(x) = x^T Q_s x + c_s^T x + d_s  over a
 * box [xL, xU] in R^n.
 *
 * Graph bounded (the lowering of one QCQP into the reference's operator
 * vocabulary):
 *     y = Linear(x; W)          W = [Q_0; ...; Q_{S-1}]  ((S n) x n, no bias)
 *     u = Unsqueeze(x)          (B, 1, n)
 *     v = Reshape(y)            (B, S, n)
 *     z = Mul(u, v)             both perturbed: coeff-aware McCormick
 *     q = ReduceSum(z, -1)      (B, S)
 *     l = Linear(x; Cm, d)      (B, S)
 *     f = Add(q, l)             specification C = eye(S), lower bounds only
 *
 * Every arithmetic step below follows the reference file:line cited next to
 * it (paths relative to solver/auto_LiRPA).  Where the reference leaves the
 * evaluation order to BLAS / einsum, this restatement fixes it:
 *   - a matrix-vector product (the IBP centre / radius / magnitude of y and
 *     the backward product beta^T Q_s) is, per output, a sequential fused
 *     multiply-add chain over the contraction index in increasing order,
 *     starting from +0.0;
 *   - a reduction to one scalar per (box, row) is "red16": 16 lane-strided
 *     partials (partial l takes indices l, l+16, l+32, ... in order, from
 *     +0.0), then an xor tree  p[l] += p[l^8], p[l^4], p[l^2], p[l^1];
 *   - every other operation is one IEEE double operation, no contraction
 *     (built with -ffp-contract=off).
 * The HIP kernels evaluate the same chains in the same order (one 16-lane
 * group per box), which is what makes their FP64 results bit-identical to
 * this oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS64 2.220446049250313080847e-16

/* rounding_error's factor (n + 16) * eps * 4 -- fp_enclosure.py:34-48 */
static double k_round(long n) { return (double)(n + 16) * EPS64 * 4.0; }

/* _higham's factor (k + 1) * eps / 2 * (1 + 1e-6) -- backward_bound.py:550-555 */
static double h_gamma(long k) { return (double)(k + 1) * EPS64 / 2.0 * (1.0 + 1e-6); }

/* finite_abs -- fp_enclosure.py:8-10 */
static double finite_abs(double v) { return isfinite(v) ? fabs(v) : 0.0; }

/* widen_bound -- fp_enclosure.py:51-57 */
static double widen(double bound, double err, double sign)
{
    if (isinf(err) && err > 0) return sign * INFINITY;
    return isfinite(bound) ? bound + sign * err : bound;
}

/* add_bound -- fp_enclosure.py:94-98 */
static double add_enclosed(double a, double b, double sign)
{
    double r = a + b;
    return widen(r, finite_abs(r) * k_round(1), sign);
}

/* red16 sum of v[0..n) */
static double red16_sum(const double *v, int n)
{
    double p[16];
    for (int l = 0; l < 16; ++l) {
        double a = 0.0;
        for (int i = l; i < n; i += 16) a = a + v[i];
        p[l] = a;
    }
    for (int off = 8; off >= 1; off >>= 1) {
        double q[16];
        for (int l = 0; l < 16; ++l) q[l] = p[l] + p[l ^ off];
        memcpy(p, q, sizeof q);
    }
    return p[0];
}

/* red16 dot: partial l = fma chain over i = l, l+16, ...; a is taken
 * through fabs() when abs_a is set (the |A| products of the reference). */
static double red16_dot(const double *a, const double *b, int n, int abs_a)
{
    double p[16];
    for (int l = 0; l < 16; ++l) {
        double acc = 0.0;
        for (int i = l; i < n; i += 16) acc = fma(abs_a ? fabs(a[i]) : a[i], b[i], acc);
        p[l] = acc;
    }
    for (int off = 8; off >= 1; off >>= 1) {
        double q[16];
        for (int l = 0; l < 16; ++l) q[l] = p[l] + p[l ^ off];
        memcpy(p, q, sizeof q);
    }
    return p[0];
}

/* McCormick lower plane chosen coeff-aware (bivariate.py:190-261) and
 * materialised by interpolated_relaxation with r in {0, 1}
 * (bivariate.py:40-75).  Returns alpha (coefficient of x), beta
 * (coefficient of y), gamma (constant). */
static void mccormick_lower(double xl, double xu, double yl, double yu,
                            double *alpha, double *beta, double *gamma)
{
    /* _affine_min_box (bivariate.py:190-196): a*x_min + b*y_min + c */
    double a1 = yl, b1 = xl, c1 = (-xl) * yl;
    double a2 = yu, b2 = xu, c2 = (-xu) * yu;
    double m1 = a1 * (a1 >= 0 ? xl : xu);
    m1 = m1 + b1 * (b1 >= 0 ? yl : yu);
    m1 = m1 + c1;
    double m2 = a2 * (a2 >= 0 ? xl : xu);
    m2 = m2 + b2 * (b2 >= 0 ? yl : yu);
    m2 = m2 + c2;
    double r = (m1 >= m2) ? 1.0 : 0.0;
    double t;
    t = (yl - yu) * r; *alpha = t + yu;
    t = (xl - xu) * r; *beta = t + xu;
    double p = yu * xu, q = yl * xl;
    t = (p - q) * r; *gamma = t - p;
}

/* Per-domain work area. */
typedef struct {
    double *x0, *ep, *mx, *yl, *yu, *mg, *al, *be, *A, *Acm, *t0, *t1;
} work_t;

/*
 * One domain, all S rows.  Layout: Q (S, n, n) with y_{s,j} = sum_i
 * Q[s][j][i] x_i; Cm (S, n); d (S); xL, xU (n); lb (S); lA (S, n) or NULL;
 * lbias (S) or NULL.
 */
static void crown_one(int n, int S, const double *Q, const double *Cm,
                      const double *d, const double *xL, const double *xU,
                      double *lb, double *lA, double *lbias, work_t *w)
{
    const double Kn = k_round(n), K2 = k_round(2);
    const double H1 = h_gamma(1), HS = h_gamma(S), HSn = h_gamma((long)S * n);
    /* box centre / radius as constraints_solving and _propagate_Linf form
     * them (concretize_func.py:548-551, linear.py:607-618); magnitude
     * max(-l, u) (backward_bound.py:739-741, fp_enclosure.py:64). */
    for (int i = 0; i < n; ++i) {
        w->x0[i] = (xU[i] + xL[i]) / 2.0;
        w->ep[i] = (xU[i] - xL[i]) / 2.0;
        w->mx[i] = finite_abs(fmax(-xL[i], xU[i]));
    }
    for (int s = 0; s < S; ++s) {
        const double *Qs = Q + (size_t)s * n * n;
        const double *cs = Cm + (size_t)s * n;
        /* ---- intermediate bounds of y_s: first-linear IBP with the
         * rounding margin (linear.py:607-653, interval_bound.py:357-390) */
        for (int j = 0; j < n; ++j) {
            const double *row = Qs + (size_t)j * n;
            double c = 0.0, dv = 0.0, mg = 0.0;
            for (int i = 0; i < n; ++i) {
                double qa = fabs(row[i]);
                c = fma(row[i], w->x0[i], c);
                dv = fma(qa, w->ep[i], dv);
                mg = fma(qa, w->mx[i], mg);
            }
            double e = mg * Kn;
            w->yl[j] = widen((c - dv) + 0.0, e, -1.0);
            w->yu[j] = widen((c + dv) + 0.0, e, 1.0);
            w->mg[j] = mg;
        }
        /* ---- Add/Sub operand rounding at f (backward_bound.py:1057-1070,
         * interval_bound.py:74-78) on the intervals of q_s and l_s.
         * Interval of z_j = x_j * y_j: four corners (bivariate.py:563-597);
         * q_s sums them (ReduceSum). */
        for (int j = 0; j < n; ++j) {
            double r0 = xL[j] * w->yl[j], r1 = xL[j] * w->yu[j];
            double r2 = xU[j] * w->yl[j], r3 = xU[j] * w->yu[j];
            w->t0[j] = fmin(fmin(r0, r1), fmin(r2, r3));
            w->t1[j] = fmax(fmax(r0, r1), fmax(r2, r3));
        }
        double ql = red16_sum(w->t0, n), qu = red16_sum(w->t1, n);
        /* interval of l_s = Linear(x; Cm, d): _bound_box (linear.py:621-653) */
        double lc = red16_dot(cs, w->x0, n, 0);
        double ldv = red16_dot(cs, w->ep, n, 1);
        double lmg = red16_dot(cs, w->mx, n, 1);
        double le = (lmg + fabs(d[s])) * Kn;
        double ll = widen((lc - ldv) + d[s], le, -1.0);
        double lu = widen((lc + ldv) + d[s], le, 1.0);
        double mag_add = (0.0 + finite_abs(fmax(fabs(ql), fabs(qu))))
                         + finite_abs(fmax(fabs(ll), fabs(lu)));
        double add_err = mag_add * K2;

        /* ---- backward pass, nodes in the reference queue order
         * f, q, l, z, u, v, y (backward_bound.py:989-1125); after each node
         * lb = add_bound(lb, node bias) (:1091). */
        double acc = 0.0;
        /* f: Add; bias 0 minus the operand-rounding charge */
        acc = add_enclosed(acc, 0.0 - add_err, -1.0);
        /* q: ReduceSum (exact expansion, bias 0 -- reduce.py:36-37) */
        acc = add_enclosed(acc, 0.0, -1.0);
        /* l: Linear with bias d; A67 terms gamma_S <|A|, |Cm| m> and
         * gamma_S |d| (backward_bound.py:611-648, 749-823) */
        double vec = (0.0 + HS * lmg) + HS * fabs(d[s]);
        acc = add_enclosed(acc, d[s] - vec, -1.0);
        /* z: Mul, McCormick planes on the y_s block (bivariate.py:401-488,
         * clampmult.py:161-216 with last_A = +1 on the block) */
        for (int j = 0; j < n; ++j)
            mccormick_lower(xL[j], xU[j], w->yl[j], w->yu[j], &w->al[j], &w->be[j], &w->t0[j]);
        double gsum = red16_sum(w->t0, n);
        acc = add_enclosed(acc, gsum, -1.0);
        /* u: Unsqueeze; merge of alpha into x's A (the first A there came
         * from l): gamma_1 <|Cm_s + alpha|, m> (backward_bound.py:736-747,
         * 801-821) */
        for (int i = 0; i < n; ++i) w->Acm[i] = cs[i] + w->al[i];
        double merge1 = red16_dot(w->Acm, w->mx, n, 1);
        acc = add_enclosed(acc, 0.0 - H1 * merge1, -1.0);
        /* v: Reshape (exact) */
        acc = add_enclosed(acc, 0.0, -1.0);
        /* y: Linear without bias: A_x += beta^T Q_s (linear.py:268-270);
         * A67 term gamma_{S n} <|beta|, |Q_s| m> plus the merge charge */
        for (int i = 0; i < n; ++i) {
            double a = 0.0;
            for (int j = 0; j < n; ++j) a = fma(w->be[j], Qs[(size_t)j * n + i], a);
            w->A[i] = w->Acm[i] + a;
        }
        for (int j = 0; j < n; ++j) w->t1[j] = HSn * w->mg[j];
        double own = red16_dot(w->be, w->t1, n, 1);
        double merge2 = red16_dot(w->A, w->mx, n, 1);
        double err_y = own + H1 * merge2;
        acc = add_enclosed(acc, 0.0 - err_y, -1.0);

        /* ---- concretization (concretize_bounds.py:267-321,
         * concretize_func.py:552-569, fp_enclosure.py:60-82) */
        if (lbias) lbias[s] = acc;
        double bias = widen(acc, finite_abs(acc) * Kn, -1.0);
        double base = red16_dot(w->A, w->x0, n, 0);
        double et = red16_dot(w->A, w->ep, n, 1);
        double ma = red16_dot(w->A, w->mx, n, 1);
        double root = widen(base - et, ma * Kn, -1.0);
        if (!isfinite(root)) root = -INFINITY;
        lb[s] = add_enclosed(bias, root, -1.0);
        if (lA) memcpy(lA + (size_t)s * n, w->A, sizeof(double) * n);
    }
}

/* Lower bounds for B boxes: lb (B, S), optional lA (B, S, n), lbias (B, S).
 * threads <= 0 uses the OpenMP default.  Returns 0. */
int gho_qcqp_crown(int B, int n, int S, const double *Q, const double *Cm,
                   const double *d, const double *xL, const double *xU,
                   double *lb, double *lA, double *lbias, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        work_t w;
        double *buf = (double *)malloc(sizeof(double) * (size_t)n * 12);
        w.x0 = buf; w.ep = buf + n; w.mx = buf + 2 * n; w.yl = buf + 3 * n;
        w.yu = buf + 4 * n; w.mg = buf + 5 * n; w.al = buf + 6 * n;
        w.be = buf + 7 * n; w.A = buf + 8 * n; w.Acm = buf + 9 * n;
        w.t0 = buf + 10 * n; w.t1 = buf + 11 * n;
#pragma omp for schedule(static)
        for (int b = 0; b < B; ++b)
            crown_one(n, S, Q, Cm, d, xL + (size_t)b * n, xU + (size_t)b * n,
                      lb + (size_t)b * S, lA ? lA + (size_t)b * S * n : NULL,
                      lbias ? lbias + (size_t)b * S : NULL, &w);
        free(buf);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Branch on the widest input (branching_heuristics.py:83-86, 'naive':
 * topk(x_U - x_L, 1)); ties go to the lowest index. */
void gho_branch_widest(int B, int n, const double *xL, const double *xU,
                       int32_t *idx)
{
    for (int b = 0; b < B; ++b) {
        const double *l = xL + (size_t)b * n, *u = xU + (size_t)b * n;
        int best = 0;
        double bw = u[0] - l[0];
        for (int i = 1; i < n; ++i) {
            double wdt = u[i] - l[i];
            if (wdt > bw) { bw = wdt; best = i; }
        }
        idx[b] = best;
    }
}

/* Smart branching, one split per box (branching_heuristics.py:25-56,
 * 59-101, 104-125, 176-200) with the GEAR presets' options (sb_sum, no
 * touch-zero score, the exhaustive guard; gear-f0.yaml:518-541):
 *   - an objective threshold of +inf (no incumbent yet) counts as the box's
 *     own bound (:68-73);
 *   - if any plane coefficient, bound or threshold of the batch is not
 *     finite, every box takes the widest input (:74-87);
 *   - else score_i = sum_s max(|lA[s][i]|, coeff_thresh) (u_i - l_i) / 2,
 *     summed over the rows in order, the largest score wins (lowest index
 *     on ties: the reference's topk order is unspecified);
 *   - the exhaustive guard (:176-200): with root widths given and
 *     min_width_ratio > 0, a choice whose width relative to the root is
 *     below min_width_ratio times the largest relative width becomes that
 *     widest (relative) input.
 * lb (B, S), thr (S), lA (B, S, n), root_w (n) or NULL. */
void gho_branch_sb(int B, int n, int S, const double *xL, const double *xU, const double *lb,
                   const double *thr, const double *lA, double coeff_thresh,
                   const double *root_w, double min_width_ratio, int32_t *idx)
{
    int finite = 1;
    for (size_t k = 0; k < (size_t)B * S * n && finite; ++k) finite = isfinite(lA[k]);
    for (int b = 0; b < B && finite; ++b)
        for (int s = 0; s < S; ++s) {
            double l = lb[(size_t)b * S + s];
            double t = (isinf(thr[s]) && thr[s] > 0) ? l : thr[s];
            if (!isfinite(l) || !isfinite(t)) { finite = 0; break; }
        }
    if (!finite) {
        gho_branch_widest(B, n, xL, xU, idx);
        return;
    }
    for (int b = 0; b < B; ++b) {
        const double *l = xL + (size_t)b * n, *u = xU + (size_t)b * n;
        const double *a = lA + (size_t)b * S * n;
        int best = 0;
        double bs = -INFINITY;
        for (int i = 0; i < n; ++i) {
            double w = u[i] - l[i], sc = 0.0;
            for (int s = 0; s < S; ++s) {
                double c = fabs(a[(size_t)s * n + i]);
                sc = sc + (fmax(c, coeff_thresh) * w) / 2;
            }
            if (sc > bs) { bs = sc; best = i; }
        }
        if (root_w && min_width_ratio > 0) {
            double rmax = -INFINITY, rc = 0.0;
            int widest = 0;
            for (int i = 0; i < n; ++i) {
                double r = root_w[i] > 0 ? (u[i] - l[i]) / fmax(root_w[i], 1e-300) : 0.0;
                if (r > rmax) { rmax = r; widest = i; }
                if (i == best) rc = r;
            }
            if (rc < min_width_ratio * rmax) best = widest;
        }
        idx[b] = best;
    }
}

/* Bisection (split.py:26-122 with split_partitions = 2, depth 1): child 0
 * is [mid, xU] and child 1 is [xL, mid] on the split dimension, with
 * mid = (1*xL + 1*xU) / 2; children are laid out [all child 0; all child 1]
 * (out arrays are (2B, n)). */
void gho_split(int B, int n, const double *xL, const double *xU,
               const int32_t *idx, double *oL, double *oU)
{
    for (int c = 0; c < 2; ++c)
        for (int b = 0; b < B; ++b) {
            const double *l = xL + (size_t)b * n, *u = xU + (size_t)b * n;
            double *ol = oL + ((size_t)c * B + b) * n;
            double *ou = oU + ((size_t)c * B + b) * n;
            memcpy(ol, l, sizeof(double) * n);
            memcpy(ou, u, sizeof(double) * n);
            int k = idx[b];
            double lo = (c + 1) * l[k] + (2 - c - 1) * u[k];
            double hi = c * l[k] + (2 - c) * u[k];
            ol[k] = lo / 2.0;
            ou[k] = hi / 2.0;
        }
}

/* Prune (branching_domains.py:447-517, solving mode): a domain is dropped
 * when some row has finite lb and threshold with lb > thr, when the
 * objective row has lb >= thr, or when the objective lb is +inf with a
 * finite threshold; boxes with xL > xU anywhere (all finite) are dropped
 * too (:534-538).  keep[b] = 1 for survivors; returns their count.
 * *disc_min receives min(thr[b][0]) over dropped domains with a finite
 * objective threshold (or is left unchanged). */
long gho_prune(int B, int S, int n, const double *lb, const double *thr,
               const double *xL, const double *xU, uint8_t *keep,
               double *disc_min)
{
    long kept = 0;
    for (int b = 0; b < B; ++b) {
        const double *L = lb + (size_t)b * S, *T = thr + (size_t)b * S;
        int ver = 0;
        for (int s = 0; s < S; ++s) {
            int gt = isfinite(L[s]) && isfinite(T[s]) && (L[s] > T[s]);
            if (s == 0) gt |= isfinite(L[0]) && isfinite(T[0]) && (L[0] >= T[0]);
            ver |= gt;
        }
        ver |= (isinf(L[0]) && L[0] > 0 && isfinite(T[0]));
        if (ver && isfinite(T[0]) && T[0] < *disc_min) *disc_min = T[0];
        int nonfinite = 0, invalid = 0;
        for (int i = 0; i < n; ++i) {
            double l = xL[(size_t)b * n + i], u = xU[(size_t)b * n + i];
            if (!isfinite(l) || !isfinite(u)) nonfinite = 1;
            if (l > u) invalid = 1;
        }
        int k = !ver && !(invalid && !nonfinite);
        keep[b] = (uint8_t)k;
        kept += k;
    }
    return kept;
}
