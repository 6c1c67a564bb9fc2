/*
 * CPU oracle (test infrastructure only) -- plain C restatement of the
 * reference's constrained concretization: the lower bound of
 * min c_h^T x over the box [xL, xU] intersected with linear cuts
 * A_k x + b_k <= 0, by the sequential one-row dual sweep.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may
 * load this library; it is the checker, never the product path.
 *
 * Reference (paths relative to solver/auto_LiRPA):
 *   _finite_constraints            concretize_func.py:78-81
 *   _sort_out_constraints          concretize_func.py:84-114
 *   sort_out_constr_batches        concretize_func.py:312-400 (centred rows
 *                                  d = A x0 + b, widened down :393-396)
 *   drop_redundant_constraints     concretize_func.py:269-310
 *   _solve_dual_var                concretize_func.py:143-227
 *   constraints_solving            concretize_func.py:431-850 (naive bound
 *                                  :594-604, sweep :788-807, final
 *                                  enclosure :810-832)
 *   rounding_error, widen_bound, box_rounding_error, finite_abs
 *                                  fp_enclosure.py:8-82
 *
 * Evaluation order (the reference leaves it to einsum / sort / cumsum; it
 * is fixed here and the HIP kernel follows it):
 *   - every sum over the X <= 64 input coordinates is "red64": term i in
 *     slot i (0 in slots i >= X), then p[l] = p[l] + p[l ^ o] for
 *     o = 32, 16, 8, 4, 2, 1;
 *   - breakpoints are sorted ascending on (q, index), NaN after every
 *     number; the cumulative weights are a Hillis-Steele inclusive scan
 *     over the 64 sorted slots: s[j] = s[j] + s[j - o] for o = 1, 2, ..., 32;
 *   - sums over the cut rows run in row order;
 *   - no contraction (built with -ffp-contract=off).
 * One deliberate difference: the final error term counts the rows a box
 * keeps (K_b) where the reference counts the largest kept-row count of its
 * chunk; both enclose the same computation.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EPS64 2.220446049250313080847e-16
#define XMAX 64

static double k_round(long n) { return (double)(n + 16) * EPS64 * 4.0; }
static double finite_abs(double v) { return isfinite(v) ? fabs(v) : 0.0; }
static double widen(double bound, double err, double sign)
{
    if (isinf(err) && err > 0) return sign * INFINITY;
    return isfinite(bound) ? bound + sign * err : bound;
}

/* red64 of p[0..64) (destroys p) */
static double red64(double *p)
{
    for (int o = 32; o >= 1; o >>= 1) {
        double q[XMAX];
        for (int l = 0; l < XMAX; ++l) q[l] = p[l] + p[l ^ o];
        memcpy(p, q, sizeof q);
    }
    return p[0];
}

/* (q, i) < (r, j): ascending, NaN last, ties by index */
static int key_less(double q, int i, double r, int j)
{
    int qn = isnan(q), rn = isnan(r);
    if (qn != rn) return rn;          /* a number before NaN */
    if (!qn && q != r) return q < r;
    return i < j;
}

/* Per box: the cut rows the sweep visits, in order, with their centred
 * offsets d (widened down).  Returns the count; *infeasible / *covered as
 * _sort_out_constraints. */
static int prepare_rows(int K, int X, const double *A, const double *b, const double *x0,
                        const double *ep, int max_rows, int *keep, double *dk,
                        int *infeasible, int *covered)
{
    double p[XMAX];
    int inf_any = 0, cov_all = 1;
    double *score = malloc(sizeof(double) * (K > 0 ? K : 1));
    int *cand = malloc(sizeof(int) * (K > 0 ? K : 1));
    int nc = 0;
    for (int k = 0; k < K; ++k) {
        const double *a = A + (size_t)k * X;
        double bk = b[k];
        int valid = isfinite(bk);
        for (int i = 0; i < X; ++i) valid = valid && isfinite(a[i]);
        double row[XMAX];
        for (int i = 0; i < X; ++i) row[i] = valid ? a[i] : 0.0;
        if (!valid) bk = -1.0;
        /* sort-out test on x0 +- eps (concretize_func.py:97-112) */
        for (int l = 0; l < XMAX; ++l) p[l] = l < X ? row[l] * x0[l] : 0.0;
        double xt = red64(p) + bk;
        for (int l = 0; l < XMAX; ++l) p[l] = l < X ? fabs(row[l]) * ep[l] : 0.0;
        double et = red64(p);
        for (int l = 0; l < XMAX; ++l) {
            if (l < X) {
                double lo = x0[l] - ep[l], hi = x0[l] + ep[l];
                p[l] = finite_abs(row[l]) * finite_abs(fmax(fabs(lo), fabs(hi)));
            } else {
                p[l] = 0.0;
            }
        }
        double err = (red64(p) + fabs(bk)) * k_round(X);
        double mn = widen(xt - et, err, -1.0), mx = widen(xt + et, err, 1.0);
        if (mn > 0) inf_any = 1;
        if (!(mx <= 0)) cov_all = 0;
        /* centred offset widened down (concretize_func.py:393-396) */
        for (int l = 0; l < XMAX; ++l) p[l] = l < X ? finite_abs(row[l]) * finite_abs(fabs(x0[l])) : 0.0;
        double derr = (red64(p) + fabs(bk)) * k_round(X);
        double d = widen(xt, derr, -1.0);
        /* _finite_constraints on (A, d) */
        if (!valid || !isfinite(d)) {
            d = -1.0;
            for (int i = 0; i < X; ++i) row[i] = 0.0;
            et = 0.0;
        }
        /* redundancy (concretize_func.py:293-296): d + |A| eps <= 0 */
        double mv = d + et;
        if (!(mv <= 0)) {
            cand[nc] = k;
            score[nc] = isnan(mv) ? INFINITY : mv;
            dk[nc] = d;
            ++nc;
        }
    }
    *infeasible = inf_any;
    *covered = K == 0 ? 1 : cov_all;
    int nk = nc;
    if (max_rows > 0 && nc > max_rows) {
        /* the max_rows largest scores (ties: lower row first), kept in row
         * order (concretize_func.py:298-303) */
        char *take = calloc(nc, 1);
        for (int t = 0; t < max_rows; ++t) {
            int best = -1;
            for (int c = 0; c < nc; ++c)
                if (!take[c] && (best < 0 || score[c] > score[best])) best = c;
            take[best] = 1;
        }
        int w = 0;
        for (int c = 0; c < nc; ++c)
            if (take[c]) { cand[w] = cand[c]; dk[w] = dk[c]; ++w; }
        nk = w;
        free(take);
    }
    for (int c = 0; c < nk; ++c) keep[c] = cand[c];
    free(score);
    free(cand);
    return nk;
}

/* _solve_dual_var for one objective vector c and one centred row (a, d) */
static double solve_beta(int X, const double *c, const double *a, double d, const double *ep)
{
    double q[XMAX], w[XMAX];
    int id[XMAX];
    for (int l = 0; l < XMAX; ++l) {
        if (l < X) {
            q[l] = -(c[l] / a[l]);
            w[l] = fabs(a[l] * ep[l]);
        } else {
            q[l] = INFINITY;
            w[l] = 0.0;
        }
        id[l] = l;
    }
    /* insertion sort on (q, index) */
    for (int i = 1; i < XMAX; ++i) {
        double kq = q[i], kw = w[i];
        int ki = id[i], j = i - 1;
        while (j >= 0 && key_less(kq, ki, q[j], id[j])) {
            q[j + 1] = q[j]; w[j + 1] = w[j]; id[j + 1] = id[j];
            --j;
        }
        q[j + 1] = kq; w[j + 1] = kw; id[j + 1] = ki;
    }
    double s[XMAX];
    memcpy(s, w, sizeof s);
    for (int o = 1; o < XMAX; o <<= 1) {
        double t[XMAX];
        for (int j = 0; j < XMAX; ++j) t[j] = j >= o ? s[j] + s[j - o] : s[j];
        memcpy(s, t, sizeof t);
    }
    double T = s[X - 1];
    int pick = X + 1;
    for (int j = 0; j <= X; ++j) {
        double cum = j == 0 ? 0.0 : s[j - 1];
        double g = (cum - T) + cum - d;
        if (g >= 0) { pick = j; break; }
    }
    double beta = pick == 0 ? 0.0 : (pick <= X ? q[pick - 1] : INFINITY);
    return beta < 0 ? 0.0 : beta;
}

/*
 * root (B, H): lower bound of min c_h^T x (no bias) over the box and the
 * cuts; flags (B): bit 0 infeasible (a cut misses the box), bit 1 every cut
 * holds on the whole box (no cut used).  obj (B, H, X), A (B, K, X),
 * b (B, K), xL, xU (B, X).  max_rows <= 0: no cap.  X <= 64.
 */
int gho_constrained_concretize(int B, int H, int K, int X, const double *obj, const double *A,
                               const double *b, const double *xL, const double *xU,
                               int max_rows, double *root, int32_t *flags, int32_t *nkept,
                               const double *bias, double *lb_io)
{
    if (X < 1 || X > XMAX || K < 0 || H < 1) return 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int bx = 0; bx < B; ++bx) {
        const double *l = xL + (size_t)bx * X, *u = xU + (size_t)bx * X;
        double x0[XMAX], ep[XMAX], mx[XMAX], p[XMAX];
        for (int i = 0; i < X; ++i) {
            ep[i] = (u[i] - l[i]) / 2.0;
            x0[i] = (u[i] + l[i]) / 2.0;
            mx[i] = finite_abs(fmax(fabs(l[i]), fabs(u[i])));
        }
        int *keep = malloc(sizeof(int) * (K > 0 ? K : 1));
        double *dk = malloc(sizeof(double) * (K > 0 ? K : 1));
        int infeasible = 0, covered = 1;
        int nk = prepare_rows(K, X, A + (size_t)bx * K * X, b + (size_t)bx * K, x0, ep, max_rows,
                              keep, dk, &infeasible, &covered);
        flags[bx] = (infeasible ? 1 : 0) | (covered ? 2 : 0);
        int active = !covered && !infeasible;
        if (nkept) nkept[bx] = active ? nk : 0;
        for (int h = 0; h < H; ++h) {
            const double *c0 = obj + ((size_t)bx * H + h) * X;
            for (int i = 0; i < XMAX; ++i) p[i] = i < X ? c0[i] * x0[i] : 0.0;
            double base = red64(p);
            for (int i = 0; i < XMAX; ++i) p[i] = i < X ? fabs(c0[i]) * ep[i] : 0.0;
            double et0 = red64(p);
            for (int i = 0; i < XMAX; ++i) p[i] = i < X ? finite_abs(c0[i]) * mx[i] : 0.0;
            double objerr = red64(p) * k_round(X);
            double naive = widen(base - et0, objerr, -1.0);
            if (!isfinite(naive)) naive = -INFINITY;
            double out = naive;
            if (active) {
                double c[XMAX], cmag[XMAX];
                for (int i = 0; i < X; ++i) { c[i] = c0[i]; cmag[i] = finite_abs(c0[i]); }
                double dual = 0.0, dmag = 0.0;
                for (int t = 0; t < nk; ++t) {
                    const double *a = A + ((size_t)bx * K + keep[t]) * X;
                    int valid = isfinite(b[(size_t)bx * K + keep[t]]);
                    for (int i = 0; i < X; ++i) valid = valid && isfinite(a[i]);
                    double ar[XMAX];
                    for (int i = 0; i < X; ++i) ar[i] = valid ? a[i] : 0.0;
                    double beta = solve_beta(X, c, ar, dk[t], ep);
                    for (int i = 0; i < X; ++i) {
                        double upd = beta * ar[i];
                        c[i] = c[i] + upd;
                        cmag[i] = cmag[i] + finite_abs(upd);
                    }
                    double bd = beta * dk[t];
                    dual = dual + bd;
                    dmag = dmag + finite_abs(bd);
                }
                for (int i = 0; i < XMAX; ++i) p[i] = i < X ? fabs(c[i]) * ep[i] : 0.0;
                double et = red64(p);
                dual = dual - et;
                dmag = dmag + finite_abs(et);
                for (int i = 0; i < XMAX; ++i) p[i] = i < X ? cmag[i] * finite_abs(ep[i]) : 0.0;
                dmag = dmag + red64(p);
                double fin = base + dual;
                double err = objerr + (dmag + finite_abs(base)) * k_round(X + nk);
                fin = widen(fin, err, -1.0);
                if (!isfinite(fin)) fin = naive;
                out = fmax(fin, naive);
            }
            if (bias) {
                /* the bound: lbias widened and added as the box
                 * concretization adds it (concretize_bounds.py:280-304),
                 * never below the given bound; +inf for an infeasible
                 * domain (bounding.py:345-350) */
                size_t bh = (size_t)bx * H + h;
                double v = INFINITY;
                if (!infeasible) {
                    double bb = bias[bh];
                    double bw = widen(bb, finite_abs(bb) * k_round(X), -1.0);
                    double r = bw + out;
                    r = widen(r, finite_abs(r) * k_round(1), -1.0);
                    v = fmax(r, lb_io[bh]);
                }
                lb_io[bh] = v;
            } else {
                root[(size_t)bx * H + h] = out;
            }
        }
        free(keep);
        free(dk);
    }
    return 0;
}

/* The rows the sweep of box 0 would visit (test hook for the row
 * selection): keep[0..return) row indices, d their centred offsets. */
int gho_constraint_rows(int K, int X, const double *A, const double *b, const double *xL,
                        const double *xU, int max_rows, int32_t *keep, double *d, int32_t *flags)
{
    double x0[XMAX], ep[XMAX];
    if (X < 1 || X > XMAX) return -1;
    for (int i = 0; i < X; ++i) {
        ep[i] = (xU[i] - xL[i]) / 2.0;
        x0[i] = (xU[i] + xL[i]) / 2.0;
    }
    int *kp = malloc(sizeof(int) * (K > 0 ? K : 1));
    int inf = 0, cov = 1;
    int nk = prepare_rows(K, X, A, b, x0, ep, max_rows, kp, d, &inf, &cov);
    for (int t = 0; t < nk; ++t) keep[t] = kp[t];
    *flags = (inf ? 1 : 0) | (cov ? 2 : 0);
    free(kp);
    return nk;
}

/* ------------------------------------------------------------------------
 * Domain clipping (complete_verifier/input_split/clip.py:147-503,
 * clip_domains with lbias given, dm_lb recomputed, one iteration, no base
 * input constraints, thresholds as given): shrink each box along every
 * coordinate to the part where some point may still be unverified by the
 * planes lA_s x + lbias_s against thresholds thr_s.  Box b takes the planes
 * of plane set b % P (the children of a split share their parent's).
 *   concretize_bounds                clip.py:490-503
 *   _underflow, _concretize_error    clip.py:126-145
 *   certified domains emptied        clip.py:395-421
 *   _clip_main_fn (i = 0)            clip.py:147-212
 * Torch's min / max propagate NaN; nan_min / nan_max below do the same.
 */
#define TINY64 2.2250738585072014e-308

static double nan_min(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a < b ? a : b); }
static double nan_max(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b); }

int gho_clip_domains(int B, int P, int n, int S, const double *lA, const double *lbias,
                     const double *thr, const double *xL, const double *xU, double *oL,
                     double *oU, int32_t *state)
{
    if (n < 1 || n > XMAX || P < 1 || S < 1) return 1;
    const double und = (double)(n + 16) * TINY64 * EPS64;
    const double cfac = (double)(n + 4) / (double)(4 * (n + 16));
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; ++b) {
        const double *l = xL + (size_t)b * n, *u = xU + (size_t)b * n;
        const double *A = lA + (size_t)(b % P) * S * n, *c = lbias + (size_t)(b % P) * S;
        double *ol = oL + (size_t)b * n, *ou = oU + (size_t)b * n;
        for (int i = 0; i < n; ++i) { ol[i] = l[i]; ou[i] = u[i]; }
        int fin = 1;
        for (int s = 0; s < S; ++s) {
            fin = fin && isfinite(c[s]) && (isfinite(thr[s]) || (isinf(thr[s]) && thr[s] > 0));
            for (int i = 0; i < n; ++i) fin = fin && isfinite(A[(size_t)s * n + i]);
        }
        if (!fin) { state[b] = 0; continue; }
        double xh[XMAX], ep[XMAX], mx[XMAX], p[XMAX];
        double accU[XMAX], accL[XMAX];
        for (int i = 0; i < n; ++i) {
            xh[i] = (u[i] + l[i]) / 2;
            ep[i] = (u[i] - l[i]) / 2;
            mx[i] = finite_abs(fmax(fabs(l[i]), fabs(u[i])));
            accU[i] = INFINITY;
            accL[i] = -INFINITY;
        }
        int cert = 0;
        for (int s = 0; s < S; ++s) {
            const double *a = A + (size_t)s * n;
            for (int i = 0; i < XMAX; ++i) p[i] = i < n ? a[i] * xh[i] : 0.0;
            double t1 = red64(p);
            for (int i = 0; i < XMAX; ++i) p[i] = i < n ? fabs(a[i]) * ep[i] : 0.0;
            double t2 = red64(p);
            double dm = (t1 - t2) + c[s];
            for (int i = 0; i < XMAX; ++i) p[i] = i < n ? finite_abs(a[i]) * mx[i] : 0.0;
            double bre = (red64(p) + finite_abs(c[s])) * k_round(n);
            if ((dm - bre) - und > thr[s]) cert = 1;
            double derr = bre * cfac + und;
            for (int i = 0; i < n; ++i) {
                double lx = a[i] * xh[i], le = fabs(a[i]) * ep[i];
                double cm1 = (dm - lx) + le;
                double num = thr[s] - cm1;
                double mag = ((finite_abs(thr[s]) + finite_abs(dm)) + finite_abs(lx)) + finite_abs(le);
                double nerr = (mag * (4 * EPS64) + und) + derr;
                double lo = num, hi = num;
                if (isfinite(num)) { lo = num - nerr; hi = num + nerr; }
                double qlo = lo / a[i], qhi = hi / a[i];
                double cu = nextafter(nan_max(qlo, qhi), INFINITY);
                double cl = nextafter(nan_min(qlo, qhi), -INFINITY);
                accU[i] = nan_min(accU[i], a[i] > 0 ? cu : INFINITY);
                accL[i] = nan_max(accL[i], a[i] < 0 ? cl : -INFINITY);
            }
        }
        if (cert) {
            ol[0] = nextafter(u[0], INFINITY);
            state[b] = 2;
        } else {
            for (int i = 0; i < n; ++i) {
                ou[i] = nan_min(accU[i], u[i]);
                ol[i] = nan_max(accL[i], l[i]);
            }
            int empty = 0;
            for (int i = 0; i < n; ++i) empty = empty || !(ol[i] <= ou[i]);
            state[b] = empty ? 3 : 1;
        }
    }
    return 0;
}
