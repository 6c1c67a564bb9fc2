// ALM-PGD primal search for QCQPs, fused (gfx950, FP64).
//
// Each restart runs the reference's ALMPGDSolver.solve_with_restarts as it
// runs one restart of one domain (complete_verifier/nlp_solver/alm.py:
// 1268-2065; inner solver _solve_subproblem_adam :551-788,
// AdamClipping_NLP.step :134-160, the presets' ALMConfig values):
//   L(x) = f_0(x) + 1/2 rho sum_s relu(g_s(x) + mu_s / rho)^2, g_s = f_s - rhs_s;
//   inner: Adam reset per outer iteration, lr = inner_lr * (largest box
//   edge) * a linear decay 1 -> 0.1, each coordinate's step clipped to
//   inner_lr * its width (over the first-moment bias correction), projected
//   onto the box; it ends when the gradient norm reaches eta_k (:764-776) or
//   L stops changing (rtol over `inner_plateau_patience` checks, :678-699);
//   outer: best point (feasible first, else least violated, :1650-1700), the
//   objective-patience and first-feasibility exits (:1800-1849), then the
//   multiplier / penalty / tolerance schedule (:1851-1914): root mode
//   (update_multipliers_always) updates mu every iteration and raises rho
//   when the violation fell by less than 1.25x; BaB mode updates mu and
//   shrinks eps_k, eta_k by rho where the violation is within eps_k, and
//   otherwise raises rho (violation > 1e-3) and resets eps_k = rho^-alpha,
//   eta_k = rho^-beta.
// With Qs = (Q + Q^T) / 2, f_s(x) = x^T Qs x + c_s x + d_s and
// grad f_s = 2 Qs x + c_s, so one pass over Qs per row gives both: lane i
// holds entries i, i + 64, ... of x, reads Qs[s][j][i] (coalesced; Qs is
// symmetric) and x_j from LDS; f_s is a wavefront sum; the gradient
// accumulates w_s (2 y + c_s) with w_0 = 1, w_s = rho relu(g_s + mu_s / rho).
// A restart's rows are split over the W wavefronts of its workgroup; for
// n <= 16 and S <= 8 one wavefront holds all of a restart's rows in
// registers (alm_pgd_small_kernel).
#include "common.h"
#include "gearhip_internal.h"

namespace {

using gh::xor32_f64;
using gh::xor16_f64;
using gh::dpp_f64;
using gh::bcast16_f64;
using gh::for_lanes16;

constexpr int kRowRor8 = 0x128, kRowRor4 = 0x124;   // lane (i + 8) / (i + 4) mod 16
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]

// Wavefront reduction as the butterfly v = op(v, v[lane ^ o]) for
// o = 32, 16, 8, 4, 2, 1, all in VALU/DPP.  After the o = 8 step lanes i and
// i ^ 8 of a row hold the same value, so lane i + 4 mod 16 (row_ror:4) holds
// what lane i ^ 4 holds; op is commutative, so every lane ends with the
// butterfly's value bit for bit.
template <class Op>
__device__ __forceinline__ double wreduce(double v, Op op)
{
    v = op(v, xor32_f64(v));
    v = op(v, xor16_f64(v));
    v = op(v, dpp_f64<kRowRor8>(v));
    v = op(v, dpp_f64<kRowRor4>(v));
    v = op(v, dpp_f64<kQuadXor2>(v));
    v = op(v, dpp_f64<kQuadXor1>(v));
    return v;
}
__device__ __forceinline__ double wsum(double v)
{
    return wreduce(v, [](double a, double b) { return a + b; });
}
__device__ __forceinline__ double wmax(double v)
{
    return wreduce(v, [](double a, double b) { return fmax(a, b); });
}

__global__ void __launch_bounds__(256)
alm_symmetrize_kernel(int n, int S, const double *__restrict__ Q, double *__restrict__ Qs)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)S * n * n;
    if (t >= total) return;
    const size_t s = t / ((size_t)n * n);
    const size_t r = t - s * n * n;
    const size_t i = r / n, j = r - i * n;
    const double *q = Q + s * n * n;
    Qs[t] = 0.5 * (q[i * n + j] + q[j * n + i]);
}

constexpr int kMaxS = 1024;  // rows held in LDS per restart
// Adam's per-step scalars (1 - beta1^t, sqrt(1 - beta2^t), the decay) are
// tabulated in LDS once per workgroup for up to this many inner steps; a
// longer schedule computes them in the step.  Same expressions either way.
constexpr int kStepTab = 256;
// Q (S n^2 doubles) is copied into LDS when it fits this many doubles and
// n <= 64; above, the rows are read from global memory every step.
constexpr int kQlds = 4096;

// the inner learning-rate factor at step t: VectorLinearLR from 1 to 0.1
// over the inner steps (complete_verifier/nlp_solver/utils.py:139-147)
__device__ __forceinline__ double lr_factor(int t, int total)
{
    const double progress = fmin(1.0, (double)t / (double)max(total, 1));
    return 1.0 + (0.1 - 1.0) * progress;
}

// rho^a as torch evaluates a tensor power (a = 1/2 is a square root)
__device__ __forceinline__ double rpow(double rho, double a)
{
    return a == 0.5 ? sqrt(rho) : pow(rho, a);
}

// One restart's outer-loop state (alm.py:1380-1390, 1410-1535), the same on
// every lane that holds it.
struct Sched {
    double rho, eps, eta;
    double best, best_viol;   // the best point: feasible first, else least violated
    double prev_best, glob_best, prev_viol;
    int best_feas, no_imp, plat_start;
};

__device__ __forceinline__ void sched_init(Sched &S_, const gh_alm_config &c, double phi0,
                                           double viol0)
{
    S_.rho = c.rho0;
    S_.eps = 1.0 / rpow(S_.rho, c.alpha);
    S_.eta = 1.0 / rpow(S_.rho, c.beta);
    S_.best = phi0;
    S_.best_viol = viol0;
    S_.best_feas = viol0 <= c.feas_tol;
    S_.prev_best = INFINITY;
    S_.glob_best = INFINITY;
    S_.prev_viol = INFINITY;
    S_.no_imp = 0;
    S_.plat_start = -1;
}

// best-point tracking (alm.py:1650-1700): true when x becomes the best point
__device__ __forceinline__ bool sched_track(Sched &S_, const gh_alm_config &c, double phi,
                                            double viol)
{
    const bool feas = viol <= c.feas_tol;
    const bool improved = feas && (!S_.best_feas || phi < S_.best);
    const bool better_viol = !feas && !S_.best_feas && viol < S_.best_viol;
    if (improved || better_viol) {
        S_.best = phi;
        S_.best_viol = viol;
        if (improved) S_.best_feas = 1;
        return true;
    }
    return false;
}

// the outer exits after iteration k (alm.py:1800-1849): objective patience,
// then first-feasibility patience
__device__ __forceinline__ bool sched_stop(Sched &S_, const gh_alm_config &c, int k)
{
    if (c.obj_patience >= 0) {
        const bool imp = S_.best < S_.prev_best - c.obj_improvement_tol;
        S_.no_imp = imp ? 0 : S_.no_imp + 1;
        S_.prev_best = fmin(S_.prev_best, S_.best);
        if (S_.no_imp >= c.obj_patience && S_.best_feas) return true;
    }
    if (c.first_feas_patience >= 0 && S_.best_feas) {
        if (S_.plat_start < 0 || S_.best < S_.glob_best - c.obj_improvement_tol) {
            S_.plat_start = k;
            S_.glob_best = S_.best;
        } else if (k - S_.plat_start >= c.first_feas_patience) {
            return true;
        }
    }
    return false;
}

// whether the multipliers move this iteration (with the rho before the update)
__device__ __forceinline__ bool sched_mu_moves(const Sched &S_, const gh_alm_config &c,
                                               double viol)
{
    return c.update_multipliers_always || viol <= S_.eps;
}

// rho, eps_k, eta_k after iteration k (alm.py:1851-1914)
__device__ __forceinline__ void sched_update(Sched &S_, const gh_alm_config &c, int k,
                                             double viol)
{
    if (c.update_multipliers_always) {
        if (k >= 1) {
            const double ratio = S_.prev_viol / fmax(viol, 1e-12);
            if (ratio < 1.25 && viol > c.feas_tol) S_.rho = fmin(c.kappa * S_.rho, c.max_rho);
        }
        S_.eps = 1.0 / rpow(S_.rho, c.alpha);
        S_.eta = 1.0 / rpow(S_.rho, c.beta);
        S_.prev_viol = viol;
    } else if (viol <= S_.eps) {
        S_.eps = S_.eps / S_.rho;
        S_.eta = S_.eta / S_.rho;
    } else {
        if (viol > 1e-3) S_.rho = fmin(c.kappa * S_.rho, c.max_rho);
        S_.eps = 1.0 / rpow(S_.rho, c.alpha);
        S_.eta = 1.0 / rpow(S_.rho, c.beta);
    }
    S_.eps = fmax(S_.eps, c.eps_final);
    S_.eta = fmax(S_.eta, c.eta_final);
}

// the per-step scalars of the inner loop
struct StepTab {
    double *bc1, *sq2, *fac;
    __device__ __forceinline__ void fill(const gh_alm_config &c, int tid, int nthr)
    {
        for (int t = tid; t < c.inner_steps && t < kStepTab; t += nthr) {
            const int step = t + 1;
            bc1[t] = 1.0 - pow(c.beta1, (double)step);
            sq2[t] = sqrt(1.0 - pow(c.beta2, (double)step));
            fac[t] = lr_factor(t, c.inner_steps);
        }
    }
    __device__ __forceinline__ void get(const gh_alm_config &c, int t, double &b1, double &s2,
                                        double &f) const
    {
        if (t < kStepTab) {
            b1 = bc1[t];
            s2 = sq2[t];
            f = fac[t];
        } else {
            const int step = t + 1;
            b1 = 1.0 - pow(c.beta1, (double)step);
            s2 = sqrt(1.0 - pow(c.beta2, (double)step));
            f = lr_factor(t, c.inner_steps);
        }
    }
};

// the outer iteration's learning-rate scale (alm.py:1600-1604)
__device__ __forceinline__ double lr_scale(const gh_alm_config &c, int k)
{
    double p = 1.0;
    for (int i = 0; i < k; ++i) p *= c.inner_lr_outer_decay;
    return fmax(p, c.inner_lr_min_factor);
}

// trace record r (0 = the start, k + 1 = after outer iteration k)
__device__ __forceinline__ double *trace_rec(double *trace, int N, int r, int rec, int S,
                                             const gh_alm_config &c)
{
    const int W = 8 + 2 * (S - 1);
    return trace + ((size_t)r * (c.outer_max_iter + 1) + rec) * W;
}

// f_s for the rows s = w, w + W, ... of this wavefront at x (in LDS), and
// optionally this wavefront's part of the ALM gradient.
template <int K, int W>
__device__ __forceinline__ void eval_rows(int n, int S, const double *__restrict__ Qs,
                                          const double *__restrict__ Cm,
                                          const double *__restrict__ d, const double *__restrict__ rhs,
                                          const double *xs, double *fs, const double *mu, double rho,
                                          bool want_grad, double grad[K])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        grad[k] = 0.0;
        x[k] = xs[lane + 64 * k];
    }
    for (int s = w; s < S; s += W) {
        const double *q = Qs + (size_t)s * n * n;
        // four partial sums per row (j mod 4) keep four FMA chains and
        // sixteen loads per lane in flight
        double y[K], ya[K], yb[K], yc[K];
        for (int k = 0; k < K; ++k) y[k] = ya[k] = yb[k] = yc[k] = 0.0;
        int j = 0;
#pragma unroll 4
        for (; j + 4 <= n; j += 4) {
            const double x0 = xs[j], x1 = xs[j + 1], x2 = xs[j + 2], x3 = xs[j + 3];
            const double *q0 = q + (size_t)j * n;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = lane + 64 * k;
                if (i < n) {
                    y[k] = fma(q0[i], x0, y[k]);
                    ya[k] = fma(q0[n + i], x1, ya[k]);
                    yb[k] = fma(q0[2 * n + i], x2, yb[k]);
                    yc[k] = fma(q0[3 * n + i], x3, yc[k]);
                }
            }
        }
        for (; j < n; ++j) {
            const double xj = xs[j];
            const double *qj = q + (size_t)j * n;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = lane + 64 * k;
                if (i < n) y[k] = fma(qj[i], xj, y[k]);
            }
        }
        double part = 0.0;
        double cv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            y[k] = (y[k] + ya[k]) + (yb[k] + yc[k]);
            cv[k] = i < n ? Cm[(size_t)s * n + i] : 0.0;
            part = fma(x[k], y[k] + cv[k], part);
        }
        const double f = wsum(part) + d[s];
        if (lane == 0) fs[s] = f;
        if (want_grad) {
            double wt = 1.0;
            if (s > 0) {
                const double g = f - rhs[s];
                wt = rho * fmax(g + mu[s] / rho, 0.0);
            }
            if (wt != 0.0) {
#pragma unroll
                for (int k = 0; k < K; ++k) grad[k] = fma(wt, 2.0 * y[k] + cv[k], grad[k]);
            }
        }
    }
}

// One restart per workgroup of W wavefronts: the rows are split over the
// wavefronts; wavefront 0 holds the iterate and the schedule, sums the
// partial gradients in a fixed order and applies the projected Adam step.
template <int K, int W, bool QL>
__global__ void __launch_bounds__(64 * W)
alm_pgd_kernel(int N, int n, int S, const double *__restrict__ Qg, const double *__restrict__ Cm,
               const double *__restrict__ d, const double *__restrict__ rhs,
               const double *__restrict__ lo_all, const double *__restrict__ hi_all,
               double *__restrict__ x_all, double *__restrict__ best_x_all,
               double *__restrict__ best_obj_all, gh_alm_config cfg, double *__restrict__ trace)
{
    __shared__ double xs[64 * K];
    __shared__ double fs[kMaxS];
    __shared__ double mu[kMaxS];
    __shared__ double gpart[W][K][64];
    __shared__ double tab_bc1[kStepTab], tab_sq2[kStepTab], tab_fac[kStepTab];
    __shared__ double rho_s;
    __shared__ int brk_s, stop_s;
    __shared__ double qlds[QL ? kQlds : 1];
    const int r = blockIdx.x;
    if (r >= N) return;
    StepTab tab{tab_bc1, tab_sq2, tab_fac};
    tab.fill(cfg, threadIdx.x, 64 * W);
    const double *Qs = Qg;
    if constexpr (QL) {
        for (int e = threadIdx.x; e < S * n * n; e += 64 * W) qlds[e] = Qg[e];
        Qs = qlds;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool owner = w == 0;
    const int m = S - 1;
    const double *lo_r = lo_all + (size_t)r * n, *hi_r = hi_all + (size_t)r * n;
    double x[K], lo[K], hi[K], ex[K], ex2[K], grad[K], bx[K];
    double edge = 0.0;
    if (owner) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            const bool v = i < n;
            lo[k] = v ? lo_r[i] : 0.0;
            hi[k] = v ? hi_r[i] : 0.0;
            x[k] = v ? x_all[(size_t)r * n + i] : 0.0;
            bx[k] = x[k];
            if (v) {
                xs[i] = x[k];
                edge = fmax(edge, hi[k] - lo[k]);
            }
        }
        edge = wmax(edge);
        if (lane == 0) {
            rho_s = cfg.rho0;
            brk_s = 0;
            stop_s = 0;
        }
    }
    for (int q = threadIdx.x; q < S; q += 64 * W) mu[q] = 0.0;
    __syncthreads();
    // the starting point is the first best point (alm.py:1505-1522)
    eval_rows<K, W>(n, S, Qs, Cm, d, rhs, xs, fs, mu, cfg.rho0, false, grad);
    __syncthreads();
    Sched sc;
    double gnorm = INFINITY;
    auto violation = [&]() {
        double v = 0.0;
        for (int q = 1 + lane; q < S; q += 64) v = fmax(v, fmax(fs[q] - rhs[q], 0.0));
        return wmax(v);
    };
    auto record = [&](int rec, double viol, int steps, int flag) {
        if (!trace) return;
        double *T = trace_rec(trace, N, r, rec, S, cfg);
        if (lane == 0) {
            T[0] = sc.rho;
            T[1] = sc.eps;
            T[2] = sc.eta;
            T[3] = viol;
            T[4] = fs[0];
            T[5] = (double)steps;
            T[6] = (double)flag;
            T[7] = gnorm;
        }
        for (int q = lane; q < m; q += 64) {
            T[8 + q] = mu[q + 1];
            T[8 + m + q] = fs[q + 1] - rhs[q + 1];
        }
    };
    if (owner) {
        const double viol0 = violation();
        sched_init(sc, cfg, fs[0], viol0);
        record(0, viol0, 0, 0);
    }
    int stopped_at = cfg.outer_max_iter;
    for (int outer = 0; outer < cfg.outer_max_iter; ++outer) {
        const double eff = cfg.inner_lr * fmax(lr_scale(cfg, outer), 0.0);
        const double base_lr = eff * fmax(edge, 1e-10);
        double prevL = 0.0;
        int have_prev = 0, plateau = 0, steps = 0;
        if (owner) {
#pragma unroll
            for (int k = 0; k < K; ++k) ex[k] = ex2[k] = 0.0;
        }
        for (int t = 0; t < cfg.inner_steps; ++t) {
            const double rho = rho_s;
            eval_rows<K, W>(n, S, Qs, Cm, d, rhs, xs, fs, mu, rho, true, grad);
            if (W > 1) {
#pragma unroll
                for (int k = 0; k < K; ++k) gpart[w][k][lane] = grad[k];
            }
            __syncthreads();  // partial gradients and f written, every x_j read
            if (owner) {
                // L and the forward's finiteness (alm.py:655-676)
                double pen = 0.0, dual = 0.0;
                bool fin = true;
                for (int q = 1 + lane; q < S; q += 64) {
                    const double g = fs[q] - rhs[q];
                    const double sh = fmax(g + mu[q] / rho, 0.0);
                    pen += sh * sh;
                    dual += mu[q] * mu[q];
                    fin = fin && isfinite(g);
                }
                pen = wsum(pen);
                dual = wsum(dual);
                fin = wmax(fin ? 0.0 : 1.0) == 0.0;
                const double phi = fs[0];
                const double L = m > 0 ? phi + 0.5 * rho * pen : phi;
                const bool ffin = fin && isfinite(phi) && isfinite(L);
                bool brk = false;
                if (cfg.inner_plateau_check > 0 && t > 0 && t % cfg.inner_plateau_check == 0) {
                    const double cur = m > 0 ? L - 0.5 / rho * dual : L;
                    if (have_prev) {
                        const double rel = fabs(cur - prevL) / (fabs(prevL) + 1e-8);
                        plateau = rel < cfg.inner_plateau_rtol ? plateau + 1 : 0;
                        brk = plateau >= cfg.inner_plateau_patience;
                    }
                    prevL = cur;
                    have_prev = 1;
                }
                if (!brk) {
                    if (W > 1) {
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            double g = gpart[0][k][lane];
                            for (int v = 1; v < W; ++v) g += gpart[v][k][lane];
                            grad[k] = g;
                        }
                    }
                    double bc1, sq2, fac;
                    tab.get(cfg, t, bc1, sq2, fac);
                    const double lr = base_lr * fac;
                    double bad = 0.0, gsq = 0.0;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const int i = lane + 64 * k;
                        double g = grad[k];
                        if (i < n && !isfinite(g)) bad = 1.0;
                        if (isnan(g) || isinf(g)) g = 0.0;  // nan_to_num
                        if (i >= n) g = 0.0;
                        gsq = fma(g, g, gsq);
                        ex[k] = ex[k] * cfg.beta1 + (1.0 - cfg.beta1) * g;
                        ex2[k] = ex2[k] * cfg.beta2 + (1.0 - cfg.beta2) * g * g;
                        const double denom = sqrt(ex2[k]) / sq2 + cfg.adam_eps;
                        double ms = eff * (hi[k] - lo[k]);
                        if (cfg.scale_eps_with_bias) ms = ms / bc1;
                        const double upd = fmin(fmax(-(ex[k] / denom) * lr, -ms), ms);
                        x[k] = fmin(fmax(x[k] + upd, lo[k]), hi[k]);
                        if (i < n) xs[i] = x[k];
                    }
                    ++steps;
                    bad = wmax(bad);
                    gnorm = (bad != 0.0 || !ffin) ? INFINITY : sqrt(wsum(gsq));
                    brk = gnorm <= sc.eta;
                }
                if (lane == 0) brk_s = brk ? 1 : 0;
            }
            __syncthreads();  // the new x is in LDS
            if (brk_s) break;
        }
        // outer bookkeeping at the new point
        eval_rows<K, W>(n, S, Qs, Cm, d, rhs, xs, fs, mu, rho_s, false, grad);
        __syncthreads();
        if (owner) {
            const double viol = violation();
            if (sched_track(sc, cfg, fs[0], viol)) {
#pragma unroll
                for (int k = 0; k < K; ++k) bx[k] = x[k];
            }
            const bool stop = sched_stop(sc, cfg, outer);
            if (!stop) {
                const double rho = sc.rho;
                if (sched_mu_moves(sc, cfg, viol))
                    for (int q = 1 + lane; q < S; q += 64) mu[q] = fmax(mu[q] + rho * (fs[q] - rhs[q]), 0.0);
                sched_update(sc, cfg, outer, viol);
            }
            record(outer + 1, viol, steps, stop ? 1 : 0);
            if (lane == 0) {
                rho_s = sc.rho;
                brk_s = 0;
                stop_s = stop ? 1 : 0;
            }
        }
        __syncthreads();
        if (stop_s) {
            stopped_at = outer + 1;
            break;
        }
    }
    if (owner) {
        if (trace)
            for (int rec = stopped_at + 1; rec <= cfg.outer_max_iter; ++rec) {
                double *T = trace_rec(trace, N, r, rec, S, cfg);
                for (int q = lane; q < 8 + 2 * m; q += 64) T[q] = q == 6 ? 2.0 : NAN;
            }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                x_all[(size_t)r * n + i] = x[k];
                best_x_all[(size_t)r * n + i] = bx[k];
            }
        }
        if (lane == 0) best_obj_all[r] = sc.best_feas ? sc.best : INFINITY;
    }
}

// ------------------------------------------------------------------------
// Small problems (n <= 16, S <= 4 G): one wavefront per restart with every
// quantity of a step in registers.  Lane 16 r + i is column i of row slot r;
// row slot r holds rows s = 4 g + r (g < G).  It keeps row i of Qs[s]
// (Qs is symmetric) and c_s[i]; x, the box, Adam's moments and the best
// point are replicated in the four row slots, so x_j reaches every lane of a
// slot by DPP row_newbcast.  f_s is a 16-lane butterfly sum, sums over rows
// add the slots by permlane swaps, and the Adam step and the schedule run
// identically in every slot: a step touches no LDS but the per-step scalars
// and needs no barrier.  Same iteration as above; only the order of the
// sums differs.

// butterfly over the 16 lanes of a row (o = 8, 4, 2, 1), as in wreduce
template <class Op>
__device__ __forceinline__ double rreduce(double v, Op op)
{
    v = op(v, dpp_f64<kRowRor8>(v));
    v = op(v, dpp_f64<kRowRor4>(v));
    v = op(v, dpp_f64<kQuadXor2>(v));
    v = op(v, dpp_f64<kQuadXor1>(v));
    return v;
}
// across the four row slots (o = 32, 16)
template <class Op>
__device__ __forceinline__ double sreduce(double v, Op op)
{
    v = op(v, xor32_f64(v));
    v = op(v, xor16_f64(v));
    return v;
}

template <int G>
__global__ void __launch_bounds__(64)
alm_pgd_small_kernel(int N, int n, int S, const double *__restrict__ Qs,
                     const double *__restrict__ Cm, const double *__restrict__ d,
                     const double *__restrict__ rhs, const double *__restrict__ lo_all,
                     const double *__restrict__ hi_all, double *__restrict__ x_all,
                     double *__restrict__ best_x_all, double *__restrict__ best_obj_all,
                     gh_alm_config cfg, double *__restrict__ trace)
{
    __shared__ double tab_bc1[kStepTab], tab_sq2[kStepTab], tab_fac[kStepTab];
    const int r = blockIdx.x;
    if (r >= N) return;
    const int lane = threadIdx.x, slot = lane >> 4, i = lane & 15;
    const bool col = i < n;
    const int m = S - 1;
    const auto add = [](double a, double b) { return a + b; };
    const auto mx = [](double a, double b) { return fmax(a, b); };
    StepTab tab{tab_bc1, tab_sq2, tab_fac};
    tab.fill(cfg, lane, 64);
    double q[G][16], c[G], dd[G], rh[G], mu[G];
    bool valid[G], con[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int s = 4 * g + slot;
        valid[g] = s < S;
        con[g] = valid[g] && s > 0;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            q[g][j] = (valid[g] && col && j < n) ? Qs[((size_t)s * n + i) * n + j] : 0.0;
        c[g] = (valid[g] && col) ? Cm[(size_t)s * n + i] : 0.0;
        dd[g] = valid[g] ? d[s] : 0.0;
        rh[g] = valid[g] ? rhs[s] : 0.0;
        mu[g] = 0.0;
    }
    const double lo = col ? lo_all[(size_t)r * n + i] : 0.0;
    const double hi = col ? hi_all[(size_t)r * n + i] : 0.0;
    double x = col ? x_all[(size_t)r * n + i] : 0.0;
    const double edge = rreduce(col ? hi - lo : 0.0, mx);
    double bx = x;
    __syncthreads();  // the step tables
    // y_s = Qs[s] x (row i), f_s at x for every row of the slot
    auto eval = [&](double (&y)[G], double (&f)[G]) {
#pragma unroll
        for (int g = 0; g < G; ++g) y[g] = 0.0;
        for_lanes16(n, [&](auto J) {
            constexpr int j = decltype(J)::value;
            const double xj = bcast16_f64<j>(x);
#pragma unroll
            for (int g = 0; g < G; ++g) y[g] = fma(q[g][j], xj, y[g]);
        });
#pragma unroll
        for (int g = 0; g < G; ++g) f[g] = rreduce(x * (y[g] + c[g]), add) + dd[g];
    };
    // f_0 lives in slot 0 (every lane of it); lane 0 is always active
    auto first = [](double v) {
        const long long b = __double_as_longlong(v);
        const int lo32 = __builtin_amdgcn_readfirstlane((int)b);
        const int hi32 = __builtin_amdgcn_readfirstlane((int)(b >> 32));
        return __longlong_as_double(((long long)hi32 << 32) | (unsigned)lo32);
    };
    auto violation = [&](const double (&f)[G]) {
        double v = 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (con[g]) v = fmax(v, fmax(f[g] - rh[g], 0.0));
        return sreduce(v, mx);
    };
    Sched sc;
    double gnorm = INFINITY;
    auto record = [&](int rec, const double (&f)[G], double viol, int steps, int flag) {
        if (!trace) return;
        double *T = trace_rec(trace, N, r, rec, S, cfg);
        if (lane == 0) {
            T[0] = sc.rho;
            T[1] = sc.eps;
            T[2] = sc.eta;
            T[3] = viol;
            T[4] = f[0];
            T[5] = (double)steps;
            T[6] = (double)flag;
            T[7] = gnorm;
        }
        if (i == 0) {
#pragma unroll
            for (int g = 0; g < G; ++g)
                if (con[g]) {
                    const int s = 4 * g + slot;
                    T[8 + s - 1] = mu[g];
                    T[8 + m + s - 1] = f[g] - rh[g];
                }
        }
    };
    {
        double y[G], f[G];
        eval(y, f);
        const double viol0 = violation(f);
        sched_init(sc, cfg, first(f[0]), viol0);
        record(0, f, viol0, 0, 0);
    }
    int stopped_at = cfg.outer_max_iter;
    for (int outer = 0; outer < cfg.outer_max_iter; ++outer) {
        const double eff = cfg.inner_lr * fmax(lr_scale(cfg, outer), 0.0);
        const double base_lr = eff * fmax(edge, 1e-10);
        const double ms0 = eff * (hi - lo);
        const double rho = sc.rho;
        double ex = 0.0, ex2 = 0.0, prevL = 0.0;
        int have_prev = 0, plateau = 0, steps = 0;
        for (int t = 0; t < cfg.inner_steps; ++t) {
            double y[G], f[G];
            eval(y, f);
            double pen = 0.0, dual = 0.0, nf = 0.0, grad = 0.0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (con[g]) {
                    const double gv = f[g] - rh[g];
                    const double sh = fmax(gv + mu[g] / rho, 0.0);
                    pen += sh * sh;
                    dual += mu[g] * mu[g];
                    if (!isfinite(gv)) nf = 1.0;
                }
                const int s = 4 * g + slot;
                double wt = 1.0;
                if (s > 0) wt = rho * fmax((f[g] - rh[g]) + mu[g] / rho, 0.0);
                if (valid[g] && wt != 0.0) grad = fma(wt, 2.0 * y[g] + c[g], grad);
            }
            // one lane per slot carries the slot's row terms
            pen = sreduce(i == 0 ? pen : 0.0, add);
            pen = rreduce(pen, add);
            dual = sreduce(i == 0 ? dual : 0.0, add);
            dual = rreduce(dual, add);
            nf = sreduce(nf, mx);
            grad = sreduce(grad, add);
            const double phi = first(f[0]);
            const double L = m > 0 ? phi + 0.5 * rho * pen : phi;
            const bool ffin = nf == 0.0 && isfinite(phi) && isfinite(L);
            if (cfg.inner_plateau_check > 0 && t > 0 && t % cfg.inner_plateau_check == 0) {
                const double cur = m > 0 ? L - 0.5 / rho * dual : L;
                bool brk = false;
                if (have_prev) {
                    const double rel = fabs(cur - prevL) / (fabs(prevL) + 1e-8);
                    plateau = rel < cfg.inner_plateau_rtol ? plateau + 1 : 0;
                    brk = plateau >= cfg.inner_plateau_patience;
                }
                prevL = cur;
                have_prev = 1;
                if (brk) break;
            }
            double bc1, sq2, fac;
            tab.get(cfg, t, bc1, sq2, fac);
            const double lr = base_lr * fac;
            double gg = col ? grad : 0.0;
            const double bad = rreduce((col && !isfinite(gg)) ? 1.0 : 0.0, mx);
            if (isnan(gg) || isinf(gg)) gg = 0.0;  // nan_to_num
            ex = ex * cfg.beta1 + (1.0 - cfg.beta1) * gg;
            ex2 = ex2 * cfg.beta2 + (1.0 - cfg.beta2) * gg * gg;
            const double denom = sqrt(ex2) / sq2 + cfg.adam_eps;
            double ms = ms0;
            if (cfg.scale_eps_with_bias) ms = ms / bc1;
            const double upd = fmin(fmax(-(ex / denom) * lr, -ms), ms);
            x = fmin(fmax(x + upd, lo), hi);
            ++steps;
            gnorm = (bad != 0.0 || !ffin) ? INFINITY : sqrt(rreduce(gg * gg, add));
            if (gnorm <= sc.eta) break;
        }
        // outer bookkeeping at the new point
        double y[G], f[G];
        eval(y, f);
        const double viol = violation(f);
        if (sched_track(sc, cfg, first(f[0]), viol)) bx = x;
        const bool stop = sched_stop(sc, cfg, outer);
        if (!stop) {
            const bool moves = sched_mu_moves(sc, cfg, viol);
#pragma unroll
            for (int g = 0; g < G; ++g)
                if (con[g] && moves) mu[g] = fmax(mu[g] + rho * (f[g] - rh[g]), 0.0);
            sched_update(sc, cfg, outer, viol);
        }
        record(outer + 1, f, viol, steps, stop ? 1 : 0);
        if (stop) {
            stopped_at = outer + 1;
            break;
        }
    }
    if (trace)
        for (int rec = stopped_at + 1; rec <= cfg.outer_max_iter; ++rec) {
            double *T = trace_rec(trace, N, r, rec, S, cfg);
            for (int qd = lane; qd < 8 + 2 * m; qd += 64) T[qd] = qd == 6 ? 2.0 : NAN;
        }
    if (slot == 0 && col) {
        x_all[(size_t)r * n + i] = x;
        best_x_all[(size_t)r * n + i] = bx;
    }
    if (lane == 0) best_obj_all[r] = sc.best_feas ? sc.best : INFINITY;
}

template <int K, int W>
void launch(int N, int n, int S, const double *Qs, const double *Cm, const double *d,
            const double *rhs, const double *lo, const double *hi, double *x, double *bx,
            double *bo, const gh_alm_config &cfg, double *trace, hipStream_t stream)
{
    if (K == 1 && (size_t)S * n * n <= (size_t)kQlds)
        hipLaunchKernelGGL((alm_pgd_kernel<K, W, K == 1>), dim3(N), dim3(64 * W), 0, stream, N, n,
                           S, Qs, Cm, d, rhs, lo, hi, x, bx, bo, cfg, trace);
    else
        hipLaunchKernelGGL((alm_pgd_kernel<K, W, false>), dim3(N), dim3(64 * W), 0, stream, N, n,
                           S, Qs, Cm, d, rhs, lo, hi, x, bx, bo, cfg, trace);
}

}  // namespace

extern "C" size_t gh_alm_plan_bytes(int n, int S)
{
    if (n < 1 || S < 1) return 0;
    return (size_t)S * n * n * sizeof(double);
}

extern "C" int gh_alm_prepare(int n, int S, const double *Q, void *plan, hipStream_t stream)
{
    if (n < 1 || S < 1 || !Q || !plan) return GH_EINVAL;
    const size_t total = (size_t)S * n * n;
    hipLaunchKernelGGL(alm_symmetrize_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, n, S, Q, static_cast<double *>(plan));
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" int gh_alm_trace_width(int S) { return S < 1 ? 0 : 8 + 2 * (S - 1); }

extern "C" int gh_alm_pgd_f64(int N, int n, int S, const void *plan, const double *Cm,
                              const double *d, const double *rhs, const double *lo,
                              const double *hi, double *x, double *best_x, double *best_obj,
                              const gh_alm_config *cfg, double *trace, hipStream_t stream)
{
    if (N == 0) return GH_OK;
    if (N < 0 || n < 1 || S < 1 || !plan || !Cm || !d || !rhs || !lo || !hi || !x || !best_x ||
        !best_obj || !cfg)
        return GH_EINVAL;
    if (S > kMaxS || n > 1024) return GH_EUNSUPPORTED;
    if (cfg->outer_max_iter < 0 || cfg->inner_steps < 0) return GH_EINVAL;
    const double *Qs = static_cast<const double *>(plan);
    if (n <= 16 && S <= 8) {
        if (S <= 4)
            hipLaunchKernelGGL(alm_pgd_small_kernel<1>, dim3(N), dim3(64), 0, stream, N, n, S, Qs,
                               Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg, trace);
        else
            hipLaunchKernelGGL(alm_pgd_small_kernel<2>, dim3(N), dim3(64), 0, stream, N, n, S, Qs,
                               Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg, trace);
        return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
    }
    const int K = (n + 63) / 64;
    // rows split over 8 wavefronts for n <= 64 and 4 (one per SIMD, so the
    // iterate's registers fit without spills) above
#define GH_ALM(KK, WW) \
    launch<KK, WW>(N, n, S, Qs, Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg, trace, stream)
    if (K <= 1) { if (S >= 8) GH_ALM(1, 8); else if (S >= 4) GH_ALM(1, 4); else GH_ALM(1, 1); }
    else if (K <= 2) { if (S >= 4) GH_ALM(2, 4); else GH_ALM(2, 1); }
    else if (K <= 4) { if (S >= 4) GH_ALM(4, 4); else GH_ALM(4, 1); }
    else if (K <= 8) { if (S >= 4) GH_ALM(8, 4); else GH_ALM(8, 1); }
    else { if (S >= 4) GH_ALM(16, 4); else GH_ALM(16, 1); }
#undef GH_ALM
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
