// ALM-PGD primal search for QCQPs, fused (gfx950, FP64).
//
// One wavefront runs one restart through every outer and inner iteration of
// the reference's ALMPGDSolver.solve_with_restarts
// (complete_verifier/nlp_solver/alm.py:1268-2065, inner solver
// _solve_subproblem_adam :551-788, AdamClipping_NLP.step :134-160):
//   L(x) = f_0(x) + 1/2 rho sum_s relu(g_s(x) + mu_s / rho)^2,
//   g_s = f_s - rhs_s, projected clipped Adam on the restart's box, then
//   mu <- max(0, mu + rho g), rho <- min(kappa rho, max_rho) where the
//   violation did not fall below 1/4 of the previous one, and the best
//   feasible point (violation <= feas_tol) is kept.
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
// wavefronts; wavefront 0 holds the iterate, sums the partial gradients in
// a fixed order, and applies the projected Adam step.
template <int K, int W, bool QL>
__global__ void __launch_bounds__(64 * W)
alm_pgd_kernel(int N, int n, int S, const double *__restrict__ Qg, const double *__restrict__ Cm,
               const double *__restrict__ d, const double *__restrict__ rhs,
               const double *__restrict__ lo_all, const double *__restrict__ hi_all,
               double *__restrict__ x_all, double *__restrict__ best_x_all,
               double *__restrict__ best_obj_all, gh_alm_config cfg)
{
    __shared__ double xs[64 * K];
    __shared__ double fs[kMaxS];
    __shared__ double mu[kMaxS];
    __shared__ double gpart[W][K][64];
    __shared__ double tab_bc1[kStepTab], tab_sq2[kStepTab], tab_decay[kStepTab];
    __shared__ double rho_s;
    __shared__ double qlds[QL ? kQlds : 1];
    const int r = blockIdx.x;
    if (r >= N) return;
    const bool tab = cfg.inner_steps <= kStepTab;
    if (tab)
        for (int t = threadIdx.x; t < cfg.inner_steps; t += 64 * W) {
            const int step = t + 1;
            tab_bc1[t] = 1.0 - pow(cfg.beta1, (double)step);
            tab_sq2[t] = sqrt(1.0 - pow(cfg.beta2, (double)step));
            tab_decay[t] = 1.0 - 0.9 * t / (double)max(cfg.inner_steps, 1);
        }
    const double *Qs = Qg;
    if constexpr (QL) {
        for (int e = threadIdx.x; e < S * n * n; e += 64 * W) qlds[e] = Qg[e];
        Qs = qlds;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool owner = w == 0;
    const double *lo_r = lo_all + (size_t)r * n, *hi_r = hi_all + (size_t)r * n;
    double x[K], lo[K], hi[K], lrv[K], ex[K], ex2[K], grad[K], bx[K];
    if (owner) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            const bool v = i < n;
            lo[k] = v ? lo_r[i] : 0.0;
            hi[k] = v ? hi_r[i] : 0.0;
            x[k] = v ? x_all[(size_t)r * n + i] : 0.0;
            lrv[k] = cfg.inner_lr * (hi[k] - lo[k]);
            bx[k] = x[k];
            xs[i] = x[k];
        }
        if (lane == 0) rho_s = cfg.rho0;
    }
    for (int s = threadIdx.x; s < S; s += 64 * W) mu[s] = 0.0;
    double best = INFINITY, prev_viol = INFINITY;
    __syncthreads();
    for (int outer = 0; outer < cfg.outer_max_iter; ++outer) {
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
            __syncthreads();  // partial gradients written, every x_j read
            if (owner) {
                if (W > 1) {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        double g = gpart[0][k][lane];
                        for (int v = 1; v < W; ++v) g += gpart[v][k][lane];
                        grad[k] = g;
                    }
                }
                double bc1, sq2, decay;
                if (tab) {
                    bc1 = tab_bc1[t];
                    sq2 = tab_sq2[t];
                    decay = tab_decay[t];
                } else {
                    const int step = t + 1;
                    bc1 = 1.0 - pow(cfg.beta1, (double)step);
                    sq2 = sqrt(1.0 - pow(cfg.beta2, (double)step));
                    decay = 1.0 - 0.9 * t / (double)max(cfg.inner_steps, 1);
                }
                const double lr = cfg.lr0 * decay;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    double g = grad[k];
                    if (isnan(g) || isinf(g)) g = 0.0;  // nan_to_num
                    ex[k] = ex[k] * cfg.beta1 + (1.0 - cfg.beta1) * g;
                    ex2[k] = ex2[k] * cfg.beta2 + (1.0 - cfg.beta2) * g * g;
                    const double denom = sqrt(ex2[k]) / sq2 + cfg.adam_eps;
                    double ms = lrv[k] * decay;
                    if (cfg.scale_eps_with_bias) ms = ms / bc1;
                    const double upd = fmin(fmax(-(ex[k] / denom) * lr, -ms), ms);
                    x[k] = fmin(fmax(x[k] + upd, lo[k]), hi[k]);
                    const int i = lane + 64 * k;
                    if (i < n) xs[i] = x[k];
                }
            }
            __syncthreads();  // the new x is in LDS
        }
        // outer bookkeeping at the new point (alm.py:1860-1914)
        eval_rows<K, W>(n, S, Qs, Cm, d, rhs, xs, fs, mu, rho_s, false, grad);
        __syncthreads();
        if (owner) {
            double viol = 0.0;
            for (int s = 1 + lane; s < S; s += 64) viol = fmax(viol, fmax(fs[s] - rhs[s], 0.0));
            viol = wmax(viol);
            const double f0 = fs[0];
            if (viol <= cfg.feas_tol && f0 < best) {
                best = f0;
#pragma unroll
                for (int k = 0; k < K; ++k) bx[k] = x[k];
            }
            const double rho = rho_s;
            for (int s = 1 + lane; s < S; s += 64) mu[s] = fmax(mu[s] + rho * (fs[s] - rhs[s]), 0.0);
            if (lane == 0 && viol > 0.25 * prev_viol) rho_s = fmin(rho * cfg.kappa, cfg.max_rho);
            prev_viol = viol;
        }
        __syncthreads();
    }
    if (owner) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                x_all[(size_t)r * n + i] = x[k];
                best_x_all[(size_t)r * n + i] = bx[k];
            }
        }
        if (lane == 0) best_obj_all[r] = best;
    }
}

// ------------------------------------------------------------------------
// Small problems (n <= 16, S <= 4 G): one wavefront per restart with every
// quantity of a step in registers.  Lane 16 r + i is column i of row slot r;
// row slot r holds rows s = 4 g + r (g < G).  It keeps row i of Qs[s]
// (Qs is symmetric) and c_s[i]; x, the box, Adam's moments and the best
// point are replicated in the four row slots, so x_j reaches every lane of a
// slot by DPP row_newbcast.  f_s is a 16-lane butterfly sum, the gradient
// sum_s w_s (2 y_s + c_s) adds the slots by permlane swaps, and the Adam
// step runs identically in every slot: a step touches no LDS but the
// per-step scalars and needs no barrier.  Same iteration as above; only the
// order of the sums differs.

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
                     gh_alm_config cfg)
{
    __shared__ double tab_bc1[kStepTab], tab_sq2[kStepTab], tab_decay[kStepTab];
    const int r = blockIdx.x;
    if (r >= N) return;
    const int lane = threadIdx.x, slot = lane >> 4, i = lane & 15;
    const bool col = i < n;
    const auto add = [](double a, double b) { return a + b; };
    const auto mx = [](double a, double b) { return fmax(a, b); };
    const bool tab = cfg.inner_steps <= kStepTab;
    if (tab)
        for (int t = lane; t < cfg.inner_steps; t += 64) {
            const int step = t + 1;
            tab_bc1[t] = 1.0 - pow(cfg.beta1, (double)step);
            tab_sq2[t] = sqrt(1.0 - pow(cfg.beta2, (double)step));
            tab_decay[t] = 1.0 - 0.9 * t / (double)max(cfg.inner_steps, 1);
        }
    double q[G][16], c[G], dd[G], rh[G], mu[G];
    bool valid[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int s = 4 * g + slot;
        valid[g] = s < S;
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
    const double lrv = cfg.inner_lr * (hi - lo);
    double bx = x, best = INFINITY, prev_viol = INFINITY, rho = cfg.rho0;
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
    for (int outer = 0; outer < cfg.outer_max_iter; ++outer) {
        double ex = 0.0, ex2 = 0.0;
        for (int t = 0; t < cfg.inner_steps; ++t) {
            double y[G], f[G];
            eval(y, f);
            double grad = 0.0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int s = 4 * g + slot;
                double wt = 1.0;
                if (s > 0) wt = rho * fmax((f[g] - rh[g]) + mu[g] / rho, 0.0);
                if (valid[g] && wt != 0.0) grad = fma(wt, 2.0 * y[g] + c[g], grad);
            }
            grad = sreduce(grad, add);
            double bc1, sq2, decay;
            if (tab) {
                bc1 = tab_bc1[t];
                sq2 = tab_sq2[t];
                decay = tab_decay[t];
            } else {
                const int step = t + 1;
                bc1 = 1.0 - pow(cfg.beta1, (double)step);
                sq2 = sqrt(1.0 - pow(cfg.beta2, (double)step));
                decay = 1.0 - 0.9 * t / (double)max(cfg.inner_steps, 1);
            }
            const double lr = cfg.lr0 * decay;
            double gg = grad;
            if (isnan(gg) || isinf(gg)) gg = 0.0;  // nan_to_num
            ex = ex * cfg.beta1 + (1.0 - cfg.beta1) * gg;
            ex2 = ex2 * cfg.beta2 + (1.0 - cfg.beta2) * gg * gg;
            const double denom = sqrt(ex2) / sq2 + cfg.adam_eps;
            double ms = lrv * decay;
            if (cfg.scale_eps_with_bias) ms = ms / bc1;
            const double upd = fmin(fmax(-(ex / denom) * lr, -ms), ms);
            x = fmin(fmax(x + upd, lo), hi);
        }
        // outer bookkeeping at the new point (alm.py:1860-1914)
        double y[G], f[G];
        eval(y, f);
        double v = 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (valid[g] && 4 * g + slot > 0) v = fmax(v, fmax(f[g] - rh[g], 0.0));
        const double viol = sreduce(v, mx);
        // f_0 lives in slot 0 (every lane of it); lane 0 is always active
        const long long f0b = __double_as_longlong(f[0]);
        const int f0lo = __builtin_amdgcn_readfirstlane((int)f0b);
        const int f0hi = __builtin_amdgcn_readfirstlane((int)(f0b >> 32));
        const double f0 = __longlong_as_double(((long long)f0hi << 32) | (unsigned)f0lo);
        if (viol <= cfg.feas_tol && f0 < best) {
            best = f0;
            bx = x;
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (valid[g] && 4 * g + slot > 0) mu[g] = fmax(mu[g] + rho * (f[g] - rh[g]), 0.0);
        if (viol > 0.25 * prev_viol) rho = fmin(rho * cfg.kappa, cfg.max_rho);
        prev_viol = viol;
    }
    if (slot == 0 && col) {
        x_all[(size_t)r * n + i] = x;
        best_x_all[(size_t)r * n + i] = bx;
    }
    if (lane == 0) best_obj_all[r] = best;
}

template <int K, int W>
void launch(int N, int n, int S, const double *Qs, const double *Cm, const double *d,
            const double *rhs, const double *lo, const double *hi, double *x, double *bx,
            double *bo, const gh_alm_config &cfg, hipStream_t stream)
{
    if (K == 1 && (size_t)S * n * n <= (size_t)kQlds)
        hipLaunchKernelGGL((alm_pgd_kernel<K, W, K == 1>), dim3(N), dim3(64 * W), 0, stream, N, n,
                           S, Qs, Cm, d, rhs, lo, hi, x, bx, bo, cfg);
    else
        hipLaunchKernelGGL((alm_pgd_kernel<K, W, false>), dim3(N), dim3(64 * W), 0, stream, N, n,
                           S, Qs, Cm, d, rhs, lo, hi, x, bx, bo, cfg);
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

extern "C" int gh_alm_pgd_f64(int N, int n, int S, const void *plan, const double *Cm,
                              const double *d, const double *rhs, const double *lo,
                              const double *hi, double *x, double *best_x, double *best_obj,
                              const gh_alm_config *cfg, hipStream_t stream)
{
    if (N == 0) return GH_OK;
    if (N < 0 || n < 1 || S < 1 || !plan || !Cm || !d || !rhs || !lo || !hi || !x || !best_x ||
        !best_obj || !cfg)
        return GH_EINVAL;
    if (S > kMaxS || n > 1024) return GH_EUNSUPPORTED;
    const double *Qs = static_cast<const double *>(plan);
    if (n <= 16 && S <= 8) {
        if (S <= 4)
            hipLaunchKernelGGL(alm_pgd_small_kernel<1>, dim3(N), dim3(64), 0, stream, N, n, S, Qs,
                               Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg);
        else
            hipLaunchKernelGGL(alm_pgd_small_kernel<2>, dim3(N), dim3(64), 0, stream, N, n, S, Qs,
                               Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg);
        return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
    }
    const int K = (n + 63) / 64;
    // rows split over 8 wavefronts for n <= 64 and 4 (one per SIMD, so the
    // iterate's registers fit without spills) above
#define GH_ALM(KK, WW) launch<KK, WW>(N, n, S, Qs, Cm, d, rhs, lo, hi, x, best_x, best_obj, *cfg, stream)
    if (K <= 1) { if (S >= 8) GH_ALM(1, 8); else if (S >= 4) GH_ALM(1, 4); else GH_ALM(1, 1); }
    else if (K <= 2) { if (S >= 4) GH_ALM(2, 4); else GH_ALM(2, 1); }
    else if (K <= 4) { if (S >= 4) GH_ALM(4, 4); else GH_ALM(4, 1); }
    else if (K <= 8) { if (S >= 4) GH_ALM(8, 4); else GH_ALM(8, 1); }
    else { if (S >= 4) GH_ALM(16, 4); else GH_ALM(16, 1); }
#undef GH_ALM
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
