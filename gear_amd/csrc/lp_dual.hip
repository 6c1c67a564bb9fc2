// FP64 LP dual certificate over boxes (gfx950).
//
// For each box b: a rigorous lower bound of
//     min c^T x + c0   s.t.   G x <= h,   l <= x <= u,
// the LP of the CROWN planes that bounding.py forms per domain
// (complete_verifier/input_split/bounding.py:489-520: c, c0 the objective
// row's plane, G, h the constraint rows' planes against their thresholds).
// As the reference's lp_bound_pdhg (complete_verifier/nlp_solver/
// dual_ascent.py:312-432) does, multipliers y >= 0 are proposed by
// diagonally preconditioned PDHG (Pock-Chambolle) on the box rescaled to
// [0, 1] after Ruiz equilibration, and the bound is the weak-duality value
//     D(y) = c0 - h^T y + sum_j min(r_j l_j, r_j u_j),   r = c + G^T y,
// evaluated with the error bound of rigorous_dual.py:1-54 (Higham Lemma 3.1;
// the sums may run in any order) at y = 0 and every `check_every`
// iterations, keeping the best.  A box whose certified value reaches a
// finite objective threshold thr[0] (the incumbent: the prune drops the box
// whatever its exact bound) stops at that checkpoint.  The proposal's precision never touches
// soundness; only the evaluation does.  (The reference's optional
// double-double recomputation of r is not done: the bound is rigorous
// without it, and at most looser where |G|^T y is large.)
//
// n, m <= 16: sixteen lanes per box, four boxes per wavefront
// (lp_dual_g16_kernel).  n, m <= 64: one wavefront per box
// (lp_dual_wave_kernel).  Larger: one
// workgroup (256 threads) per box; the per-box vectors live in dynamic LDS
// sized by (n, m), and G too when m n <= kGlds.
#include "common.h"
#include "gearhip_internal.h"

namespace {

using gh::bcast16_f64;

constexpr double U = 1.1102230246251565404236316680908203125e-16;  // 2^-53
constexpr double ETA = 4.9406564584124654e-324;                     // 2^-1074
constexpr int kThreads = 256;
constexpr int kMaxN = 512, kMaxM = 512;
constexpr int kGlds = 4096;  // doubles of G staged in LDS

__device__ __forceinline__ double gamma_k(double k) { return k * U / (1.0 - k * U); }
// the certified value reached the objective threshold: the box is pruned
__device__ __forceinline__ bool reached(double best, double thr0)
{
    return isfinite(thr0) && best >= thr0;
}

// block-wide sum of one double per thread (all threads return it)
__device__ double block_sum(double v, double *red)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < kThreads / 64; ++k) s += red[k];
    return s;
}
struct LpView {
    const double *G;  // row i at G + i * ldg
    int ldg;
    const double *c, *l, *u;
    double c0;
    int m, n;
};

// Rigorous lower bound of D(y) (rigorous_dual.py _reduced_float + _bound,
// single pass).  h, y in LDS (m); r / rho / k computed per column.
__device__ double rigorous_dual(const LpView &L, const double *Gl, const double *h,
                                const double *y, const unsigned char *act, double *red)
{
    const int m = L.m, n = L.n;
    const double infl = 1.0 + gamma_k(2.0 * (m + n) + 16.0);
    double t_sum = 0.0, t_abs = 0.0, me = 0.0, nz_t = 0.0;
    bool ok = true;
    for (int j = threadIdx.x; j < n; j += kThreads) {
        double r = 0.0, rho = 0.0, k = 0.0;
        for (int i = 0; i < m; ++i) {
            const double yi = y[i];
            const double g = Gl ? Gl[i * n + j] : L.G[(size_t)i * L.ldg + j];
            if (yi != 0.0 && g != 0.0 && act[i]) {
                const double p = g * yi;
                r += p;
                rho += fabs(p);
                k += 1.0;
            }
        }
        const double cj = L.c[j];
        r = r + cj;
        rho = rho + fabs(cj);
        const double kk = k > 0 ? gamma_k(k + 1.0) : 0.0;
        const double eps = kk * rho + 4.0 * ETA * (k + 2.0);
        const double lj = L.l[j], uj = L.u[j];
        const bool fl = isfinite(lj), fu = isfinite(uj);
        const double eps_sign = eps * infl;
        const bool pos = (r > eps_sign) && fl, neg = (r < -eps_sign) && fu;
        const bool zero = (rho == 0.0) && (k == 0.0);
        const double l0 = fl ? lj : 0.0, u0 = fu ? uj : 0.0;
        double t, M;
        if (fl && fu) {
            t = fmin(r * l0, r * u0);
            M = fmax(fabs(l0), fabs(u0));
        } else if (pos) {
            t = r * l0;
            M = fabs(l0);
        } else if (neg) {
            t = r * u0;
            M = fabs(u0);
        } else {
            t = 0.0;
            M = 0.0;
            if (!zero) ok = false;
        }
        t_sum += t;
        t_abs += fabs(t);
        me += M * eps;
        nz_t += (t != 0.0) ? 1.0 : 0.0;
    }
    double hy_sum = 0.0, hy_abs = 0.0, ny = 0.0;
    for (int i = threadIdx.x; i < m; i += kThreads) {
        if (y[i] > 0 && act[i]) {
            const double hy = h[i] * y[i];
            hy_sum += hy;
            hy_abs += fabs(hy);
            ny += 1.0;
        }
    }
    t_sum = block_sum(t_sum, red);
    t_abs = block_sum(t_abs, red);
    me = block_sum(me, red);
    nz_t = block_sum(nz_t, red);
    hy_sum = block_sum(hy_sum, red);
    hy_abs = block_sum(hy_abs, red);
    ny = block_sum(ny, red);
    const double okd = block_sum(ok ? 0.0 : 1.0, red);
    const double val = (L.c0 - hy_sum) + t_sum;
    const double N = 1.0 + ny + nz_t;
    const double under = (m + n + 3) * ETA;
    double E = gamma_k(N + 2.0) * (fabs(L.c0) + hy_abs + t_abs) + me + under;
    E = E * infl;
    E = E + 4.0 * U * (fabs(val) + E);
    const double lb = val - E;
    return (okd == 0.0 && isfinite(lb)) ? lb : -INFINITY;
}

__global__ void __launch_bounds__(kThreads)
lp_dual_kernel(int B, int m, int n, int S, const double *__restrict__ lA,
               const double *__restrict__ lbias, const double *__restrict__ thr,
               const double *__restrict__ xL, const double *__restrict__ xU, int iters,
               int ruiz_iters, int check_every, double *__restrict__ lb_out,
               double *__restrict__ y_out)
{
    // dynamic LDS: 6 n + 7 m + 4 doubles, G (m n doubles) when it fits, m flags
    extern __shared__ double smem[];
    double *cs = smem, *zs = cs + n, *zn = zs + n, *zub = zn + n, *tau = zub + n, *wDc = tau + n;
    double *ys = wDc + n, *hs = ys + m, *sig = hs + m, *Dr = sig + m, *h = Dr + m, *yv = h + m,
           *ybest = yv + m;
    double *red = ybest + m;
    double *Gl_s = red + kThreads / 64;
    const bool g_in_lds = m * n <= kGlds;
    unsigned char *act = reinterpret_cast<unsigned char *>(Gl_s + (g_in_lds ? m * n : 0));
    const int b = blockIdx.x;
    if (b >= B) return;
    const int t = threadIdx.x;
    LpView L;
    L.m = m;
    L.n = n;
    L.ldg = n;
    L.G = lA + (size_t)b * S * n + n;  // rows s = 1 .. S - 1
    L.c = lA + (size_t)b * S * n;      // row 0: the objective plane
    L.c0 = lbias[(size_t)b * S];
    L.l = xL + (size_t)b * n;
    L.u = xU + (size_t)b * n;
    // h = threshold - plane constant; rows with a non-finite rhs or row are
    // inactive (y = 0)
    for (int i = t; i < m; i += kThreads) {
        const double hi = thr[i + 1] - lbias[(size_t)b * S + i + 1];
        bool a = isfinite(hi);
        for (int j = 0; j < n && a; ++j) a = isfinite(L.G[(size_t)i * n + j]);
        act[i] = a ? 1 : 0;
        h[i] = a ? hi : 1.0;
        yv[i] = 0.0;
        ybest[i] = 0.0;
    }
    __syncthreads();  // the row flags are read across threads below
    const double *Gl = nullptr;
    if (g_in_lds) {
        for (int e = t; e < m * n; e += kThreads) {
            const int i = e / n;
            Gl_s[e] = act[i] ? L.G[e] : 0.0;
        }
        Gl = Gl_s;
    }
    __syncthreads();
    auto Gat = [&](int i, int j) -> double {
        return Gl ? Gl[i * n + j] : (act[i] ? L.G[(size_t)i * n + j] : 0.0);
    };
    double best = rigorous_dual(L, Gl, h, yv, act, red);
    const double thr0 = thr[0];
    if (iters > 0 && m > 0 && !reached(best, thr0)) {
        // x = l + w z, z in [0, 1]; hs = h - G l
        for (int j = t; j < n; j += kThreads) wDc[j] = fmax(L.u[j] - L.l[j], 0.0);  // w, then w Dc
        for (int i = t; i < m; i += kThreads) {
            double gl = 0.0;
            for (int j = 0; j < n; ++j) gl += Gat(i, j) * L.l[j];
            hs[i] = h[i] - gl;
            Dr[i] = 1.0;
        }
        for (int j = t; j < n; j += kThreads) zub[j] = 1.0;  // Dc, until the end
        __syncthreads();
        // Ruiz equilibration of Gs = Dr G diag(w) Dc
        for (int it = 0; it < ruiz_iters; ++it) {
            for (int i = t; i < m; i += kThreads) {
                double rn = 0.0;
                for (int j = 0; j < n; ++j) rn = fmax(rn, fabs(Gat(i, j) * wDc[j] * Dr[i] * zub[j]));
                sig[i] = rn > 0 ? sqrt(rn) : 1.0;
            }
            for (int j = t; j < n; j += kThreads) {
                double cn = 0.0;
                for (int i = 0; i < m; ++i) cn = fmax(cn, fabs(Gat(i, j) * wDc[j] * Dr[i] * zub[j]));
                tau[j] = cn > 0 ? sqrt(cn) : 1.0;
            }
            __syncthreads();
            for (int i = t; i < m; i += kThreads) Dr[i] = Dr[i] / sig[i];
            for (int j = t; j < n; j += kThreads) zub[j] = zub[j] / tau[j];
            __syncthreads();
        }
        // cs = c w Dc, hs = hs Dr, zub = 1 / Dc; wDc = w Dc; step sizes
        for (int j = t; j < n; j += kThreads) {
            const double w = wDc[j], dc = zub[j];
            wDc[j] = w * dc;
            cs[j] = L.c[j] * w * dc;
            zub[j] = 1.0 / dc;
        }
        for (int i = t; i < m; i += kThreads) hs[i] = hs[i] * Dr[i];
        __syncthreads();
        for (int j = t; j < n; j += kThreads) {
            double a = 0.0;
            for (int i = 0; i < m; ++i) a += fabs(Gat(i, j) * Dr[i] * wDc[j]);
            tau[j] = 1.0 / fmax(a, 1e-12);
            zs[j] = 0.5 * zub[j];
        }
        for (int i = t; i < m; i += kThreads) {
            double a = 0.0;
            for (int j = 0; j < n; ++j) a += fabs(Gat(i, j) * Dr[i] * wDc[j]);
            sig[i] = 1.0 / fmax(a, 1e-12);
            ys[i] = 0.0;
        }
        __syncthreads();
        for (int k = 0; k < iters; ++k) {
            // zn = clamp(zs - tau (cs + Gs^T ys), 0, zub)
            for (int j = t; j < n; j += kThreads) {
                double g = 0.0;
                for (int i = 0; i < m; ++i) g += Gat(i, j) * (Dr[i] * ys[i]);
                g = cs[j] + wDc[j] * g;
                zn[j] = fmin(fmax(zs[j] - tau[j] * g, 0.0), zub[j]);
            }
            __syncthreads();
            // ys = max(0, ys + sig (Gs (2 zn - zs) - hs)); one wavefront per row
            const int w = t >> 6, lane = t & 63;
            for (int i = w; i < m; i += kThreads / 64) {
                double a = 0.0;
                for (int j = lane; j < n; j += 64) a += Gat(i, j) * (wDc[j] * (2.0 * zn[j] - zs[j]));
                for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o);
                if (lane == 0) ys[i] = fmax(ys[i] + sig[i] * (Dr[i] * a - hs[i]), 0.0);
            }
            __syncthreads();
            for (int j = t; j < n; j += kThreads) zs[j] = zn[j];
            if ((k + 1) % check_every == 0 || k + 1 == iters) {
                for (int i = t; i < m; i += kThreads) yv[i] = act[i] ? ys[i] * Dr[i] : 0.0;
                __syncthreads();
                const double d = rigorous_dual(L, Gl, h, yv, act, red);
                if (isfinite(d) && (!isfinite(best) || d > best)) {
                    best = d;
                    for (int i = t; i < m; i += kThreads) ybest[i] = yv[i];
                }
            }
            __syncthreads();
            if (reached(best, thr0)) break;  // best is the same in every thread
        }
    }
    __syncthreads();
    if (t == 0) lb_out[b] = best;
    if (y_out)
        for (int i = t; i < m; i += kThreads) y_out[(size_t)b * m + i] = ybest[i];
}


// ------------------------------------------------------------------------
// One wavefront per box for n <= 64 and m <= 64: lane j owns column j
// (gradient, primal step), lane i owns row i (dual step); G, the vectors
// and the scalings live in the wavefront's LDS, so an iteration needs no
// workgroup barrier.  Same iteration and same rigorous evaluation as above.

// sum over the wavefront as the xor butterfly (o = 32, 16, 8, 4, 2, 1) on
// permlane swaps and DPP: every lane ends with the same value
__device__ __forceinline__ double wsum64(double v)
{
    v = v + gh::xor32_f64(v);
    v = v + gh::xor16_f64(v);
    return gh::tree16(v);
}

// sum_k a[k * sa] * b[k] for k < len, in order, from LDS: eight entries of
// each are loaded before their products are added, so eight loads are in
// flight where a plain loop waits for every load in turn
__device__ __forceinline__ double dot_lds(int len, const double *a, int sa, const double *b,
                                          int sb = 1)
{
    double acc = 0.0;
    int k = 0;
    for (; k + 8 <= len; k += 8) {
        double av[8], bv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            av[u] = a[(k + u) * sa];
            bv[u] = b[(k + u) * sb];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += av[u] * bv[u];
    }
    for (; k < len; ++k) acc += a[k * sa] * b[k * sb];
    return acc;
}

// per-wavefront LDS: G (m x n, row stride ld = n rounded up to odd), then
// y[64], v[64].  Lane j reads column j as G[i ld + j] (consecutive) and lane
// i row i as G[i ld + j]: an odd stride puts the 32 lanes of a half-wave on
// distinct bank pairs, so the row reads have no bank conflicts either.
__host__ __device__ __forceinline__ int wave_ld(int n) { return n | 1; }
struct WaveLds {
    double *G, *y, *v;
    int ld;
};

// rigorous D(y) of the wavefront's LP; y in LDS, lane j handles column j
__device__ double rigorous_dual_wave(int m, int n, const WaveLds W, double c0, double cj,
                                     double lj, double uj, double hi, bool act_i)
{
    const int lane = threadIdx.x & 63;
    const double infl = 1.0 + gamma_k(2.0 * (m + n) + 16.0);
    double t = 0.0, M = 0.0, eps = 0.0;
    bool ok = true;
    if (lane < n) {
        double r = 0.0, rho = 0.0, k = 0.0;
        for (int i = 0; i < m; ++i) {
            const double yi = W.y[i], g = W.G[i * W.ld + lane];
            if (yi != 0.0 && g != 0.0) {
                const double p = g * yi;
                r += p;
                rho += fabs(p);
                k += 1.0;
            }
        }
        r = r + cj;
        rho = rho + fabs(cj);
        const double kk = k > 0 ? gamma_k(k + 1.0) : 0.0;
        eps = kk * rho + 4.0 * ETA * (k + 2.0);
        const bool fl = isfinite(lj), fu = isfinite(uj);
        const double eps_sign = eps * infl;
        const bool pos = (r > eps_sign) && fl, neg = (r < -eps_sign) && fu;
        const bool zero = (rho == 0.0) && (k == 0.0);
        const double l0 = fl ? lj : 0.0, u0 = fu ? uj : 0.0;
        if (fl && fu) {
            t = fmin(r * l0, r * u0);
            M = fmax(fabs(l0), fabs(u0));
        } else if (pos) {
            t = r * l0;
            M = fabs(l0);
        } else if (neg) {
            t = r * u0;
            M = fabs(u0);
        } else if (!zero) {
            ok = false;
        }
    }
    double hy = 0.0, ny = 0.0;
    if (lane < m && act_i && W.y[lane] > 0) {
        hy = hi * W.y[lane];
        ny = 1.0;
    }
    const double t_sum = wsum64(t), t_abs = wsum64(fabs(t)), me = wsum64(M * eps);
    const double nz_t = wsum64(t != 0.0 ? 1.0 : 0.0);
    const double hy_sum = wsum64(hy), hy_abs = wsum64(fabs(hy)), nyy = wsum64(ny);
    const double bad = wsum64(ok ? 0.0 : 1.0);
    const double val = (c0 - hy_sum) + t_sum;
    const double N = 1.0 + nyy + nz_t;
    double E = gamma_k(N + 2.0) * (fabs(c0) + hy_abs + t_abs) + me + (m + n + 3) * ETA;
    E = E * infl;
    E = E + 4.0 * U * (fabs(val) + E);
    const double lb = val - E;
    return (bad == 0.0 && isfinite(lb)) ? lb : -INFINITY;
}

__global__ void __launch_bounds__(64)
lp_dual_wave_kernel(int B, int m, int n, int S, const double *__restrict__ lA,
                    const double *__restrict__ lbias, const double *__restrict__ thr,
                    const double *__restrict__ xL, const double *__restrict__ xU, int iters,
                    int ruiz_iters, int check_every, double *__restrict__ lb_out,
                    double *__restrict__ y_out)
{
    extern __shared__ double wl[];
    const int ld = wave_ld(n);
    WaveLds W{wl, wl + (size_t)m * ld, wl + (size_t)m * ld + 64, ld};
    const int b = blockIdx.x;
    if (b >= B) return;
    const int lane = threadIdx.x;
    const double *Gb = lA + (size_t)b * S * n + n;
    const double c0 = lbias[(size_t)b * S];
    const double cj = lane < n ? lA[(size_t)b * S * n + lane] : 0.0;
    const double lj = lane < n ? xL[(size_t)b * n + lane] : 0.0;
    const double uj = lane < n ? xU[(size_t)b * n + lane] : 0.0;
    // row lane: threshold, activity
    double hi = 1.0;
    bool act_i = false;
    if (lane < m) {
        const double h0 = thr[lane + 1] - lbias[(size_t)b * S + lane + 1];
        bool a = isfinite(h0);
        for (int j = 0; j < n && a; ++j) a = isfinite(Gb[(size_t)lane * n + j]);
        act_i = a;
        hi = a ? h0 : 1.0;
    }
    for (int i = 0; i < m; ++i) {
        const bool ai = __shfl(act_i ? 1 : 0, i) != 0;
        if (lane < n) W.G[i * ld + lane] = ai ? Gb[(size_t)i * n + lane] : 0.0;
    }
    W.y[lane] = 0.0;
    __syncthreads();
    double best = rigorous_dual_wave(m, n, W, c0, cj, lj, uj, hi, act_i);
    double ybest = 0.0;
    const double thr0 = thr[0];
    if (iters > 0 && m > 0 && !reached(best, thr0)) {
        // scalings: lane j holds w_j and Dc_j, lane i holds Dr_i
        const double w = lane < n ? fmax(uj - lj, 0.0) : 0.0;
        double dc = 1.0, dr = 1.0;
        // hs = h - G l (l broadcast through LDS: every lane reads all of it)
        double hs = 0.0;
        W.v[lane] = lj;
        __syncthreads();
        if (lane < m) {
            double gl = 0.0;
            for (int j = 0; j < n; ++j) gl += W.G[lane * ld + j] * W.v[j];
            hs = hi - gl;
        }
        __syncthreads();
        for (int it = 0; it < ruiz_iters; ++it) {
            // broadcast current scalings through LDS: v = w Dc (columns), y = Dr (rows)
            W.v[lane] = w * dc;
            W.y[lane] = dr;
            __syncthreads();
            double rn = 0.0, cn = 0.0;
            if (lane < m)
                for (int j = 0; j < n; ++j) rn = fmax(rn, fabs(W.G[lane * ld + j] * W.v[j] * dr));
            if (lane < n)
                for (int i = 0; i < m; ++i) cn = fmax(cn, fabs(W.G[i * ld + lane] * W.y[i] * w * dc));
            __syncthreads();
            if (lane < m) dr = dr / (rn > 0 ? sqrt(rn) : 1.0);
            if (lane < n) dc = dc / (cn > 0 ? sqrt(cn) : 1.0);
        }
        const double wdc = w * dc;
        const double cs = cj * w * dc, zub = 1.0 / dc;
        hs = hs * dr;
        W.v[lane] = lane < n ? wdc : 0.0;
        W.y[lane] = lane < m ? dr : 0.0;
        __syncthreads();
        double tau = 0.0, sig = 0.0;
        if (lane < n) {
            double a = 0.0;
            for (int i = 0; i < m; ++i) a += fabs(W.G[i * ld + lane] * W.y[i] * wdc);
            tau = 1.0 / fmax(a, 1e-12);
        }
        if (lane < m) {
            double a = 0.0;
            for (int j = 0; j < n; ++j) a += fabs(W.G[lane * ld + j] * dr * W.v[j]);
            sig = 1.0 / fmax(a, 1e-12);
        }
        double zs = lane < n ? 0.5 * zub : 0.0, ys = 0.0;
        // the dual step's row sums split over P <= 4 lane groups of m lanes
        const int P = min(4, 64 / m), rp = lane / m, ri = lane - rp * m;
        __syncthreads();
        for (int k = 0; k < iters; ++k) {
            // lane i publishes Dr_i ys_i; lane j: gradient and primal step
            W.y[lane] = lane < m ? dr * ys : 0.0;
            __syncthreads();
            double zn = 0.0;
            if (lane < n) {
                const double g = cs + wdc * dot_lds(m, W.G + lane, ld, W.y);
                zn = fmin(fmax(zs - tau * g, 0.0), zub);
            }
            // lane j publishes w_j Dc_j (2 zn_j - zs_j); dual step of row i:
            // lanes i + p m (p < P) sum the columns j = p, p + P, ..., lane i
            // adds the parts
            W.v[lane] = lane < n ? wdc * (2.0 * zn - zs) : 0.0;
            __syncthreads();
            const double part =
                rp < P ? dot_lds((n - rp + P - 1) / P, W.G + ri * ld + rp, P, W.v + rp, P) : 0.0;
            double a = part;
            for (int p = 1; p < P; ++p) a += __shfl(part, ri + p * m);
            if (lane < m) ys = fmax(ys + sig * (dr * a - hs), 0.0);
            zs = zn;
            if ((k + 1) % check_every == 0 || k + 1 == iters) {
                const double yi = (lane < m && act_i) ? ys * dr : 0.0;
                __syncthreads();
                W.y[lane] = yi;
                __syncthreads();
                const double d = rigorous_dual_wave(m, n, W, c0, cj, lj, uj, hi, act_i);
                if (isfinite(d) && (!isfinite(best) || d > best)) {
                    best = d;
                    ybest = yi;
                }
            }
            __syncthreads();
            if (reached(best, thr0)) break;  // best is the same in every lane
        }
    }
    if (lane == 0) lb_out[b] = best;
    if (y_out && lane < m) y_out[(size_t)b * m + lane] = ybest;
}

// ------------------------------------------------------------------------
// Sixteen lanes per box for n <= 16 and m <= 16: four boxes per wavefront,
// sixteen per 256-thread workgroup.  Lane s of a group holds column s of G
// (its gradient entry) and row s of G (its dual step) in registers; the
// vectors each step needs from the other lanes of the group come by DPP
// row_newbcast (one 16-lane row of the wavefront is one box), so the
// iteration touches neither LDS nor a barrier.  Same iteration (its two
// matrix-vector sums as fused multiply-adds: the proposal's rounding never
// touches soundness) and the same rigorous evaluation as the kernels above
// (the evaluation's sums run over the group's lanes; its error bound holds
// for any summation order).

constexpr int kG = 16;              // lanes per box
constexpr int kGThreads = 256;      // 16 boxes per workgroup

// sum over the box's 16 lanes (xor butterfly o = 8, 4, 2, 1 on DPP)
__device__ __forceinline__ double gsum16(double v) { return gh::tree16(v); }

// rigorous D(y) of the group's LP: lane s holds column s of G (gc), y_s (row
// lanes) and its column's data
template <int NM>
__device__ double rigorous_dual_g16(int m, int n, const double (&gc)[NM], double ys_eval,
                                    double c0, double cj, double lj, double uj, double hi,
                                    bool act_i)
{
    const int s = threadIdx.x & (kG - 1);
    const double infl = 1.0 + gamma_k(2.0 * (m + n) + 16.0);
    double r = 0.0, rho = 0.0, k = 0.0;
    gh::for_lanes_upto<NM>(m, [&](auto I) {
        const double yi = bcast16_f64<decltype(I)::value>(ys_eval), g = gc[decltype(I)::value];
        if (yi != 0.0 && g != 0.0) {
            const double p = g * yi;
            r += p;
            rho += fabs(p);
            k += 1.0;
        }
    });
    double t = 0.0, M = 0.0, eps = 0.0;
    bool ok = true;
    if (s < n) {
        r = r + cj;
        rho = rho + fabs(cj);
        const double kk = k > 0 ? gamma_k(k + 1.0) : 0.0;
        eps = kk * rho + 4.0 * ETA * (k + 2.0);
        const bool fl = isfinite(lj), fu = isfinite(uj);
        const double eps_sign = eps * infl;
        const bool pos = (r > eps_sign) && fl, neg = (r < -eps_sign) && fu;
        const bool zero = (rho == 0.0) && (k == 0.0);
        const double l0 = fl ? lj : 0.0, u0 = fu ? uj : 0.0;
        if (fl && fu) {
            t = fmin(r * l0, r * u0);
            M = fmax(fabs(l0), fabs(u0));
        } else if (pos) {
            t = r * l0;
            M = fabs(l0);
        } else if (neg) {
            t = r * u0;
            M = fabs(u0);
        } else if (!zero) {
            ok = false;
        }
    }
    double hy = 0.0, ny = 0.0;
    if (s < m && act_i && ys_eval > 0) {
        hy = hi * ys_eval;
        ny = 1.0;
    }
    const double t_sum = gsum16(t), t_abs = gsum16(fabs(t)), me = gsum16(M * eps);
    const double nz_t = gsum16(t != 0.0 ? 1.0 : 0.0);
    const double hy_sum = gsum16(hy), hy_abs = gsum16(fabs(hy)), nyy = gsum16(ny);
    const double bad = gsum16(ok ? 0.0 : 1.0);
    const double val = (c0 - hy_sum) + t_sum;
    const double N = 1.0 + nyy + nz_t;
    double E = gamma_k(N + 2.0) * (fabs(c0) + hy_abs + t_abs) + me + (m + n + 3) * ETA;
    E = E * infl;
    E = E + 4.0 * U * (fabs(val) + E);
    const double lb = val - E;
    return (bad == 0.0 && isfinite(lb)) ? lb : -INFINITY;
}

template <int NM>
__global__ void __launch_bounds__(kGThreads)
lp_dual_g16_kernel(int B, int m, int n, int S, const double *__restrict__ lA,
                   const double *__restrict__ lbias, const double *__restrict__ thr,
                   const double *__restrict__ xL, const double *__restrict__ xU, int iters,
                   int ruiz_iters, int check_every, double *__restrict__ lb_out,
                   double *__restrict__ y_out)
{
    const int b = blockIdx.x * (kGThreads / kG) + (threadIdx.x >> 4);
    if (b >= B) return;  // a whole 16-lane row leaves: no broadcast reads it
    const int s = threadIdx.x & (kG - 1);
    const bool col = s < n, row = s < m;
    const double *Gb = lA + (size_t)b * S * n + n;
    const double c0 = lbias[(size_t)b * S];
    const double cj = col ? lA[(size_t)b * S * n + s] : 0.0;
    const double lj = col ? xL[(size_t)b * n + s] : 0.0;
    const double uj = col ? xU[(size_t)b * n + s] : 0.0;
    // row s (threshold, activity, entries) and column s of G, inactive rows zeroed
    double gr[NM], gc[NM];
    double hi = 1.0;
    bool act_i = false;
#pragma unroll
    for (int j = 0; j < NM; ++j) gr[j] = (row && j < n) ? Gb[(size_t)s * n + j] : 0.0;
    if (row) {
        const double h0 = thr[s + 1] - lbias[(size_t)b * S + s + 1];
        bool a = isfinite(h0);
#pragma unroll
        for (int j = 0; j < NM; ++j) a = a && isfinite(gr[j]);
        act_i = a;
        hi = a ? h0 : 1.0;
    }
#pragma unroll
    for (int j = 0; j < NM; ++j) gr[j] = act_i ? gr[j] : 0.0;
    const double actd = act_i ? 1.0 : 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) gc[i] = 0.0;
    gh::for_lanes_upto<NM>(m, [&](auto I) {
        constexpr int i = decltype(I)::value;
        const double g = col ? Gb[(size_t)i * n + s] : 0.0;
        gc[i] = bcast16_f64<i>(actd) != 0.0 ? g : 0.0;
    });
    double best = rigorous_dual_g16(m, n, gc, 0.0, c0, cj, lj, uj, hi, act_i);
    double ybest = 0.0;
    const double thr0 = thr[0];
    if (iters > 0 && m > 0 && !reached(best, thr0)) {
        // scalings: lane j holds w_j and Dc_j (column), lane i holds Dr_i (row)
        const double w = col ? fmax(uj - lj, 0.0) : 0.0;
        double dc = 1.0, dr = 1.0;
        // hs = h - G l
        double gl = 0.0;
        gh::for_lanes_upto<NM>(n, [&](auto J) {
            gl += gr[decltype(J)::value] * bcast16_f64<decltype(J)::value>(lj);
        });
        double hs = row ? hi - gl : 0.0;
        for (int it = 0; it < ruiz_iters; ++it) {
            const double v = w * dc, y = dr;
            double rn = 0.0, cn = 0.0;
            gh::for_lanes_upto<NM>(n, [&](auto J) {
                constexpr int j = decltype(J)::value;
                rn = fmax(rn, fabs(gr[j] * bcast16_f64<j>(v) * dr));
            });
            gh::for_lanes_upto<NM>(m, [&](auto I) {
                constexpr int i = decltype(I)::value;
                cn = fmax(cn, fabs(gc[i] * bcast16_f64<i>(y) * w * dc));
            });
            if (row) dr = dr / (rn > 0 ? sqrt(rn) : 1.0);
            if (col) dc = dc / (cn > 0 ? sqrt(cn) : 1.0);
        }
        const double wdc = col ? w * dc : 0.0;
        const double cs = cj * w * dc, zub = 1.0 / dc;
        hs = hs * dr;
        const double drr = row ? dr : 0.0;
        double tau = 0.0, sig = 0.0;
        {
            double a = 0.0, q = 0.0;
            gh::for_lanes_upto<NM>(m, [&](auto I) {
                constexpr int i = decltype(I)::value;
                a += fabs(gc[i] * bcast16_f64<i>(drr) * wdc);
            });
            gh::for_lanes_upto<NM>(n, [&](auto J) {
                constexpr int j = decltype(J)::value;
                q += fabs(gr[j] * dr * bcast16_f64<j>(wdc));
            });
            if (col) tau = 1.0 / fmax(a, 1e-12);
            if (row) sig = 1.0 / fmax(q, 1e-12);
        }
        double zs = col ? 0.5 * zub : 0.0, ys = 0.0;
        for (int k = 0; k < iters; ++k) {
            // gradient of column s from the rows' Dr_i ys_i; primal step
            const double yd = row ? dr * ys : 0.0;
            double g = 0.0;
            gh::for_lanes_upto<NM>(m, [&](auto I) {
                constexpr int i = decltype(I)::value;
                g = fma(gc[i], bcast16_f64<i>(yd), g);
            });
            g = cs + wdc * g;
            const double zn = col ? fmin(fmax(zs - tau * g, 0.0), zub) : 0.0;
            // dual step of row s from the columns' w_j Dc_j (2 zn_j - zs_j)
            const double v = col ? wdc * (2.0 * zn - zs) : 0.0;
            double a = 0.0;
            gh::for_lanes_upto<NM>(n, [&](auto J) {
                constexpr int j = decltype(J)::value;
                a = fma(gr[j], bcast16_f64<j>(v), a);
            });
            if (row) ys = fmax(ys + sig * (dr * a - hs), 0.0);
            zs = zn;
            if ((k + 1) % check_every == 0 || k + 1 == iters) {
                const double yi = (row && act_i) ? ys * dr : 0.0;
                const double d = rigorous_dual_g16(m, n, gc, yi, c0, cj, lj, uj, hi, act_i);
                if (isfinite(d) && (!isfinite(best) || d > best)) {
                    best = d;
                    ybest = yi;
                }
                // best is the same in the group's 16 lanes; the group leaves
                // as a whole, and no other row's broadcasts read its lanes
                if (reached(best, thr0)) break;
            }
        }
    }
    if (s == 0) lb_out[b] = best;
    if (y_out && row) y_out[(size_t)b * m + s] = ybest;
}

}  // namespace

extern "C" int gh_lp_dual_bound_f64(int B, int n, int S, const double *lA, const double *lbias,
                                    const double *thr, const double *xL, const double *xU,
                                    int iters, double *lb, double *y, hipStream_t stream)
{
    if (B == 0) return GH_OK;
    if (B < 0 || n < 1 || S < 1 || !lA || !lbias || !thr || !xL || !xU || !lb || iters < 0)
        return GH_EINVAL;
    const int m = S - 1;
    if (n > kMaxN || m > kMaxM) return GH_EUNSUPPORTED;
    if (n <= kG && m <= kG) {
        // register arrays of 8 entries for n, m <= 8 (fewer VGPRs, more
        // wavefronts per SIMD), 16 above
        const int per = kGThreads / kG;
        const dim3 grid((B + per - 1) / per);
        if (n <= 8 && m <= 8)
            hipLaunchKernelGGL(lp_dual_g16_kernel<8>, grid, dim3(kGThreads), 0, stream, B, m, n,
                               S, lA, lbias, thr, xL, xU, iters, 10, 50, lb, y);
        else
            hipLaunchKernelGGL(lp_dual_g16_kernel<16>, grid, dim3(kGThreads), 0, stream, B, m, n,
                               S, lA, lbias, thr, xL, xU, iters, 10, 50, lb, y);
        return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
    }
    if (n <= 64 && m <= 64) {
        const size_t wlds = sizeof(double) * ((size_t)m * wave_ld(n) + 128);
        hipLaunchKernelGGL(lp_dual_wave_kernel, dim3(B), dim3(64), wlds, stream, B, m, n, S, lA,
                           lbias, thr, xL, xU, iters, 10, 50, lb, y);
        return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
    }
    const size_t lds = sizeof(double) * (6 * (size_t)n + 7 * (size_t)m + kThreads / 64 +
                                         ((size_t)m * n <= kGlds ? (size_t)m * n : 0)) + m + 8;
    hipLaunchKernelGGL(lp_dual_kernel, dim3(B), dim3(kThreads), lds, stream, B, m, n, S, lA, lbias,
                       thr, xL, xU, iters, 10, 50, lb, y);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
