// Fused FP64 CROWN lower bound of a QCQP over a batch of boxes (gfx950).
//
// Rows s = 0..S-1 are quadratics f_s(x) = x^T Q_s x + c_s^T x + d_s.  The
// bound follows the reference's plain CROWN pass over the lowered graph
//   y = Linear(x; [Q_0; ..; Q_{S-1}]),  z = Mul(x, y) (coeff-aware McCormick),
//   q = ReduceSum(z),  l = Linear(x; Cm, d),  f = Add(q, l),  C = eye(S)
// (auto_LiRPA/backward_bound.py:880-1206 driver; operators/linear.py:171-480
// and :607-653 Linear backward / IBP; operators/bivariate.py:40-75, 190-261,
// 401-488 Mul; concretize_bounds.py:267-321 and concretize_func.py:552-569
// box concretization; fp_enclosure.py and backward_bound.py:550-823
// rounding enclosures).
//
// Mapping.  One 16-lane group owns RB boxes; lane l owns the variable
// indices l, l+16, l+32, l+48 of each (n <= 16 * NCH).  A block of 256
// threads is 16 groups = 16*RB boxes and walks all S rows, so the box data is
// staged once and each Q_s tile (staged in LDS, padded stride n+1) is reused
// by all 16*RB boxes.  Per (box, row):
//   1. IBP of y_s = Q_s x: centre / radius / magnitude, three FMA chains
//      over i per output j (lane-private, index order);
//   2. McCormick plane choice and coefficients, interval of z, in registers;
//   3. beta -> LDS, backward A = beta^T Q_s (FMA chain over j per output i);
//   4. merges, A67 charges and concretization as 16-lane partials + xor tree.
// The oracle (oracle/crown_qcqp.c) fixes the same evaluation order, so the
// result is bit-identical to it.
#include "common.h"
#include "gearhip_internal.h"

namespace {

struct QcqpConsts {
    double kn, k1, k2, h1, hs, hsn;
};

template <int RB, int NCH>
__global__ __launch_bounds__(256) void qcqp_crown_f64_kernel(
    int B, int n, int S, const double* __restrict__ Q, const double* __restrict__ Cm,
    const double* __restrict__ d, const double* __restrict__ xL, const double* __restrict__ xU,
    double* __restrict__ lb, double* __restrict__ lA, double* __restrict__ lbias, QcqpConsts k)
{
    extern __shared__ double smem[];
    constexpr int BOXES = 16 * RB;
    const int NJ = 16 * NCH;
    const int ldq = n + 1;
    double* sQ = smem;                                // NJ * ldq + 64
    double* sX0 = sQ + NJ * ldq + 64;                 // BOXES * n
    double* sEP = sX0 + BOXES * n;
    double* sMX = sEP + BOXES * n;
    double* sBeta = sMX + BOXES * n;                  // BOXES * n

    const int tid = threadIdx.x;
    const int g = tid >> 4, l = tid & 15;
    const int box0 = blockIdx.x * BOXES;

    // Box centre / radius / magnitude (concretize_func.py:548-551,
    // linear.py:607-611, backward_bound.py:739-741).
    for (int t = tid; t < BOXES * n; t += 256) {
        int slot = t / n, i = t - slot * n;
        int b = box0 + slot;
        double x0 = 0.0, ep = 0.0, mx = 0.0;
        if (b < B) {
            double lo = xL[(size_t)b * n + i], hi = xU[(size_t)b * n + i];
            x0 = (hi + lo) / 2.0;
            ep = (hi - lo) / 2.0;
            mx = gh::finite_abs(fmax(-lo, hi));
        }
        sX0[t] = x0;
        sEP[t] = ep;
        sMX[t] = mx;
    }

    for (int s = 0; s < S; ++s) {
        __syncthreads();
        // Stage Q_s, zero rows n..NJ-1.
        const double* Qs = Q + (size_t)s * n * n;
        for (int t = tid; t < NJ * n; t += 256) {
            int j = t / n, i = t - j * n;
            sQ[j * ldq + i] = (j < n) ? Qs[t] : 0.0;
        }
        __syncthreads();

        // ---- 1. IBP of y_s for the lane's rows j = l + 16c.
        double cc[RB][NCH], dv[RB][NCH], mg[RB][NCH];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < NCH; ++c) cc[r][c] = dv[r][c] = mg[r][c] = 0.0;
        for (int i = 0; i < n; ++i) {
            double q[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) q[c] = sQ[(l + 16 * c) * ldq + i];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                int slot = g * RB + r;
                double x0 = sX0[slot * n + i], ep = sEP[slot * n + i], mx = sMX[slot * n + i];
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    double qa = fabs(q[c]);
                    cc[r][c] = fma(q[c], x0, cc[r][c]);
                    dv[r][c] = fma(qa, ep, dv[r][c]);
                    mg[r][c] = fma(qa, mx, mg[r][c]);
                }
            }
        }

        // ---- 2. Elementwise: intermediate bounds, interval of z, planes.
        const double* cs = Cm + (size_t)s * n;
        double alpha[RB][NCH];
        double p_ql[RB], p_qu[RB], p_g[RB], p_own[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            int slot = g * RB + r, b = box0 + slot;
            p_ql[r] = p_qu[r] = p_g[r] = p_own[r] = 0.0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                int j = l + 16 * c;
                alpha[r][c] = 0.0;
                if (j < n) {
                    double e = mg[r][c] * k.kn;
                    double yl = gh::widen((cc[r][c] - dv[r][c]) + 0.0, e, -1.0);
                    double yu = gh::widen((cc[r][c] + dv[r][c]) + 0.0, e, 1.0);
                    double xl = 0.0, xu = 0.0;
                    if (b < B) { xl = xL[(size_t)b * n + j]; xu = xU[(size_t)b * n + j]; }
                    // interval of z_j (bivariate.py:563-597)
                    double r0 = xl * yl, r1 = xl * yu, r2 = xu * yl, r3 = xu * yu;
                    p_ql[r] = p_ql[r] + fmin(fmin(r0, r1), fmin(r2, r3));
                    p_qu[r] = p_qu[r] + fmax(fmax(r0, r1), fmax(r2, r3));
                    // coeff-aware McCormick lower plane (bivariate.py:190-261, 40-75)
                    double a1 = yl, b1 = xl, c1 = (-xl) * yl;
                    double a2 = yu, b2 = xu, c2 = (-xu) * yu;
                    double m1 = a1 * (a1 >= 0 ? xl : xu);
                    m1 = m1 + b1 * (b1 >= 0 ? yl : yu);
                    m1 = m1 + c1;
                    double m2 = a2 * (a2 >= 0 ? xl : xu);
                    m2 = m2 + b2 * (b2 >= 0 ? yl : yu);
                    m2 = m2 + c2;
                    double rr = (m1 >= m2) ? 1.0 : 0.0;
                    double t = (yl - yu) * rr;
                    double al = t + yu;
                    t = (xl - xu) * rr;
                    double be = t + xu;
                    double pp = yu * xu, qq = yl * xl;
                    t = (pp - qq) * rr;
                    double ga = t - pp;
                    alpha[r][c] = al;
                    sBeta[slot * n + j] = be;
                    p_g[r] = p_g[r] + ga;
                    p_own[r] = fma(fabs(be), k.hsn * mg[r][c], p_own[r]);
                }
            }
        }
        __syncthreads();

        // ---- 3. Backward product A = beta^T Q_s for the lane's i = l + 16c.
        double acc[RB][NCH];
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int c = 0; c < NCH; ++c) acc[r][c] = 0.0;
        for (int j = 0; j < n; ++j) {
            double q[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) q[c] = sQ[j * ldq + l + 16 * c];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                double bj = sBeta[(g * RB + r) * n + j];
#pragma unroll
                for (int c = 0; c < NCH; ++c) acc[r][c] = fma(bj, q[c], acc[r][c]);
            }
        }

        // ---- 4. Merges, charges and concretization.
        double csv[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            int i = l + 16 * c;
            csv[c] = (i < n) ? cs[i] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            int slot = g * RB + r, b = box0 + slot;
            double p_lc = 0.0, p_ldv = 0.0, p_lmg = 0.0, p_m1 = 0.0, p_m2 = 0.0;
            double p_base = 0.0, p_et = 0.0, p_ma = 0.0;
            double Af[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                int i = l + 16 * c;
                Af[c] = 0.0;
                if (i < n) {
                    double x0 = sX0[slot * n + i], ep = sEP[slot * n + i], mx = sMX[slot * n + i];
                    double ca = fabs(csv[c]);
                    p_lc = fma(csv[c], x0, p_lc);
                    p_ldv = fma(ca, ep, p_ldv);
                    p_lmg = fma(ca, mx, p_lmg);
                    double acm = csv[c] + alpha[r][c];
                    p_m1 = fma(fabs(acm), mx, p_m1);
                    double a = acm + acc[r][c];
                    Af[c] = a;
                    p_m2 = fma(fabs(a), mx, p_m2);
                    p_base = fma(a, x0, p_base);
                    p_et = fma(fabs(a), ep, p_et);
                    p_ma = fma(fabs(a), mx, p_ma);
                }
            }
            double ql = gh::tree16(p_ql[r]), qu = gh::tree16(p_qu[r]);
            double lc = gh::tree16(p_lc), ldv = gh::tree16(p_ldv), lmg = gh::tree16(p_lmg);
            double gsum = gh::tree16(p_g[r]), own = gh::tree16(p_own[r]);
            double merge1 = gh::tree16(p_m1), merge2 = gh::tree16(p_m2);
            double base = gh::tree16(p_base), et = gh::tree16(p_et), ma = gh::tree16(p_ma);
            if (b >= B) continue;
            double ds = d[s];
            double le = (lmg + fabs(ds)) * k.kn;
            double ll = gh::widen((lc - ldv) + ds, le, -1.0);
            double lu = gh::widen((lc + ldv) + ds, le, 1.0);
            double mag_add = (0.0 + gh::finite_abs(fmax(fabs(ql), fabs(qu))))
                             + gh::finite_abs(fmax(fabs(ll), fabs(lu)));
            double add_err = mag_add * k.k2;
            double a = 0.0;
            a = gh::add_enclosed(a, 0.0 - add_err, -1.0, k.k1);                    // f
            a = gh::add_enclosed(a, 0.0, -1.0, k.k1);                              // q
            double vec = (0.0 + k.hs * lmg) + k.hs * fabs(ds);
            a = gh::add_enclosed(a, ds - vec, -1.0, k.k1);                         // l
            a = gh::add_enclosed(a, gsum, -1.0, k.k1);                             // z
            a = gh::add_enclosed(a, 0.0 - k.h1 * merge1, -1.0, k.k1);              // u
            a = gh::add_enclosed(a, 0.0, -1.0, k.k1);                              // v
            a = gh::add_enclosed(a, 0.0 - (own + k.h1 * merge2), -1.0, k.k1);      // y
            double bias = gh::widen(a, gh::finite_abs(a) * k.kn, -1.0);
            double root = gh::widen(base - et, ma * k.kn, -1.0);
            if (!isfinite(root)) root = -INFINITY;
            double out = gh::add_enclosed(bias, root, -1.0, k.k1);
            if (l == 0) {
                lb[(size_t)b * S + s] = out;
                if (lbias) lbias[(size_t)b * S + s] = a;
            }
            if (lA) {
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    int i = l + 16 * c;
                    if (i < n) lA[((size_t)b * S + s) * n + i] = Af[c];
                }
            }
        }
    }
}

template <int RB, int NCH>
int launch_f64(int B, int n, int S, const double* Q, const double* Cm, const double* d,
               const double* xL, const double* xU, double* lb, double* lA, double* lbias,
               const QcqpConsts& k, hipStream_t stream)
{
    constexpr int BOXES = 16 * RB;
    size_t lds = sizeof(double) * ((size_t)16 * NCH * (n + 1) + 64 + (size_t)4 * BOXES * n);
    if (lds > 160 * 1024) return GH_EUNSUPPORTED;
    auto kern = qcqp_crown_f64_kernel<RB, NCH>;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return GH_ELAUNCH;
    }
    dim3 grid((B + BOXES - 1) / BOXES);
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, stream, B, n, S, Q, Cm, d, xL, xU, lb, lA,
                       lbias, k);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

}  // namespace

extern "C" int gh_qcqp_crown_f64(int B, int n, int S, const double* Q, const double* Cm,
                                 const double* d, const double* xL, const double* xU, double* lb,
                                 double* lA, double* lbias, hipStream_t stream)
{
    if (B < 0 || n <= 0 || S <= 0) return GH_EINVAL;
    if (B == 0) return GH_OK;
    if (!Q || !Cm || !d || !xL || !xU || !lb) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    const double eps = 2.220446049250313080847e-16;
    QcqpConsts k;
    k.kn = (double)(n + 16) * eps * 4.0;
    k.k1 = (double)(1 + 16) * eps * 4.0;
    k.k2 = (double)(2 + 16) * eps * 4.0;
    k.h1 = (double)(1 + 1) * eps / 2.0 * (1.0 + 1e-6);
    k.hs = (double)(S + 1) * eps / 2.0 * (1.0 + 1e-6);
    k.hsn = (double)((long)S * n + 1) * eps / 2.0 * (1.0 + 1e-6);
    int nch = (n + 15) / 16;
    switch (nch) {
    case 1: return launch_f64<2, 1>(B, n, S, Q, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    case 2: return launch_f64<2, 2>(B, n, S, Q, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    case 3: return launch_f64<2, 3>(B, n, S, Q, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    default: return launch_f64<2, 4>(B, n, S, Q, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    }
}
