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
// Mapping.  One wavefront per 16 boxes (one block each, so there is no
// block-level synchronisation).  Q is packed once per problem into MFMA
// operand order (gh_qcqp_pack_f64), so every B operand is one coalesced
// load that stays L2-resident across the batch.  Per row s:
//   1. IBP of y_s = Q_s x on FP64 MFMA (16x16x4): rows = boxes, columns =
//      j, contraction over i; centre, radius and magnitude share Q;
//   2. McCormick plane choice and coefficients and the interval of z on the
//      MFMA output layout (lane (q, r), item v <-> box q + 4v, j = 16c + r);
//   3. beta -> LDS, backward A = beta^T Q_s on MFMA (contraction over j);
//   4. merges, A67 charges and concretization as 16-lane partials (column
//      r of every 16-wide tile) plus an xor tree over the 16 lanes of a box.
// An FP64 MFMA k-step is the FMA chain over its 4 k in order, so each GEMM
// output is the sequential FMA chain the oracle (oracle/crown_qcqp.c)
// prescribes; with the same reduction tree the result is bit-identical.
#include "common.h"
#include "gearhip_internal.h"

namespace {

struct QcqpConsts {
    double kn, k1, k2, h1, hs, hsn;
};

typedef double d4 __attribute__((ext_vector_type(4)));

// D = C + A * B on a 16x16x4 FP64 tile.  Operand layout (lane l): A holds
// A[l % 16][l / 16], B holds B[l / 16][l % 16]; D element v of lane l is
// D[l / 16 + 4 v][l % 16].  Each output is the FMA chain over k = 0..3 in
// order, so a K loop of these is the sequential FMA chain over K.
__device__ __forceinline__ d4 mfma16(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Packed Q operands (built once per problem by gh_qcqp_pack_f64):
//   QB[s][c][t][l] = Q[s][16c + l%16][4t + l/16]   IBP B operand (Q^T tiles)
//   QA[s][c][t][l] = Q[s][4t + l/16][16c + l%16]   backward B operand
// zero outside 0 <= row, col < n; c < NCH, t < KT = ceil(n / 4).  Each MFMA
// B operand is then one coalesced 512-byte load per wave.
__global__ void qcqp_pack_kernel(int n, int S, int nch, int kt, const double* __restrict__ Q,
                                 double* __restrict__ QB, double* __restrict__ QA)
{
    size_t per = (size_t)nch * kt * 64;
    size_t total = (size_t)S * per;
    size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= total) return;
    int s = (int)(u / per);
    size_t r = u - (size_t)s * per;
    int c = (int)(r / ((size_t)kt * 64));
    r -= (size_t)c * kt * 64;
    int t = (int)(r / 64), l = (int)(r - (size_t)t * 64);
    const double* Qs = Q + (size_t)s * n * n;
    int j = 16 * c + (l & 15), i = 4 * t + (l >> 4);
    QB[u] = (j < n && i < n) ? Qs[(size_t)j * n + i] : 0.0;
    int j2 = 4 * t + (l >> 4), i2 = 16 * c + (l & 15);
    QA[u] = (j2 < n && i2 < n) ? Qs[(size_t)j2 * n + i2] : 0.0;
}

// One wavefront per 16 boxes; no block-level synchronisation.  LDS per wave:
// beta and alpha (16 x K4 each, zero padded); box rows are read from global
// memory (L1/L2 resident for the wave's lifetime).
template <int NCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void qcqp_crown_f64_kernel(
    int B, int n, int S, int kt, const double* __restrict__ QB, const double* __restrict__ QA,
    const double* __restrict__ Cm, const double* __restrict__ d, const double* __restrict__ xL,
    const double* __restrict__ xU, double* __restrict__ lb, double* __restrict__ lA,
    double* __restrict__ lbias, QcqpConsts k)
{
    extern __shared__ double smem[];
    const int K4 = 4 * kt;
    double* sBeta = smem;              // 16 * K4
    double* sAlpha = sBeta + 16 * K4;  // 16 * K4

    const int l = threadIdx.x;
    const int lr = l & 15, lq = l >> 4;
    const int box0 = blockIdx.x * 16;

    for (int t = l; t < 16 * K4; t += 64) sBeta[t] = 0.0;
    __syncthreads();
    // Box rows this lane reads: the A-operand box (lr) and, on the output
    // layout, boxes lq + 4v.  Rows past B read box B-1 (never stored).
    const int bA = min(box0 + lr, B - 1);
    const double* aL = xL + (size_t)bA * n;
    const double* aU = xU + (size_t)bA * n;

    const size_t rowstride = (size_t)NCH * kt * 64;
    for (int s = 0; s < S; ++s) {
        const double* qb = QB + (size_t)s * rowstride + l;
        const double* qa = QA + (size_t)s * rowstride + l;

        // ---- 1. IBP of y_s on MFMA: rows = the 16 boxes, columns = j,
        // contraction over i (concretize_func.py:548-551, linear.py:607-653,
        // backward_bound.py:739-741 for centre / radius / magnitude).
        d4 Dc[NCH], Dd[NCH], Dm[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) Dc[c] = Dd[c] = Dm[c] = (d4){0.0, 0.0, 0.0, 0.0};
        double qn[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) qn[c] = qb[(size_t)c * kt * 64];
        for (int t = 0; t < kt; ++t) {
            double q[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) q[c] = qn[c];
            if (t + 1 < kt) {
#pragma unroll
                for (int c = 0; c < NCH; ++c) qn[c] = qb[((size_t)c * kt + t + 1) * 64];
            }
            int i = 4 * t + lq;
            double lo = 0.0, hi = 0.0;
            if (i < n) { lo = aL[i]; hi = aU[i]; }
            double x0 = (hi + lo) / 2.0;
            double ep = (hi - lo) / 2.0;
            double mx = gh::finite_abs(fmax(-lo, hi));
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                double qa_ = fabs(q[c]);
                Dc[c] = mfma16(x0, q[c], Dc[c]);
                Dd[c] = mfma16(ep, qa_, Dd[c]);
                Dm[c] = mfma16(mx, qa_, Dm[c]);
            }
        }

        // ---- 2. Elementwise on the MFMA output layout: item v of lane
        // (lq, lr) is box lq + 4v, column j = 16c + lr.
        const double* cs = Cm + (size_t)s * n;
        double p_ql[4], p_qu[4], p_g[4], p_own[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int slot = lq + 4 * v;
            const int bo = min(box0 + slot, B - 1);
            p_ql[v] = p_qu[v] = p_g[v] = p_own[v] = 0.0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int j = 16 * c + lr;
                if (j < n) {
                    double cc = Dc[c][v], dv = Dd[c][v], mg = Dm[c][v];
                    double e = mg * k.kn;
                    double yl = gh::widen((cc - dv) + 0.0, e, -1.0);
                    double yu = gh::widen((cc + dv) + 0.0, e, 1.0);
                    double xl = xL[(size_t)bo * n + j], xu = xU[(size_t)bo * n + j];
                    // interval of z_j (bivariate.py:563-597)
                    double r0 = xl * yl, r1 = xl * yu, r2 = xu * yl, r3 = xu * yu;
                    p_ql[v] = p_ql[v] + fmin(fmin(r0, r1), fmin(r2, r3));
                    p_qu[v] = p_qu[v] + fmax(fmax(r0, r1), fmax(r2, r3));
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
                    sAlpha[slot * K4 + j] = al;
                    sBeta[slot * K4 + j] = be;
                    p_g[v] = p_g[v] + ga;
                    p_own[v] = fma(fabs(be), k.hsn * mg, p_own[v]);
                }
            }
        }
        __syncthreads();   // one wave per block: orders the beta writes and reads

        // ---- 3. Backward product A = beta^T Q_s on MFMA: rows = boxes,
        // columns = i, contraction over j.
        d4 Da[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) Da[c] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < NCH; ++c) qn[c] = qa[(size_t)c * kt * 64];
        for (int t = 0; t < kt; ++t) {
            double q[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) q[c] = qn[c];
            if (t + 1 < kt) {
#pragma unroll
                for (int c = 0; c < NCH; ++c) qn[c] = qa[((size_t)c * kt + t + 1) * 64];
            }
            double a = sBeta[lr * K4 + 4 * t + lq];
#pragma unroll
            for (int c = 0; c < NCH; ++c) Da[c] = mfma16(a, q[c], Da[c]);
        }

        // ---- 4. Merges, charges and concretization per box.
        double csv[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            int i = 16 * c + lr;
            csv[c] = (i < n) ? cs[i] : 0.0;
        }
        const double ds = d[s];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int slot = lq + 4 * v, b = box0 + slot;
            const int bo = min(b, B - 1);
            double p_lc = 0.0, p_ldv = 0.0, p_lmg = 0.0, p_m1 = 0.0, p_m2 = 0.0;
            double p_base = 0.0, p_et = 0.0, p_ma = 0.0;
            double Af[NCH];
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                int i = 16 * c + lr;
                Af[c] = 0.0;
                if (i < n) {
                    double lo = xL[(size_t)bo * n + i], hi = xU[(size_t)bo * n + i];
                    double x0 = (hi + lo) / 2.0, ep = (hi - lo) / 2.0;
                    double mx = gh::finite_abs(fmax(-lo, hi));
                    double ca = fabs(csv[c]);
                    p_lc = fma(csv[c], x0, p_lc);
                    p_ldv = fma(ca, ep, p_ldv);
                    p_lmg = fma(ca, mx, p_lmg);
                    double acm = csv[c] + sAlpha[slot * K4 + i];
                    p_m1 = fma(fabs(acm), mx, p_m1);
                    double a = acm + Da[c][v];
                    Af[c] = a;
                    p_m2 = fma(fabs(a), mx, p_m2);
                    p_base = fma(a, x0, p_base);
                    p_et = fma(fabs(a), ep, p_et);
                    p_ma = fma(fabs(a), mx, p_ma);
                }
            }
            double ql = gh::tree16(p_ql[v]), qu = gh::tree16(p_qu[v]);
            double lc = gh::tree16(p_lc), ldv = gh::tree16(p_ldv), lmg = gh::tree16(p_lmg);
            double gsum = gh::tree16(p_g[v]), own = gh::tree16(p_own[v]);
            double merge1 = gh::tree16(p_m1), merge2 = gh::tree16(p_m2);
            double base = gh::tree16(p_base), et = gh::tree16(p_et), ma = gh::tree16(p_ma);
            if (b >= B) continue;
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
            if (lr == 0) {
                lb[(size_t)b * S + s] = out;
                if (lbias) lbias[(size_t)b * S + s] = a;
            }
            if (lA) {
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    int i = 16 * c + lr;
                    if (i < n) lA[((size_t)b * S + s) * n + i] = Af[c];
                }
            }
        }
        __syncthreads();   // beta of this row is consumed before the next row writes it
    }
}

template <int NCH>
int launch_f64(int B, int n, int S, const double* QB, const double* QA, const double* Cm,
               const double* d, const double* xL, const double* xU, double* lb, double* lA,
               double* lbias, const QcqpConsts& k, hipStream_t stream)
{
    const int kt = (n + 3) / 4;
    size_t lds = sizeof(double) * (size_t)2 * 16 * 4 * kt;
    dim3 grid((B + 15) / 16);
    hipLaunchKernelGGL(qcqp_crown_f64_kernel<NCH>, grid, dim3(64), lds, stream, B, n, S, kt, QB,
                       QA, Cm, d, xL, xU, lb, lA, lbias, k);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

inline int nch_of(int n) { return (n + 15) / 16; }

}  // namespace

extern "C" size_t gh_qcqp_pack_bytes(int n, int S)
{
    if (n <= 0 || S <= 0 || n > 64) return 0;
    return sizeof(double) * 2 * (size_t)S * nch_of(n) * ((n + 3) / 4) * 64;
}

extern "C" int gh_qcqp_pack_f64(int n, int S, const double* Q, void* packed, hipStream_t stream)
{
    if (n <= 0 || S <= 0 || !Q || !packed) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    int nch = nch_of(n), kt = (n + 3) / 4;
    size_t per = (size_t)S * nch * kt * 64;
    double* QB = (double*)packed;
    double* QA = QB + per;
    hipLaunchKernelGGL(qcqp_pack_kernel, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, stream,
                       n, S, nch, kt, Q, QB, QA);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" int gh_qcqp_crown_f64(int B, int n, int S, const void* packed, const double* Cm,
                                 const double* d, const double* xL, const double* xU, double* lb,
                                 double* lA, double* lbias, hipStream_t stream)
{
    if (B < 0 || n <= 0 || S <= 0) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    if (B == 0) return GH_OK;
    if (!packed || !Cm || !d || !xL || !xU || !lb) return GH_EINVAL;
    const double eps = 2.220446049250313080847e-16;
    QcqpConsts k;
    k.kn = (double)(n + 16) * eps * 4.0;
    k.k1 = (double)(1 + 16) * eps * 4.0;
    k.k2 = (double)(2 + 16) * eps * 4.0;
    k.h1 = (double)(1 + 1) * eps / 2.0 * (1.0 + 1e-6);
    k.hs = (double)(S + 1) * eps / 2.0 * (1.0 + 1e-6);
    k.hsn = (double)((long)S * n + 1) * eps / 2.0 * (1.0 + 1e-6);
    int nch = nch_of(n), kt = (n + 3) / 4;
    const double* QB = (const double*)packed;
    const double* QA = QB + (size_t)S * nch * kt * 64;
    switch (nch) {
    case 1: return launch_f64<1>(B, n, S, QB, QA, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    case 2: return launch_f64<2>(B, n, S, QB, QA, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    case 3: return launch_f64<3>(B, n, S, QB, QA, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    default: return launch_f64<4>(B, n, S, QB, QA, Cm, d, xL, xU, lb, lA, lbias, k, stream);
    }
}
