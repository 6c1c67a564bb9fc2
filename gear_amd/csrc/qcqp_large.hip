// FP32-sound CROWN lower bound of a QCQP of any size (gfx950).
//
// The same relaxation as the FP64 path (first-linear IBP of y_s = Q_s x,
// coefficient-aware McCormick lower planes of z = x * y, backward product
// A = Cm + alpha + Q_s^T beta, box concretization;
// auto_LiRPA/bound_general.py:1275-1593, operators/bivariate.py:190-334,
// backward_bound.py:880-1206, concretize_bounds.py:267-321), arranged for
// large n as two fused MFMA kernels per chunk of boxes:
//   1. boxes are rounded outward to FP32 and written as centre / radius
//      columns (x0, ep) with box in [x0 - ep, x0 + ep];
//   2. forward (large_fwd_kernel): per row s, 64 outputs j and 64 boxes,
//      C = Q32_s x0, DV = |Q32_s| ep and CA = |Q32_s| |x0| on FP32 MFMA from
//      one pass over the Q tile (|Q| formed in registers), then in the same
//      kernel, per (s, b, j) in FP64: y bounds c -/+ dv -/+ E, with E
//      covering the FP32 chains' rounding for any summation order (gamma_n =
//      n u / (1 - n u)), Q rounded to FP32 and the FP64 path's own margin;
//      the plane choice of the FP64 path (elements within a rigorous
//      distance of a tie take the FP64 path's exact choice); beta rounded to
//      FP32, Ainit = fl32(alpha + Cm); per-tile sums of gamma (rounded down),
//      |beta| mg and |Ainit| mx;
//   3. backward (large_bwd_kernel): A_s = Q32_s^T beta_s + Ainit_s on FP32
//      MFMA (chains starting at Ainit), then per-tile box minimum and
//      magnitude sums in FP64;
//   4. large_final_kernel: lb = d + sum gamma + min_box(A x) - E_c - slack,
//      E_c = (gamma_{n+1} + 2.05 u) (sum |beta| mg + sum |Ainit| mx) (the
//      chains' rounding, Q and beta rounded to FP32, Ainit rounded), slack
//      1e-12 of the magnitudes plus twice the FP64 path's enclosure margin,
//      the tiles' sums added in order.
// With the same plane choices, each FP32 plane lies below the FP64 plane on
// the box (looser y bounds move P1 by (yl - yl')(x - xl) >= 0 and P2 by
// (yu' - yu)(x - xu) <= 0), so the bound is not above the FP64 bound.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "gearhip_internal.h"
#include "qcqp_choice.h"

namespace {

constexpr double U32 = 5.9604644775390625e-08;        // 2^-24
constexpr double EPS64 = 2.220446049250313080847e-16;

__device__ __forceinline__ float down_f32(double v)
{
    float f = (float)v;
    return ((double)f > v) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float up_f32(double v)
{
    float f = (float)v;
    return ((double)f < v) ? nextafterf(f, INFINITY) : f;
}

// Wavefront reductions as the xor butterfly (o = 32, 16, 8, 4, 2, 1) on
// permlane swaps and DPP (no LDS); every lane ends with the same value.
__device__ __forceinline__ double wave_sum(double v)
{
    v = v + gh::xor32_f64(v);
    v = v + gh::xor16_f64(v);
    return gh::tree16(v);
}
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, gh::xor32_f64(v));
    v = fmax(v, gh::xor16_f64(v));
    v = fmax(v, gh::dpp_f64<0x128>(v));   // row_ror:8 (lane ^ 8)
    v = fmax(v, gh::dpp_f64<0x124>(v));   // row_ror:4 (a lane holding lane ^ 4's value)
    v = fmax(v, gh::dpp_f64<0x4E>(v));    // quad_perm [2,3,0,1]
    return fmax(v, gh::dpp_f64<0xB1>(v)); // quad_perm [1,0,3,2]
}
// 1 when any lane's flag (0 or 1) is set
__device__ __forceinline__ int wave_any(int v) { return __ballot(v != 0) != 0ull ? 1 : 0; }

// Q32 = fl32(Q), row-major (S, n, n).
__global__ void __launch_bounds__(256)
large_prepare_kernel(size_t total, const double *__restrict__ Q, float *__restrict__ Q32)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < total) Q32[t] = (float)Q[t];
}

// Per box: FP32 centre / radius with box ⊆ [x0 - ep, x0 + ep], and the
// largest magnitude of the outward FP32 box.
__global__ void __launch_bounds__(64)
large_boxes_kernel(int B, int n, const double *__restrict__ xL, const double *__restrict__ xU,
                   float *__restrict__ X0, float *__restrict__ EP, float *__restrict__ XA,
                   double *__restrict__ mxmax)
{
    const int b = blockIdx.x;
    if (b >= B) return;
    double m = 0.0;
    for (int j = threadIdx.x; j < n; j += 64) {
        double l = xL[(size_t)b * n + j], u = xU[(size_t)b * n + j];
        float l32 = down_f32(l), u32 = up_f32(u);
        float x0 = (float)(((double)l32 + (double)u32) * 0.5);
        // both differences are exact in FP64
        float ep = up_f32(fmax((double)u32 - (double)x0, (double)x0 - (double)l32));
        X0[(size_t)b * n + j] = x0;
        EP[(size_t)b * n + j] = ep;
        XA[(size_t)b * n + j] = fabsf(x0);
        m = fmax(m, fmax(fabs((double)l32), fabs((double)u32)));
    }
    m = wave_max(m);
    if (threadIdx.x == 0) mxmax[b] = m;
}

struct LargeConsts {
    double gn;     // gamma_n for the IBP products
    double gn1;    // gamma_{n+1} for the backward product (with Ainit)
    double uq;     // rounding of Q to FP32 (relative to |Q32|)
    double kn64;   // FP64 path's enclosure factor (n + 16) 4 eps
};

// The fused kernels below tile (row s, 64 outputs, 64 boxes) per workgroup
// of four wavefronts; each wavefront owns 16 boxes and four 16 x 16 output
// tiles of v_mfma_f32_16x16x4f32 (lane l: A[l % 16][l / 16],
// B[l / 16][l % 16], D element v at [4 (l / 16) + v][l % 16]), with the
// operand tiles staged in LDS, 32 contraction indices per stage, double
// buffered through registers.
constexpr int TT = 64;   // outputs and boxes per workgroup
constexpr int TK = 32;   // contraction indices per LDS stage

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma16f(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Step 2 + 3 fused, per (s, 64 outputs j, 64 boxes): the three products
// c = Q32 x0, dv = |Q32| ep, ca = |Q32| |x0| on MFMA (each a sequential FMA
// chain over i), then per element in FP64: y bounds, plane choice, beta and
// Ainit (written), and per box the partial sums over this tile's outputs
// [gamma, |gamma|, |beta| mg, |Ainit| mx] (written per tile; the final
// kernel adds the tiles in order).
//
// Errors of the FP32 products (any summation order), with ca = |Q32| |x0|:
//   |c - Q x0|   <= e_c  = (gamma_n + uq) (1 + gamma_n) ca
//   |dv - |Q| ep| <= e_dv = (gamma_n + uq) (1 + 2 gamma_n) dv
// (uq covers Q rounded to FP32).  The planes use y bounds widened by
// E = e_c + e_dv + the FP64 path's margin + an FP64 rounding slack.  The
// choice is decided on the unwidened estimate (c -/+ dv), whose m1 - m2
// differs from the FP64 path's by at most
//   |D_l + D_u| e_c + |D_u - D_l| e_d + FP64 rounding of m1, m2,
// D_l, D_u being the derivatives of m1 - m2 in yl and yu; within that
// distance of a tie the element takes the FP64 path's exact choice.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
large_fwd_kernel(int Bc, int n, int S, int ntj, const float *__restrict__ Q32,
                 const float *__restrict__ X0, const float *__restrict__ EP,
                 const double *__restrict__ Q64,
                 const double *__restrict__ Cm, const double *__restrict__ xL,
                 const double *__restrict__ xU, float *__restrict__ BETA,
                 float *__restrict__ AIN, double *__restrict__ part, int *__restrict__ flag,
                 LargeConsts k)
{
    // 8 wavefronts: wave w owns boxes 16 (w % 4) .. + 15 and outputs
    // 64 (w / 4) .. + 63 of the workgroup's 64 boxes x 128 outputs
    __shared__ float sQ[2][2 * TT][TK + 1];
    __shared__ float sX[2][2][TT][TK + 1];
    const int jt = blockIdx.x, bt = blockIdx.y, s = blockIdx.z;
    const int j0 = jt * 2 * TT, b0 = bt * TT;
    const int t = threadIdx.x, l = t & 63, wv = t >> 6, w = wv & 3, jh = wv >> 2;
    const int lr = l & 15, lq = l >> 4;
    const float *Qs = Q32 + (size_t)s * n * n;
    const int nk = (n + TK - 1) / TK;
    // stage loads (row r, column kk): 4,096 Q entries (8 per thread) and
    // 2,048 of x0 and of ep (4 each)
    float rq[8], rx[2][4];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = t + 512 * e, r = idx >> 5, kk = idx & 31;
            const int kg = k0 + kk, j = j0 + r;
            rq[e] = (j < n && kg < n) ? Qs[(size_t)j * n + kg] : 0.0f;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 512 * e, r = idx >> 5, kk = idx & 31;
            const int kg = k0 + kk, b = b0 + (r & 63);
            const bool okb = b < Bc && kg < n && r < TT;
            rx[0][e] = (okb && e < 4) ? X0[(size_t)b * n + kg] : 0.0f;
            rx[1][e] = okb ? EP[(size_t)b * n + kg] : 0.0f;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = t + 512 * e, r = idx >> 5, kk = idx & 31;
            sQ[buf][r][kk] = rq[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 512 * e, r = idx >> 5, kk = idx & 31;
            if (r < TT) {
                sX[buf][0][r][kk] = rx[0][e];
                sX[buf][1][r][kk] = rx[1][e];
            }
        }
    };
    f4 Dc[4], Dd[4], Dm[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) Dc[q] = Dd[q] = Dm[q] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * TK);
        // lane slot lq takes the stage's contraction indices 8 lq .. 8 lq + 7
        // (any order of the products is covered by the error terms), so each
        // operand row is two 16-byte LDS reads
#pragma unroll
        for (int k4 = 0; k4 < TK / 4; ++k4) {
            const int kk = 4 * k4 + lq;
            const float x0 = sX[cur][0][16 * w + lr][kk];
            const float ep = sX[cur][1][16 * w + lr][kk];
            const float xa = fabsf(x0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float qv = sQ[cur][64 * jh + 16 * q + lr][kk];
                const float qa = fabsf(qv);
                Dc[q] = mfma16f(qv, x0, Dc[q]);
                Dd[q] = mfma16f(qa, ep, Dd[q]);
                Dm[q] = mfma16f(qa, xa, Dm[q]);
            }
        }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    // epilogue: lane (lq, lr) holds box b0 + 16 w + lr, outputs
    // j0 + 64 jh + 16 q + 4 lq + v
    const int b = b0 + 16 * w + lr;
    const bool live = b < Bc;
    const double *lo = xL + (size_t)(live ? b : 0) * n, *hi = xU + (size_t)(live ? b : 0) * n;
    double gs = 0.0, ga = 0.0, bm = 0.0, am = 0.0;
    int tie = 0;
    const double f = k.gn + k.uq;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = j0 + 64 * jh + 16 * q + 4 * lq + v;
            if (!live || i >= n) continue;
            const double c = Dc[q][v], dv = Dd[q][v], ca = Dm[q][v];
            const double e_c = f * (1.0 + k.gn) * ca;
            const double e_dv = f * (1.0 + 2.0 * k.gn) * dv;
            // |Q| mx <= |Q| (|x0| + ep) on the FP32 box, computed values rounded up
            const double mg = (ca + dv) * (1.0 + 2.0 * k.gn) * (1.0 + 4.0 * EPS64);
            const double m64 = 2.0 * k.kn64 * mg;
            const double E = (e_c + e_dv) * (1.0 + 1e-9) + m64 + 1e-12 * (fabs(c) + dv + ca);
            const double yl = (c - dv) - E, yu = (c + dv) + E;
            const double xl = lo[i], xu = hi[i];
            // choice on the unwidened estimate, the FP64 path's selection
            // arithmetic (bivariate.py:205-261)
            const double el = c - dv, eu = c + dv;
            const double X1 = el >= 0 ? xl : xu, Y1 = xl >= 0 ? el : eu;
            const double X2 = eu >= 0 ? xl : xu, Y2 = xu >= 0 ? el : eu;
            double m1 = el * X1;
            m1 = m1 + xl * Y1;
            m1 = m1 + (-xl) * el;
            double m2 = eu * X2;
            m2 = m2 + xu * Y2;
            m2 = m2 + (-xu) * eu;
            const double d1l = X1 - xl + (xl >= 0 ? xl : 0.0), d1u = xl >= 0 ? 0.0 : xl;
            const double d2l = xu >= 0 ? xu : 0.0, d2u = X2 - xu + (xu >= 0 ? 0.0 : xu);
            const double Dl = d1l - d2l, Du = d1u - d2u;
            const double ed = e_dv + m64;
            double tau = fabs(Dl + Du) * e_c + fabs(Du - Dl) * ed;
            const double xm = fmax(fabs(xl), fabs(xu));
            // a sign of yl / yu may differ: |D_l|, |D_u| <= 4 max|x| in every region
            if (fabs(el) <= e_c + ed || fabs(eu) <= e_c + ed)
                tau = fmax(tau, 8.0 * xm * (e_c + ed));
            const double T = 3.0 * xm * (fabs(el) + fabs(eu));
            tau = (tau + 16.0 * EPS64 * T) * (1.0 + 1e-6) + 1e-300;
            bool r1 = m1 >= m2;
            if (fabs(m1 - m2) <= tau || !(fabs(m1 - m2) < INFINITY)) {
                r1 = gh::fp64_choice(n, Q64 + ((size_t)s * n + i) * n, lo, hi, i, k.kn64);
                tie = 1;
            }
            const double alpha = r1 ? yl : yu, beta = r1 ? xl : xu;
            double g = -(beta * alpha);
            g = g - fabs(g) * (2.0 * EPS64) - 1e-300;  // rounded down
            const float b32 = (float)beta;
            const float a32 = (float)(alpha + Cm[(size_t)s * n + i]);
            const size_t o = ((size_t)s * Bc + b) * n + i;
            BETA[o] = b32;
            AIN[o] = a32;
            gs += g;
            ga += fabs(g);
            bm += fabs((double)b32) * mg;
            const double mx32 = fmax(fabs((double)down_f32(xl)), fabs((double)up_f32(xu)));
            am += fabs((double)a32) * mx32;
        }
    }
    // the box's four lanes (lq = 0..3) add their parts
    gs = gs + __shfl_xor(gs, 16);
    ga = ga + __shfl_xor(ga, 16);
    bm = bm + __shfl_xor(bm, 16);
    am = am + __shfl_xor(am, 16);
    gs = gs + __shfl_xor(gs, 32);
    ga = ga + __shfl_xor(ga, 32);
    bm = bm + __shfl_xor(bm, 32);
    am = am + __shfl_xor(am, 32);
    tie = __shfl_xor(tie, 16) | tie;
    tie = __shfl_xor(tie, 32) | tie;
    if (live && lq == 0) {
        // two 64-output tiles per workgroup: tile 2 jt + jh
        double *o = part + (((size_t)s * ntj + 2 * jt + jh) * Bc + b) * 6;
        o[0] = gs;
        o[1] = ga;
        o[2] = bm;
        o[3] = am;
        if (tie) flag[b] = 1;  // a plain store: every writer stores 1
    }
}

// Step 4 + 5 fused, per (s, 64 inputs i, 64 boxes): A = Q32^T beta + Ainit
// on MFMA (the chain starts at Ainit and adds the n products), then per box
// the partial box minimum sum_i A_i (A_i >= 0 ? l_i : u_i) and magnitude
// sum_i |A_i| max(|l_i|, |u_i|) over this tile's inputs, in FP64.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
large_bwd_kernel(int Bc, int n, int S, int ntj, const float *__restrict__ Q32,
                 const float *__restrict__ BETA, const float *__restrict__ AIN,
                 const double *__restrict__ xL, const double *__restrict__ xU,
                 double *__restrict__ part)
{
    // 8 wavefronts: wave w owns boxes 16 (w % 4) .. + 15 and inputs
    // 64 (w / 4) .. + 63 of the workgroup's 64 boxes x 128 inputs
    __shared__ __align__(16) float sQ[2][2 * TT][TK + 4];   // [i][k]: the tile transposed
    __shared__ __align__(16) float sB[2][TT][TK + 4];
    const int it = blockIdx.x, bt = blockIdx.y, s = blockIdx.z;
    const int i0 = it * 2 * TT, b0 = bt * TT;
    const int t = threadIdx.x, l = t & 63, wv = t >> 6, w = wv & 3, ih = wv >> 2;
    const int lr = l & 15, lq = l >> 4;
    const float *Qs = Q32 + (size_t)s * n * n;
    const float *Bs = BETA + (size_t)s * Bc * n;
    const int nk = (n + TK - 1) / TK;
    float rq[8], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            // Q rows k (j) of 128 consecutive columns i
            const int idx = t + 512 * e, kr = idx >> 7, ic = idx & 127;
            const int kg = k0 + kr, ig = i0 + ic;
            rq[e] = (kg < n && ig < n) ? Qs[(size_t)kg * n + ig] : 0.0f;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // beta rows b of 32 consecutive j
            const int idx = t + 512 * e, br = idx >> 5, kk = idx & 31;
            const int bg = b0 + br, kg2 = k0 + kk;
            rb[e] = (bg < Bc && kg2 < n) ? Bs[(size_t)bg * n + kg2] : 0.0f;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = t + 512 * e;
            sQ[buf][idx & 127][idx >> 7] = rq[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 512 * e;
            sB[buf][idx >> 5][idx & 31] = rb[e];
        }
    };
    const int b = b0 + 16 * w + lr;
    const bool live = b < Bc;
    // the chains start at Ainit
    f4 D[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + 64 * ih + 16 * q + 4 * lq + v;
            D[q][v] = (live && i < n) ? AIN[((size_t)s * Bc + b) * n + i] : 0.0f;
        }
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * TK);
        // lane slot lq takes the stage's contraction indices 8 lq .. 8 lq + 7
        const f4 *br = reinterpret_cast<const f4 *>(&sB[cur][16 * w + lr][8 * lq]);
        const f4 bv4[2] = {br[0], br[1]};
        f4 qv4[4][2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const f4 *qr = reinterpret_cast<const f4 *>(&sQ[cur][64 * ih + 16 * q + lr][8 * lq]);
            qv4[q][0] = qr[0];
            qv4[q][1] = qr[1];
        }
#pragma unroll
        for (int k8 = 0; k8 < 8; ++k8) {
            const float bv = bv4[k8 >> 2][k8 & 3];
#pragma unroll
            for (int q = 0; q < 4; ++q) D[q] = mfma16f(qv4[q][k8 >> 2][k8 & 3], bv, D[q]);
        }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    const double *lo = xL + (size_t)(live ? b : 0) * n, *hi = xU + (size_t)(live ? b : 0) * n;
    double mn = 0.0, am = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + 64 * ih + 16 * q + 4 * lq + v;
            if (!live || i >= n) continue;
            const double a = D[q][v];
            const double lv = lo[i], hv = hi[i];
            mn += a * (a >= 0 ? lv : hv);
            am += fabs(a) * fmax(fabs(lv), fabs(hv));
        }
    mn = mn + __shfl_xor(mn, 16);
    am = am + __shfl_xor(am, 16);
    mn = mn + __shfl_xor(mn, 32);
    am = am + __shfl_xor(am, 32);
    if (live && lq == 0) {
        double *o = part + (((size_t)s * ntj + 2 * it + ih) * Bc + b) * 6;
        o[4] = mn;
        o[5] = am;
    }
}

// Step 5: the tiles' partial sums, in tile order, and the error terms.
__global__ void __launch_bounds__(256)
large_final_kernel(int Bc, int S, int ntj, const double *__restrict__ part,
                   const double *__restrict__ d, double *__restrict__ lb, LargeConsts k)
{
    const size_t sb = (size_t)blockIdx.x * 256 + threadIdx.x;  // s * Bc + b
    if (sb >= (size_t)S * Bc) return;
    const int s = (int)(sb / Bc), b = (int)(sb - (size_t)s * Bc);
    double gs = 0.0, ga = 0.0, bm = 0.0, aim = 0.0, mn = 0.0, am = 0.0;
    for (int jt = 0; jt < ntj; ++jt) {
        const double *o = part + (((size_t)s * ntj + jt) * Bc + b) * 6;
        gs += o[0];
        ga += o[1];
        bm += o[2];
        aim += o[3];
        mn += o[4];
        am += o[5];
    }
    const double dd = d[s];
    const double f = k.gn1 + 2.05 * U32;
    const double Ec = (f * bm + f * aim) * (1.0 + 1e-9);
    const double mag = am + ga + fabs(dd) + fabs(mn) + fabs(gs);
    const double slack = 1e-12 * mag + 2.0 * k.kn64 * mag + 1e-300;
    double v = ((dd + gs) + mn) - Ec - slack;
    if (!(v == v)) v = -INFINITY;
    lb[(size_t)b * S + s] = v;
}

__global__ void zero_i32_kernel(int *p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

struct LargePlanLayout {
    size_t q32, total;
};
LargePlanLayout plan_layout(int n, int S)
{
    LargePlanLayout L;
    L.q32 = 0;
    L.total = align256((size_t)S * n * n * sizeof(float));
    return L;
}

// per-box workspace: X0, EP, |x0| (n floats each), BETA and Ainit (S n
// floats each), the tiles' partial sums (6 doubles per (s, tile)) and the
// box's largest magnitude (1 double)
size_t per_box_bytes(int n, int S)
{
    const size_t ntj = 2 * (size_t)((n + 2 * TT - 1) / (2 * TT));
    return (size_t)3 * n * 4 + (size_t)2 * S * n * 4 + (size_t)6 * S * ntj * 8 + 8;
}

}  // namespace

extern "C" size_t gh_qcqp_large_plan_bytes(int n, int S)
{
    if (n < 1 || S < 1) return 0;
    return plan_layout(n, S).total;
}

extern "C" int gh_qcqp_large_prepare(int n, int S, const double *Q, void *plan, hipStream_t stream)
{
    if (n < 1 || S < 1 || !Q || !plan) return GH_EINVAL;
    const size_t total = (size_t)S * n * n;
    hipLaunchKernelGGL(large_prepare_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, total, Q, static_cast<float *>(plan));
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" size_t gh_qcqp_large_workspace_bytes(int n, int S, int B)
{
    if (n < 1 || S < 1 || B < 1) return 0;
    // alignment padding for the eight arrays of a chunk
    return per_box_bytes(n, S) * (size_t)B + 8 * 256;
}

extern "C" int gh_qcqp_crown_f32s_large(int B, int n, int S, const void *plan, const double *Q64,
                                        const double *Cm, const double *d, const double *xL,
                                        const double *xU, double *lb, int32_t *flag, void *ws,
                                        size_t ws_bytes, hipStream_t stream)
{
    if (B == 0) return GH_OK;
    if (B < 0 || n < 1 || S < 1 || !plan || !Q64 || !Cm || !d || !xL || !xU || !lb || !flag || !ws)
        return GH_EINVAL;
    const size_t pb = per_box_bytes(n, S);
    if (ws_bytes < pb + 8 * 256) return GH_EINVAL;
    const int chunk = (int)std::min<size_t>((size_t)B, (ws_bytes - 8 * 256) / pb);
    const LargePlanLayout L = plan_layout(n, S);
    const char *pp = static_cast<const char *>(plan);
    const float *Q32 = reinterpret_cast<const float *>(pp + L.q32);
    LargeConsts k;
    const double u = U32;
    k.gn = n * u / (1.0 - n * u);
    k.gn1 = (n + 1) * u / (1.0 - (n + 1) * u);
    k.uq = u * (1.0 + u) * 1.01;
    k.kn64 = (double)(n + 16) * EPS64 * 4.0;
    // 64-output tiles, an even number (the forward kernel takes them in pairs)
    const int ntj = 2 * ((n + 2 * TT - 1) / (2 * TT));
    hipLaunchKernelGGL(zero_i32_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, flag, B);
    for (int b0 = 0; b0 < B; b0 += chunk) {
        const int Bc = std::min(chunk, B - b0);
        char *w = static_cast<char *>(ws);
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char *r = w + off;
            off = align256(off + bytes);
            return r;
        };
        float *X0 = reinterpret_cast<float *>(take((size_t)Bc * n * 4));
        float *EP = reinterpret_cast<float *>(take((size_t)2 * Bc * n * 4));
        float *XA = EP + (size_t)Bc * n;
        float *BETA = reinterpret_cast<float *>(take((size_t)S * Bc * n * 4));
        float *AIN = reinterpret_cast<float *>(take((size_t)S * Bc * n * 4));
        double *part = reinterpret_cast<double *>(take((size_t)S * ntj * Bc * 6 * 8));
        double *mxm = reinterpret_cast<double *>(take((size_t)Bc * 8));
        const double *cxL = xL + (size_t)b0 * n, *cxU = xU + (size_t)b0 * n;
        hipLaunchKernelGGL(large_boxes_kernel, dim3(Bc), dim3(64), 0, stream, Bc, n, cxL, cxU, X0,
                           EP, XA, mxm);
        const dim3 gfwd((unsigned)(ntj / 2), (unsigned)((Bc + TT - 1) / TT), (unsigned)S);
        hipLaunchKernelGGL(large_fwd_kernel, gfwd, dim3(512), 0, stream, Bc, n, S, ntj, Q32, X0,
                           EP, Q64, Cm, cxL, cxU, BETA, AIN, part, flag + b0, k);
        hipLaunchKernelGGL(large_bwd_kernel, gfwd, dim3(512), 0, stream, Bc, n, S, ntj, Q32, BETA,
                           AIN, cxL, cxU, part);
        hipLaunchKernelGGL(large_final_kernel, dim3((unsigned)(((size_t)S * Bc + 255) / 256)),
                           dim3(256), 0, stream, Bc, S, ntj, part, d, lb + (size_t)b0 * S, k);
    }
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
