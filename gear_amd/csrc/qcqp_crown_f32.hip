// FP32-sound CROWN lower bound of a QCQP over a batch of boxes (gfx950).
//
// The same relaxation as the FP64 kernel (qcqp_crown.hip: first-linear IBP of
// y_s = Q_s x, coeff-aware McCormick lower planes of z = x * y, backward
// product, box concretization), evaluated in FP32 on MFMA with a rigorous
// outward error term instead of the reference's FP64 rounding charges.  With
// u = 2^-24 (FP32 MFMA steps and VALU operations round to nearest; the
// probe behind this was an in-bounds comparison on the MI355X) and
// g = (n + 1) u, per row:
//   * boxes are rounded outward to FP32; Q, Cm, d to nearest (|Q64 - Q32| <=
//     1.0000001 u |Q32|);
//   * y bounds: c -/+ dv -/+ E_y with E_y = KY (mg + dv), KY covering the FMA
//     chains (2g), the data rounding (u), the centre / radius rounding (4u)
//     and the final subtractions (4u), times 1.01;
//   * lower planes P1 = yl x + xl y - xl yl or P2 = yu x + xu y - xu yu,
//     constants rounded down, so each plane is below x * y on the box;
//   * the concretized bound is lowered by E_c, the sum of the backward
//     product's rounding and data error (KB sum_j |beta_j| mg_j), the merge
//     roundings (3u sum_i (|Cm_i| + |alpha_i|) mx_i), the concretization
//     chains (KC sum_i |A_i| mx_i), the bias chain (g sum_j |gamma_j| + u |d|)
//     and the last additions (4u (|bias| + |root|)), times 1.01.
// Since yl32 <= yl64 and yu32 >= yu64, each FP32 plane lies below the FP64
// plane of the same choice on the box, so the FP32 bound is not above the
// FP64 one when both pick the same planes (a relative slack of 1e-12 on the
// magnitudes covers the FP64 path's own rounding).  The magnitude term is
// mg = ca + dv with ca = |Q32| |x0| (a third MFMA chain), which bounds
// |Q32| mx on the FP32 box.  The plane choice is decided on the unwidened
// estimate c -/+ dv; where its m1 - m2 is within the rigorous distance
// |D_l + D_u| e_c + |D_u - D_l| e_d (+ rounding) of a tie, D_l, D_u being
// the derivatives of m1 - m2 in yl and yu, e_c and e_d the errors of the
// centre and radius, the lane recomputes that element's y interval with the
// FP64 path's exact FMA chains and takes the FP64 choice, so the choices
// always agree.
#include "common.h"
#include "gearhip_internal.h"
#include "qcqp_choice.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// D = C + A * B on a 16x16x4 FP32 tile.  Lane l: A holds A[l % 16][l / 16],
// B holds B[l / 16][l % 16]; D element v of lane l is D[4 (l / 16) + v][l % 16].
__device__ __forceinline__ f4 mfma16f(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// FP32 packing with the MFMA row permutation R -> 16c + 4 (R % 4) + R / 4, so
// item v of lane (q, r) in tile c is index q + 4 (4c + v), as in the FP64
// kernel:
//   QB[s][c][t][l] = Q[s][16c + 4 (l%16 % 4) + (l%16) / 4][4t + l/16]
//   QA[s][c][t][l] = Q[s][4t + l/16][16c + 4 (l%16 % 4) + (l%16) / 4]
__global__ void qcqp_pack_f32_kernel(int n, int S, int nch, int kt, const double* __restrict__ Q,
                                     float* __restrict__ QB, float* __restrict__ QA)
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
    int R = l & 15, K = l >> 4;
    int row = 16 * c + 4 * (R & 3) + (R >> 2);
    int j = row, i = 4 * t + K;
    QB[u] = (j < n && i < n) ? (float)Qs[(size_t)j * n + i] : 0.0f;
    int j2 = 4 * t + K, i2 = row;
    QA[u] = (j2 < n && i2 < n) ? (float)Qs[(size_t)j2 * n + i2] : 0.0f;
}

struct F32Consts {
    float ky;      // y-bound error factor
    float kb;      // backward product factor
    float kc;      // concretization factor
    float g;       // (n + 1) u
    float u;       // 2^-24
    float tiny;    // absolute slack for underflow
    float f64slack;  // covers the FP64 path's own rounding (relative)
    double kn64;   // FP64 y-bound margin factor (n + 16) * 4 * eps
};

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

__device__ __forceinline__ float tree4x4f(const float p[4])
{
    float q0 = p[0] + p[2], q1 = p[1] + p[3];
    float r = q0 + q1;
    r = r + __builtin_bit_cast(float, (int)__builtin_amdgcn_permlane32_swap(
                __builtin_bit_cast(unsigned, r), __builtin_bit_cast(unsigned, r), false,
                false)[(threadIdx.x & 63) < 32 ? 1 : 0]);
    r = r + __builtin_bit_cast(float, (int)__builtin_amdgcn_permlane16_swap(
                __builtin_bit_cast(unsigned, r), __builtin_bit_cast(unsigned, r), false,
                false)[((threadIdx.x >> 4) & 1) == 0 ? 1 : 0]);
    return r;
}

// STAGED: 8 wavefronts (128 boxes) per workgroup share the current row's
// FP32 Q tiles in LDS while the next row's are copied in (double buffer);
// the arithmetic is the same either way.
template <int KT, bool STAGED>
__global__ __launch_bounds__(STAGED ? 512 : 64) __attribute__((amdgpu_waves_per_eu(2))) void qcqp_crown_f32s_kernel(
    int B, int n, int S, const float* __restrict__ QB, const float* __restrict__ QA,
    const double* __restrict__ Q64, const double* __restrict__ Cm, const double* __restrict__ d,
    const double* __restrict__ xL, const double* __restrict__ xU,
    const int64_t* __restrict__ rows, double* __restrict__ lb, int* __restrict__ flag,
    double* __restrict__ lA, double* __restrict__ lbias, int PR, F32Consts k, int rpb)
{
    constexpr int NCH = (KT + 3) / 4;
    constexpr int NM = KT;
    constexpr int G = 2;
    constexpr int W = STAGED ? 8 : 1;
    extern __shared__ float qsf[];       // STAGED: 2 x [QB row | QA row]
    const int l = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int lr = l & 15, lq = l >> 4;
    const int box = (blockIdx.x * W + wv) * 16 + lr;
    const int bo = min(box, B - 1);
    const bool live = box < B;
    // the box's row of xL / xU: rows[bo] when an index is given
    const size_t xr = rows ? (size_t)rows[bo] : (size_t)bo;

    float xl[NM], xu[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        int i = lq + 4 * m;
        xl[m] = (i < n) ? down_f32(xL[xr * n + i]) : 0.0f;
        xu[m] = (i < n) ? up_f32(xU[xr * n + i]) : 0.0f;
    }
    int ambiguous = 0;

    const size_t rowstride = (size_t)NCH * KT * 64;
    // this workgroup's rows (blockIdx.y): the spec rows are independent
    const int s0 = blockIdx.y * rpb, s1 = min(S, s0 + rpb);
    if constexpr (STAGED) {
        gh::stage_row<float, W>(QB, QA, rowstride, s0, qsf + (s0 & 1) * 2 * rowstride, wv, l);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    for (int s = s0; s < s1; ++s) {
        const float* qb;
        const float* qa;
        if constexpr (STAGED) {
            if (s + 1 < s1)
                gh::stage_row<float, W>(QB, QA, rowstride, s + 1,
                                        qsf + ((s + 1) & 1) * 2 * rowstride, wv, l);
            const float* buf = qsf + (s & 1) * 2 * rowstride;
            qb = buf + l;
            qa = buf + rowstride + l;
        } else {
            qb = QB + (size_t)s * rowstride + l;
            qa = QA + (size_t)s * rowstride + l;
        }
        const double* cs = Cm + (size_t)s * n;

        float acm[NM], beta[NM];
        float p_g[4], p_ga[4], p_own[4], p_m1[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) p_g[v] = p_ga[v] = p_own[v] = p_m1[v] = 0.0f;

#pragma unroll
        for (int c0 = 0; c0 < NCH; c0 += G) {
            f4 Dc[G], Dd[G], Dm[G];
#pragma unroll
            for (int g = 0; g < G; ++g)
                Dc[g] = Dd[g] = Dm[g] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                float x0 = (xu[t] + xl[t]) * 0.5f;
                float ep = (xu[t] - xl[t]) * 0.5f;
                ep = ep + (ep + fabsf(x0)) * (4.0f * k.u);   // covers both roundings
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (c0 + g < NCH) {
                        float q = qb[((size_t)(c0 + g) * KT + t) * 64];
                        float qa_ = fabsf(q);
                        Dc[g] = mfma16f(q, x0, Dc[g]);
                        Dd[g] = mfma16f(qa_, ep, Dd[g]);
                        Dm[g] = mfma16f(qa_, fabsf(x0), Dm[g]);
                    }
                }
                if (t % 4 == 3) __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = c0 + g, m = 4 * c + v;
                    if (c < NCH && m < NM) {
                        const int j = lq + 4 * m;
                        acm[m] = 0.0f;
                        beta[m] = 0.0f;
                        if (j < n) {
                            // ca = |Q32| |x0|; |Q32| mx <= |Q32| (|x0| + ep) = ca + dv
                            // on the FP32 box, the computed sums rounded up
                            float cc = Dc[g][v], dv = Dd[g][v], ca = Dm[g][v];
                            float mg = (ca + dv) * (1.0f + 2.0f * k.g + 2.0f * k.u);
                            float ey = k.ky * (mg + dv) + k.tiny;
                            float yl = (cc - dv) - ey, yu = (cc + dv) + ey;
                            float xlj = xl[m], xuj = xu[m];
                            // the choice is decided on the unwidened estimate, with
                            // the FP64 path's selection arithmetic (bivariate.py:205-261)
                            float el = cc - dv, eu = cc + dv;
                            float X1 = el >= 0 ? xlj : xuj, Y1 = xlj >= 0 ? el : eu;
                            float X2 = eu >= 0 ? xlj : xuj, Y2 = xuj >= 0 ? el : eu;
                            float m1 = el * X1 + xlj * Y1 + (-xlj) * el;
                            float m2 = eu * X2 + xuj * Y2 + (-xuj) * eu;
                            // distance of the estimate's m1 - m2 from the FP64 path's:
                            // errors of the centre (e_c, from the chain and the box
                            // rounding, proportional to ca) and of the radius (e_d,
                            // proportional to dv, with the FP64 margin), through the
                            // derivatives D_l, D_u of m1 - m2 in yl, yu
                            float mxj = fmaxf(-xlj, xuj);
                            float e_c = ((k.g + 4.0f * k.u) * (1.0f + 2.0f * k.g) * ca
                                         + 4.0f * k.u * dv) * 1.01f;
                            float e_d = ((k.g + 8.0f * k.u) * (1.0f + 2.0f * k.g) * dv
                                         + 8.0f * k.u * ca + 2.0f * (float)k.kn64 * mg) * 1.01f;
                            float d1l = X1 - xlj + (xlj >= 0 ? xlj : 0.0f);
                            float d1u = xlj >= 0 ? 0.0f : xlj;
                            float d2l = xuj >= 0 ? xuj : 0.0f;
                            float d2u = X2 - xuj + (xuj >= 0 ? 0.0f : xuj);
                            float Dl = d1l - d2l, Du = d1u - d2u;
                            float tau = fabsf(Dl + Du) * e_c + fabsf(Du - Dl) * e_d;
                            float xm = fmaxf(fabsf(xlj), fabsf(xuj));
                            if (fabsf(el) <= e_c + e_d || fabsf(eu) <= e_c + e_d)
                                tau = fmaxf(tau, 8.0f * xm * (e_c + e_d));
                            float T = 3.0f * xm * (fabsf(el) + fabsf(eu));
                            tau = (tau + 16.0f * k.u * T) * 1.01f + k.tiny;
                            bool r1 = m1 >= m2;
                            if (!(fabsf(m1 - m2) > tau)) {
                                // near a tie: take the FP64 path's choice
                                ambiguous = 1;
                                r1 = gh::fp64_choice(n, Q64 + ((size_t)s * n + j) * n,
                                                 xL + xr * n, xU + xr * n, j,
                                                 k.kn64);
                            }
                            float al = r1 ? yl : yu;
                            float be = r1 ? xlj : xuj;
                            float ga = r1 ? -(xlj * yl) : -(xuj * yu);
                            ga = ga - fabsf(ga) * (2.0f * k.u) - k.tiny;   // rounded down
                            float cm = (float)cs[j];
                            acm[m] = cm + al;
                            beta[m] = be;
                            p_g[v] = p_g[v] + ga;
                            p_ga[v] = p_ga[v] + fabsf(ga);
                            p_own[v] = fmaf(fabsf(be), mg, p_own[v]);
                            p_m1[v] = fmaf(fabsf(cm) + fabsf(al), mxj, p_m1[v]);
                        }
                    }
                }
            }
        }
        const float gsum = tree4x4f(p_g), gabs = tree4x4f(p_ga);
        const float own = tree4x4f(p_own), m1s = tree4x4f(p_m1);

        f4 Da[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) Da[c] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < KT; ++t) {
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                Da[c] = mfma16f(qa[((size_t)c * KT + t) * 64], beta[t], Da[c]);
            if (t % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }

        float p_base[4], p_et[4], p_ma[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) p_base[v] = p_et[v] = p_ma[v] = 0.0f;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 4 * c + v, i = lq + 4 * m;
                if (m < NM && i < n) {
                    float lo = xl[m], hi = xu[m];
                    float x0 = (hi + lo) * 0.5f;
                    float ep = (hi - lo) * 0.5f;
                    ep = ep + (ep + fabsf(x0)) * (4.0f * k.u);
                    float mx = fmaxf(-lo, hi);
                    float a = acm[m] + Da[c][v];
                    if (lA && live && s < PR) lA[((size_t)box * PR + s) * n + i] = (double)a;
                    p_base[v] = fmaf(a, x0, p_base[v]);
                    p_et[v] = fmaf(fabsf(a), ep, p_et[v]);
                    p_ma[v] = fmaf(fabsf(a), mx, p_ma[v]);
                }
            }
        }
        const float base = tree4x4f(p_base), et = tree4x4f(p_et), ma = tree4x4f(p_ma);
        if (live && lq == 0) {
            float ds = (float)d[s];
            float bias = gsum + ds;
            float root = base - et;
            float mags = own + m1s + ma + gabs + fabsf(bias) + fabsf(root);
            float ec = k.kb * own + 3.0f * k.u * m1s + k.kc * ma + k.g * gabs
                       + k.u * fabsf(ds) + 4.0f * k.u * (fabsf(bias) + fabsf(root))
                       + k.f64slack * mags + k.tiny;
            ec = ec * 1.01f;
            float out = (bias + root) - ec;
            double o = isfinite(out) ? (double)out : -INFINITY;
            lb[(size_t)box * S + s] = o;
            if (lbias && s < PR) {
                // the plane a x + lbias is below f_s on the box: the bound's
                // error without its concretization terms (kc ma, root)
                float ep = (k.kb * own + 3.0f * k.u * m1s + k.g * gabs + k.u * fabsf(ds)
                            + 4.0f * k.u * fabsf(bias) + k.tiny) * 1.01f;
                double pb = (double)bias - (double)ep;
                lbias[(size_t)box * PR + s] = isfinite(pb) ? pb : -INFINITY;
            }
        }
        if constexpr (STAGED) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
    // one flag per box: some plane choice was near a tie and was taken
    // from FP64 (diagnostic); box r's lanes are r, r + 16, r + 32, r + 48
    unsigned long long any = __ballot(ambiguous);
    if (live && lq == 0 && (any & (0x0001000100010001ull << lr))) atomicOr(&flag[box], 1);
}

__global__ void __launch_bounds__(256) zero_i32_kernel(int n, int* __restrict__ p)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

int compute_units_f32()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            v > 0)
            cus = v;
        else
            cus = 256;
    }
    return cus;
}

// rows per workgroup so that about `target` workgroups run
inline int rows_per_block_f32(int S, long nwg, long target)
{
    long chunks = (target + nwg - 1) / nwg;
    if (chunks < 1) chunks = 1;
    if (chunks > S) chunks = S;
    return (int)((S + chunks - 1) / chunks);
}

template <int KT>
int launch_f32(int B, int n, int S, const float* QB, const float* QA, const double* Q64,
               const double* Cm, const double* d, const double* xL, const double* xU,
               const int64_t* rows, double* lb, int* flag, double* lA, double* lbias, int PR,
               const F32Consts& k, hipStream_t stream)
{
    // the flags are OR-ed over the row blocks of a box
    hipLaunchKernelGGL(zero_i32_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, B, flag);
    if (B >= 4096) {   // the LDS-staged form, 128 boxes per workgroup
        const size_t rowstride = (size_t)((KT + 3) / 4) * KT * 64;
        const size_t lds = 4 * rowstride * sizeof(float);
        static bool attr = false;   // once per instantiation (same value every call)
        if (!attr) {
            if (hipFuncSetAttribute((const void*)qcqp_crown_f32s_kernel<KT, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                return GH_ELAUNCH;
            attr = true;
        }
        const long nwg = (B + 127) / 128;
        const int rpb = rows_per_block_f32(S, nwg, compute_units_f32());
        hipLaunchKernelGGL((qcqp_crown_f32s_kernel<KT, true>),
                           dim3((unsigned)nwg, (unsigned)((S + rpb - 1) / rpb)), dim3(512),
                           lds, stream, B, n, S, QB, QA, Q64, Cm, d, xL, xU, rows, lb, flag, lA,
                           lbias, PR, k, rpb);
    } else {
        const long nwg = (B + 15) / 16;
        const int rpb = rows_per_block_f32(S, nwg, 4L * compute_units_f32());
        hipLaunchKernelGGL((qcqp_crown_f32s_kernel<KT, false>),
                           dim3((unsigned)nwg, (unsigned)((S + rpb - 1) / rpb)), dim3(64), 0,
                           stream, B, n, S, QB, QA, Q64, Cm, d, xL, xU, rows, lb, flag, lA, lbias,
                           PR, k, rpb);
    }
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

inline int nch_of(int n) { return (n + 15) / 16; }

}  // namespace

extern "C" size_t gh_qcqp_pack_f32_bytes(int n, int S)
{
    if (n <= 0 || S <= 0 || n > 64) return 0;
    return sizeof(float) * 2 * (size_t)S * nch_of(n) * ((n + 3) / 4) * 64;
}

extern "C" int gh_qcqp_pack_f32(int n, int S, const double* Q, void* packed, hipStream_t stream)
{
    if (n <= 0 || S <= 0 || !Q || !packed) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    int nch = nch_of(n), kt = (n + 3) / 4;
    size_t per = (size_t)S * nch * kt * 64;
    float* QB = (float*)packed;
    float* QA = QB + per;
    hipLaunchKernelGGL(qcqp_pack_f32_kernel, dim3((unsigned)((per + 255) / 256)), dim3(256), 0,
                       stream, n, S, nch, kt, Q, QB, QA);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

namespace {

int crown_f32s(int B, int n, int S, const void* packed, const double* Q, const double* Cm,
               const double* d, const double* xL, const double* xU, const int64_t* rows,
               double* lb, int32_t* flag, double* lA, double* lbias, int plane_rows,
               hipStream_t stream)
{
    if (B < 0 || n <= 0 || S <= 0) return GH_EINVAL;
    if ((lA || lbias) && (plane_rows < 1 || plane_rows > S)) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    if (B == 0) return GH_OK;
    if (!packed || !Q || !Cm || !d || !xL || !xU || !lb || !flag) return GH_EINVAL;
    const double u = 5.9604644775390625e-08;   // 2^-24
    const double g = (double)(n + 1) * u / (1.0 - (double)(n + 1) * u);
    F32Consts k;
    k.u = (float)u;
    k.g = (float)(g * 1.01);
    // y bounds: 2g (two chains) + 1.0000001 u (Q rounding) + 4u (centre / radius) + 4u (final
    // subtractions), on mg + dv with mg_true <= mg (1 + 2g)
    k.ky = (float)((2.0 * g + 9.0 * u) * (1.0 + 2.0 * g) * 1.01);
    // backward: chain g + Q rounding u + data of beta*mg (1 + 2g)
    k.kb = (float)((g + 2.0 * u) * (1.0 + 2.0 * g) * 1.01 + 2.0 * u);
    // concretization: two chains (base, et) 2g plus the rounding of A (2u)
    k.kc = (float)((2.0 * g + 2.0 * u) * 1.01);
    k.tiny = 1e-30f;
    k.f64slack = 1e-12f;
    k.kn64 = (double)(n + 16) * 2.220446049250313080847e-16 * 4.0;
    int nch = nch_of(n), kt = (n + 3) / 4;
    const float* QB = (const float*)packed;
    const float* QA = QB + (size_t)S * nch * kt * 64;
#define GH_KT(K) \
    case K: return launch_f32<K>(B, n, S, QB, QA, Q, Cm, d, xL, xU, rows, lb, flag, lA, \
                                 lbias, plane_rows, k, stream);
    switch (kt) {
        GH_KT(1) GH_KT(2) GH_KT(3) GH_KT(4) GH_KT(5) GH_KT(6) GH_KT(7) GH_KT(8)
        GH_KT(9) GH_KT(10) GH_KT(11) GH_KT(12) GH_KT(13) GH_KT(14) GH_KT(15) GH_KT(16)
    default: return GH_EUNSUPPORTED;
    }
#undef GH_KT
}

}  // namespace

extern "C" int gh_qcqp_crown_f32s(int B, int n, int S, const void* packed, const double* Q,
                                  const double* Cm, const double* d, const double* xL,
                                  const double* xU, double* lb, int32_t* flag, double* lA,
                                  double* lbias, int plane_rows, hipStream_t stream)
{
    return crown_f32s(B, n, S, packed, Q, Cm, d, xL, xU, nullptr, lb, flag, lA, lbias,
                      plane_rows, stream);
}

extern "C" int gh_qcqp_crown_rows_f32s(int B, int n, int S, const void* packed, const double* Q,
                                       const double* Cm, const double* d, const double* xL,
                                       const double* xU, const int64_t* rows, double* lb,
                                       int32_t* flag, double* lA, double* lbias, int plane_rows,
                                       hipStream_t stream)
{
    if (B > 0 && !rows) return GH_EINVAL;
    return crown_f32s(B, n, S, packed, Q, Cm, d, xL, xU, rows, lb, flag, lA, lbias, plane_rows,
                      stream);
}
