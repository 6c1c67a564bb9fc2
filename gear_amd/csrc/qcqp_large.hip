// FP32-sound CROWN lower bound of a QCQP of any size (gfx950).
//
// The same relaxation as the FP64 path (first-linear IBP of y_s = Q_s x,
// coefficient-aware McCormick lower planes of z = x * y, backward product
// A = Cm + alpha + Q_s^T beta, box concretization;
// auto_LiRPA/bound_general.py:1275-1593, operators/bivariate.py:190-334,
// backward_bound.py:880-1206, concretize_bounds.py:267-321), arranged for
// large n:
//   1. boxes are rounded outward to FP32 and written as centre / radius
//      columns (x0, ep) with box ⊆ [x0 - ep, x0 + ep];
//   2. C_s = Q32_s x0, DV_s = |Q32_s| ep and CA_s = |Q32_s| |x0| for every
//      row s and box: two strided-batched FP32 GEMMs (MFMA; rocBLAS with its
//      default IEEE math mode), S batches of (n x n) (n x B) and
//      (n x n) (n x 2B);
//   3. per (s, b, i), in FP64 from the FP32 products: y bounds
//      c -/+ dv -/+ E, with E covering the GEMM rounding for any summation
//      order (gamma_n = n u / (1 - n u)), the rounding of Q to FP32 and the
//      FP64 path's own margin; the plane choice of the FP64 path (elements
//      within a rigorous distance of a tie take the FP64 path's exact
//      choice); beta rounded to FP32, Ainit = fl32(alpha + Cm); per (s, b)
//      sums of gamma (rounded down), |beta| mg and |Ainit| mx;
//   4. A_s = Q32_s^T beta_s + Ainit_s: a third strided-batched FP32 GEMM;
//   5. per (s, b): lb = d + sum gamma + min_box(A x) - E_c - slack, with
//      E_c = (gamma_{n+1} + 2.05 u) (sum |beta| mg + sum |Ainit| mx) (the
//      GEMM rounding, Q and beta rounded to FP32, Ainit rounded), and a slack of 1e-12 of the magnitudes plus twice
//      the FP64 path's enclosure margin.
// With the same plane choices, each FP32 plane lies below the FP64 plane on
// the box (looser y bounds move P1 by (yl - yl')(x - xl) >= 0 and P2 by
// (yu' - yu)(x - xu) <= 0), so the bound is not above the FP64 bound.
#include <rocblas/rocblas.h>

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

// Q32 = fl32(Q), absQ = |Q32|, rowabs[s][i] >= sum_j |Q32[s][i][j]|.
__global__ void __launch_bounds__(64)
large_prepare_kernel(int n, int S, const double *__restrict__ Q, float *__restrict__ Q32,
                     float *__restrict__ absQ, float *__restrict__ rowabs)
{
    const size_t row = blockIdx.x;  // s * n + i
    if (row >= (size_t)S * n) return;
    const double *q = Q + row * n;
    double acc = 0.0;
    for (int j = threadIdx.x; j < n; j += 64) {
        float f = (float)q[j];
        Q32[row * n + j] = f;
        absQ[row * n + j] = fabsf(f);
        acc += fabs((double)f);
    }
    acc = wave_sum(acc);
    // FP64 sum of n FP32 magnitudes: relative error below n * eps64
    if (threadIdx.x == 0) rowabs[row] = up_f32(acc * (1.0 + 1e-9));
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

// Step 3: y bounds, plane choice, beta / Ainit and the (s, b) sums.
// One wavefront per (s, b); sums: [gamma, |gamma|, |beta| mg, |Ainit| mx].
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
__global__ void __launch_bounds__(64)
large_relax_kernel(int B, int n, int S, const float *__restrict__ C, const float *__restrict__ DVCA,
                   const double *__restrict__ mxmax, const double *__restrict__ Q64,
                   const double *__restrict__ Cm, const double *__restrict__ xL,
                   const double *__restrict__ xU, float *__restrict__ BETA, float *__restrict__ AOUT,
                   double *__restrict__ sums, int *__restrict__ flag, LargeConsts k)
{
    const size_t sb = blockIdx.x;  // s * B + b
    if (sb >= (size_t)S * B) return;
    const int s = (int)(sb / B), b = (int)(sb - (size_t)s * B);
    (void)mxmax;
    const double *lo = xL + (size_t)b * n, *hi = xU + (size_t)b * n;
    double gs = 0.0, ga = 0.0, bm = 0.0, am = 0.0;
    int tie = 0;
    const size_t base = sb * n;
    // DVCA per row s: columns [0, B) are |Q| ep, [B, 2B) are |Q| |x0|
    const float *DV = DVCA + (size_t)s * 2 * B * n + (size_t)b * n;
    const float *CA = DV + (size_t)B * n;
    const double f = k.gn + k.uq;
    for (int i = threadIdx.x; i < n; i += 64) {
        const double c = C[base + i], dv = DV[i], ca = CA[i];
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
        const double aplus = alpha + Cm[(size_t)s * n + i];
        const float a32 = (float)aplus;
        BETA[base + i] = b32;
        AOUT[base + i] = a32;
        gs += g;
        ga += fabs(g);
        bm += fabs((double)b32) * mg;
        const double mx32 = fmax(fabs((double)down_f32(xl)), fabs((double)up_f32(xu)));
        am += fabs((double)a32) * mx32;
    }
    gs = wave_sum(gs);
    ga = wave_sum(ga);
    bm = wave_sum(bm);
    am = wave_sum(am);
    tie = wave_any(tie);
    if (threadIdx.x == 0) {
        double *o = sums + sb * 4;
        o[0] = gs;
        o[1] = ga;
        o[2] = bm;
        o[3] = am;
        if (tie) flag[b] = 1;  // a plain store: every writer stores 1
    }
}

// Step 5: concretization and the error terms.
__global__ void __launch_bounds__(64)
large_concretize_kernel(int B, int n, int S, const float *__restrict__ A,
                        const double *__restrict__ sums, const double *__restrict__ d,
                        const double *__restrict__ xL, const double *__restrict__ xU,
                        double *__restrict__ lb, LargeConsts k)
{
    const size_t sb = blockIdx.x;
    if (sb >= (size_t)S * B) return;
    const int s = (int)(sb / B), b = (int)(sb - (size_t)s * B);
    const double *lo = xL + (size_t)b * n, *hi = xU + (size_t)b * n;
    const size_t base = sb * n;
    double mn = 0.0, am = 0.0;
    for (int j = threadIdx.x; j < n; j += 64) {
        const double a = A[base + j];
        const double l = lo[j], h = hi[j];
        mn += a * (a >= 0 ? l : h);
        am += fabs(a) * fmax(fabs(l), fabs(h));
    }
    mn = wave_sum(mn);
    am = wave_sum(am);
    if (threadIdx.x == 0) {
        const double *o = sums + sb * 4;
        const double gs = o[0], ga = o[1], bm = o[2], aim = o[3];
        const double dd = d[s];
        const double f = k.gn1 + 2.05 * U32;
        const double Ec = (f * bm + f * aim) * (1.0 + 1e-9);
        const double mag = am + ga + fabs(dd) + fabs(mn) + fabs(gs);
        const double slack = 1e-12 * mag + 2.0 * k.kn64 * mag + 1e-300;
        double v = ((dd + gs) + mn) - Ec - slack;
        if (!(v == v)) v = -INFINITY;
        lb[(size_t)b * S + s] = v;
    }
}

__global__ void zero_i32_kernel(int *p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// rocBLAS handles, one per device, created on first use.
std::mutex g_mu;
rocblas_handle g_handles[64] = {};

rocblas_handle handle_for_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_handles[dev]) {
        rocblas_handle h;
        if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
        rocblas_set_math_mode(h, rocblas_default_math);
        g_handles[dev] = h;
    }
    return g_handles[dev];
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

struct LargePlanLayout {
    size_t q32, absq, rowabs, total;
};
LargePlanLayout plan_layout(int n, int S)
{
    LargePlanLayout L;
    size_t nn = (size_t)S * n * n * sizeof(float);
    L.q32 = 0;
    L.absq = align256(nn);
    L.rowabs = L.absq + align256(nn);
    L.total = L.rowabs + align256((size_t)S * n * sizeof(float));
    return L;
}

// per-box workspace: X0, EP, |x0| (n floats each), C, BETA, A (S n floats
// each), DV and CA (2 S n floats), sums (4 S doubles), mxmax (1 double)
size_t per_box_bytes(int n, int S)
{
    return (size_t)3 * n * 4 + (size_t)5 * S * n * 4 + (size_t)4 * S * 8 + 8;
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
    LargePlanLayout L = plan_layout(n, S);
    char *p = static_cast<char *>(plan);
    hipLaunchKernelGGL(large_prepare_kernel, dim3((unsigned)((size_t)S * n)), dim3(64), 0, stream, n,
                       S, Q, reinterpret_cast<float *>(p + L.q32),
                       reinterpret_cast<float *>(p + L.absq), reinterpret_cast<float *>(p + L.rowabs));
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
    rocblas_handle h = handle_for_current_device();
    if (!h) return GH_ELAUNCH;
    if (rocblas_set_stream(h, stream) != rocblas_status_success) return GH_ELAUNCH;
    const LargePlanLayout L = plan_layout(n, S);
    const char *pp = static_cast<const char *>(plan);
    const float *Q32 = reinterpret_cast<const float *>(pp + L.q32);
    const float *absQ = reinterpret_cast<const float *>(pp + L.absq);
    const float *rowabs = reinterpret_cast<const float *>(pp + L.rowabs);
    LargeConsts k;
    const double u = U32;
    k.gn = n * u / (1.0 - n * u);
    k.gn1 = (n + 1) * u / (1.0 - (n + 1) * u);
    k.uq = u * (1.0 + u) * 1.01;
    k.kn64 = (double)(n + 16) * EPS64 * 4.0;
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
        // EP and |x0| adjacent: one GEMM with 2 Bc columns gives DV and CA
        float *EP = reinterpret_cast<float *>(take((size_t)2 * Bc * n * 4));
        float *XA = EP + (size_t)Bc * n;
        float *Cb = reinterpret_cast<float *>(take((size_t)S * Bc * n * 4));
        float *DVb = reinterpret_cast<float *>(take((size_t)2 * S * Bc * n * 4));
        float *BETA = reinterpret_cast<float *>(take((size_t)S * Bc * n * 4));
        float *AOUT = reinterpret_cast<float *>(take((size_t)S * Bc * n * 4));
        double *sums = reinterpret_cast<double *>(take((size_t)S * Bc * 4 * 8));
        double *mxm = reinterpret_cast<double *>(take((size_t)Bc * 8));
        const double *cxL = xL + (size_t)b0 * n, *cxU = xU + (size_t)b0 * n;
        hipLaunchKernelGGL(large_boxes_kernel, dim3(Bc), dim3(64), 0, stream, Bc, n, cxL, cxU, X0,
                           EP, XA, mxm);
        const float one = 1.0f, zero = 0.0f;
        const rocblas_stride sq = (rocblas_stride)n * n, sc = (rocblas_stride)n * Bc;
        // C_s (n x Bc, column-major) = Q32_s x0; Q32_s row-major is Q32_s^T column-major
        rocblas_status st = rocblas_sgemm_strided_batched(
            h, rocblas_operation_transpose, rocblas_operation_none, n, Bc, n, &one, Q32, n, sq, X0,
            n, 0, &zero, Cb, n, sc, S);
        if (st == rocblas_status_success)  // [DV | CA]_s = |Q32_s| [ep | |x0|]
            st = rocblas_sgemm_strided_batched(h, rocblas_operation_transpose,
                                               rocblas_operation_none, n, 2 * Bc, n, &one, absQ, n,
                                               sq, EP, n, 0, &zero, DVb, n, 2 * sc, S);
        if (st != rocblas_status_success) return GH_ELAUNCH;
        hipLaunchKernelGGL(large_relax_kernel, dim3((unsigned)((size_t)S * Bc)), dim3(64), 0, stream,
                           Bc, n, S, Cb, DVb, mxm, Q64, Cm, cxL, cxU, BETA, AOUT, sums, flag + b0,
                           k);
        // A_s = Q32_s^T beta_s + Ainit_s
        st = rocblas_sgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, Bc,
                                           n, &one, Q32, n, sq, BETA, n, sc, &one, AOUT, n, sc, S);
        if (st != rocblas_status_success) return GH_ELAUNCH;
        hipLaunchKernelGGL(large_concretize_kernel, dim3((unsigned)((size_t)S * Bc)), dim3(64), 0,
                           stream, Bc, n, S, AOUT, sums, d, cxL, cxU, lb + (size_t)b0 * S, k);
    }
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
