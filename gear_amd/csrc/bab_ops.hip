// Branch-and-bound step kernels: widest-variable branching, bisection,
// prune + order-preserving append to the device frontier, strided min.
// The predicates and the arithmetic follow the reference step functions
// cited in include/gearhip.h; the oracle (oracle/crown_qcqp.c) restates the
// same rules and the tests compare the two bit for bit.
#include "common.h"
#include "gearhip_internal.h"

namespace {

__global__ void branch_widest_kernel(int B, int n, const double* __restrict__ xL,
                                     const double* __restrict__ xU, int32_t* __restrict__ idx)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const double* l = xL + (size_t)b * n;
    const double* u = xU + (size_t)b * n;
    int best = 0;
    double bw = u[0] - l[0];
    for (int i = 1; i < n; ++i) {
        double w = u[i] - l[i];
        if (w > bw) { bw = w; best = i; }
    }
    idx[b] = best;
}

// One thread per (child, box, variable): coalesced copies of the parent row
// with the split coordinate replaced (split.py:74-107 with P = 2).
__global__ void split_kernel(int B, int n, const double* __restrict__ xL,
                             const double* __restrict__ xU, const int32_t* __restrict__ idx,
                             double* __restrict__ oL, double* __restrict__ oU)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t per = (size_t)B * n;
    if (t >= 2 * per) return;
    int c = t >= per;
    size_t r = t - c * per;
    int b = (int)(r / n), i = (int)(r - (size_t)b * n);
    double l = xL[r], u = xU[r];
    if (i == idx[b]) {
        double lo = (c + 1) * l + (2 - c - 1) * u;
        double hi = c * l + (2 - c) * u;
        l = lo / 2.0;
        u = hi / 2.0;
    }
    oL[t] = l;
    oU[t] = u;
}

// Verification predicate of get_remaining_index (solving mode) plus the
// invalid-box check.  Returns keep; *verified_thr gets thr[0] when the box
// is verified with a finite objective threshold, +inf otherwise.
__device__ __forceinline__ bool keep_box(int b, int S, int n, const double* __restrict__ lb,
                                         const double* __restrict__ thr,
                                         const double* __restrict__ xL,
                                         const double* __restrict__ xU, double* verified_thr)
{
    const double* L = lb + (size_t)b * S;
    bool ver = false;
    for (int s = 0; s < S; ++s) {
        double Ls = L[s], Ts = thr[s];
        bool gt = isfinite(Ls) && isfinite(Ts) && (Ls > Ts);
        if (s == 0) gt = gt || (isfinite(Ls) && isfinite(Ts) && (Ls >= Ts));
        ver = ver || gt;
    }
    ver = ver || (isinf(L[0]) && L[0] > 0 && isfinite(thr[0]));
    *verified_thr = (ver && isfinite(thr[0])) ? thr[0] : INFINITY;
    bool nonfinite = false, invalid = false;
    const double* l = xL + (size_t)b * n;
    const double* u = xU + (size_t)b * n;
    for (int i = 0; i < n; ++i) {
        double a = l[i], c = u[i];
        nonfinite = nonfinite || !isfinite(a) || !isfinite(c);
        invalid = invalid || (a > c);
    }
    return !ver && !(invalid && !nonfinite);
}

__device__ void atomic_min_f64(double* addr, double v)
{
    unsigned long long* p = (unsigned long long*)addr;
    unsigned long long old = *p;
    while (v < __longlong_as_double((long long)old)) {
        unsigned long long assumed = old;
        old = atomicCAS(p, assumed, (unsigned long long)__double_as_longlong(v));
        if (old == assumed) break;
    }
}

__device__ __forceinline__ double block_min(double v, double* red)
{
    for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    double m = INFINITY;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmin(m, red[i]);
    __syncthreads();
    return m;
}

// Pass 1: keep flags, per-block survivor counts, dropped-threshold minimum.
__global__ __launch_bounds__(256) void filter_count_kernel(
    int B, int S, int n, const double* __restrict__ lb, const double* __restrict__ thr,
    const double* __restrict__ xL, const double* __restrict__ xU, int32_t* __restrict__ dst,
    int32_t* __restrict__ counts, double* __restrict__ disc_min)
{
    __shared__ double red[4];
    int b = blockIdx.x * 256 + threadIdx.x;
    double vt = INFINITY;
    bool k = false;
    if (b < B) {
        k = keep_box(b, S, n, lb, thr, xL, xU, &vt);
        dst[b] = k ? 1 : 0;
    }
    int cnt = __syncthreads_count(k);
    double m = block_min(vt, red);
    if (threadIdx.x == 0) {
        counts[blockIdx.x] = cnt;
        if (m < INFINITY) atomic_min_f64(disc_min, m);
    }
}

// Pass 2: exclusive scan of the block counts (one block), total survivors.
__global__ __launch_bounds__(1024) void filter_scan_kernel(int nblk, int32_t* __restrict__ counts,
                                                           int64_t* __restrict__ total)
{
    __shared__ int64_t part[1024];
    int t = threadIdx.x;
    int per = (nblk + 1023) / 1024;
    int lo = t * per, hi = min(nblk, lo + per);
    int64_t sum = 0;
    for (int i = lo; i < hi; ++i) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - sum;  // exclusive prefix of this thread's chunk
    for (int i = lo; i < hi; ++i) {
        int32_t c = counts[i];
        counts[i] = (int32_t)run;
        run += c;
    }
    if (t == 1023) *total = part[1023];
}

// Pass 3: rank of each survivor inside its block -> destination row
// (relative to base), -1 for dropped boxes.
__global__ __launch_bounds__(256) void filter_rank_kernel(int B, int32_t* __restrict__ dst,
                                                          const int32_t* __restrict__ offsets)
{
    __shared__ int wsum[4];
    int b = blockIdx.x * 256 + threadIdx.x;
    int k = (b < B) ? dst[b] : 0;
    unsigned long long mask = __ballot(k);
    int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long below = (l == 0) ? 0ull : (mask & ((~0ull) >> (64 - l)));
    int rank = __popcll(below);
    if (l == 0) wsum[w] = __popcll(mask);
    __syncthreads();
    int pre = 0;
    for (int i = 0; i < w; ++i) pre += wsum[i];
    if (b < B) dst[b] = k ? offsets[blockIdx.x] + pre + rank : -1;
}

// Pass 4: coalesced copy of surviving rows into the frontier.
__global__ void filter_copy_kernel(int B, int S, int n, const double* __restrict__ lb,
                                   const double* __restrict__ cL, const double* __restrict__ cU,
                                   const int32_t* __restrict__ dst, double* __restrict__ FL,
                                   double* __restrict__ FU, double* __restrict__ Flb,
                                   int64_t base)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nx = (size_t)B * n, ns = (size_t)B * S;
    if (t < nx) {
        int b = (int)(t / n), i = (int)(t - (size_t)b * n);
        int r = dst[b];
        if (r >= 0) {
            size_t o = (size_t)(base + r) * n + i;
            FL[o] = cL[t];
            FU[o] = cU[t];
        }
    } else if (t < nx + ns) {
        size_t u = t - nx;
        int b = (int)(u / S), s = (int)(u - (size_t)b * S);
        int r = dst[b];
        if (r >= 0) Flb[(size_t)(base + r) * S + s] = lb[u];
    }
}

__global__ __launch_bounds__(256) void min_strided_kernel(int64_t N, int64_t stride,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ out)
{
    __shared__ double red[4];
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * 256)
        m = fmin(m, x[i * stride]);
    m = block_min(m, red);
    if (threadIdx.x == 0 && m < INFINITY) atomic_min_f64(out, m);
}

__global__ void set_i64_kernel(int64_t* p, int64_t v) { *p = v; }

// Payload row of each survivor (row k % P of pay) to its frontier row.
__global__ void filter_payload_kernel(int B, int W, int P, const double* __restrict__ pay,
                                      const int32_t* __restrict__ dst, double* __restrict__ Fpay,
                                      int64_t base)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)B * W) return;
    int k = (int)(t / W), w = (int)(t - (size_t)k * W);
    int r = dst[k];
    if (r >= 0) Fpay[(size_t)(base + r) * W + w] = pay[(size_t)(k % P) * W + w];
}

inline int status() { return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH; }

}  // namespace

extern "C" int gh_version(void) { return 100; }

extern "C" int gh_branch_widest_f64(int B, int n, const double* xL, const double* xU, int32_t* idx,
                                    hipStream_t stream)
{
    if (B < 0 || n <= 0 || (B > 0 && (!xL || !xU || !idx))) return GH_EINVAL;
    if (B == 0) return GH_OK;
    hipLaunchKernelGGL(branch_widest_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, B, n, xL,
                       xU, idx);
    return status();
}

extern "C" int gh_split_f64(int B, int n, const double* xL, const double* xU, const int32_t* idx,
                            double* oL, double* oU, hipStream_t stream)
{
    if (B < 0 || n <= 0 || (B > 0 && (!xL || !xU || !idx || !oL || !oU))) return GH_EINVAL;
    if (B == 0) return GH_OK;
    size_t total = 2 * (size_t)B * n;
    hipLaunchKernelGGL(split_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, B,
                       n, xL, xU, idx, oL, oU);
    return status();
}

extern "C" size_t gh_filter_workspace_bytes(int B)
{
    if (B < 0) return 0;
    size_t nblk = ((size_t)B + 255) / 256;
    return sizeof(int32_t) * ((size_t)B + nblk) + 256;
}

extern "C" int gh_filter_append_f64(int B, int S, int n, const double* lb, const double* thr,
                                    const double* cL, const double* cU, double* FL, double* FU,
                                    double* Flb, int64_t base, int64_t cap, int64_t* total,
                                    double* disc_min, void* ws, hipStream_t stream)
{
    if (B < 0 || S <= 0 || n <= 0 || base < 0 || !total || !disc_min) return GH_EINVAL;
    if (B > 0 && (!lb || !thr || !cL || !cU || !FL || !FU || !Flb || !ws)) return GH_EINVAL;
    if (base + (int64_t)B > cap) return GH_EINVAL;
    if (B == 0) {
        // a fill kernel rather than a memset node, so a captured graph
        // replays it faithfully
        hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, stream, total, (int64_t)0);
        return status();
    }
    int nblk = (B + 255) / 256;
    int32_t* dst = (int32_t*)ws;
    int32_t* counts = dst + B;
    hipLaunchKernelGGL(filter_count_kernel, dim3(nblk), dim3(256), 0, stream, B, S, n, lb, thr, cL,
                       cU, dst, counts, disc_min);
    hipLaunchKernelGGL(filter_scan_kernel, dim3(1), dim3(1024), 0, stream, nblk, counts, total);
    hipLaunchKernelGGL(filter_rank_kernel, dim3(nblk), dim3(256), 0, stream, B, dst, counts);
    size_t items = (size_t)B * (n + S);
    hipLaunchKernelGGL(filter_copy_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0,
                       stream, B, S, n, lb, cL, cU, dst, FL, FU, Flb, base);
    return status();
}

extern "C" int gh_filter_payload_f64(int B, int W, int P, const double* pay, const void* ws,
                                     double* Fpay, int64_t base, int64_t cap, hipStream_t stream)
{
    if (B < 0 || W < 0 || P < 1 || base < 0) return GH_EINVAL;
    if (B == 0 || W == 0) return GH_OK;
    if (!pay || !ws || !Fpay || base + (int64_t)B > cap) return GH_EINVAL;
    size_t items = (size_t)B * W;
    hipLaunchKernelGGL(filter_payload_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0,
                       stream, B, W, P, pay, (const int32_t*)ws, Fpay, base);
    return status();
}

extern "C" int gh_min_strided_f64(int64_t N, int64_t stride, const double* x, double* out,
                                  hipStream_t stream)
{
    if (N < 0 || stride <= 0 || !out || (N > 0 && !x)) return GH_EINVAL;
    if (N == 0) return GH_OK;
    int64_t blocks = (N + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(min_strided_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, N, stride,
                       x, out);
    return status();
}

// ------------------------------------------------------------ child keys
namespace {

// Child k (of 2K, [upper halves; lower halves]) of parent k mod K: row 0 is
// max(parent bound, the parent's objective plane minimised over the child
// box and rounded down); rows >= 1 keep the parent's bound
// (batch_branch_and_bound.py:189-225 _plane_minimum_down / _raise_child_keys).
// The minimum sum_i (a_i c_i - |a_i| r_i) + b is lowered by
// (n + 4) u (sum_i |a_i| (|c_i| + r_i) + |b|) and one step toward -inf.
__global__ void __launch_bounds__(256)
child_keys_kernel(int K, int n, int S, int R, const double *__restrict__ plb,
                  const double *__restrict__ lA, const double *__restrict__ lbias,
                  const double *__restrict__ cL, const double *__restrict__ cU,
                  double *__restrict__ keys)
{
    // one thread per child; the sums run in index order, the order the key
    // is specified in, so keys (and the sort and prune that read them) are
    // the same on every device and in the restatement
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * K) return;
    const int p = k % K;
    const double *a = lA + (size_t)p * R * n;  // objective row
    const double *xl = cL + (size_t)k * n, *xu = cU + (size_t)k * n;
    double pm = 0.0, mg = 0.0;
    bool bad = false;
    for (int i = 0; i < n; ++i) {
        const double l = xl[i], u = xu[i];
        const double c = (l + u) / 2, r = (u - l) / 2;
        const double ai = a[i];
        pm += ai * c - fabs(ai) * r;
        mg += fabs(ai) * (fabs(c) + r);
        bad |= !(l <= u);
    }
    for (int s = 0; s < S; ++s) keys[(size_t)k * S + s] = plb[(size_t)p * S + s];
    const double b = lbias[(size_t)p * R];
    const double u53 = 1.1102230246251565404236316680908203125e-16;
    pm = pm + b;
    const double margin = (n + 4) * u53 * (mg + fabs(b));
    const double key = nextafter(pm - margin, -INFINITY);
    if (!bad && isfinite(key)) keys[(size_t)k * S] = fmax(plb[(size_t)p * S], key);
}

}  // namespace

extern "C" int gh_child_keys_f64(int K, int n, int S, int R, const double *plb,
                                 const double *lA, const double *lbias, const double *cL,
                                 const double *cU, double *keys, hipStream_t stream)
{
    if (K == 0) return GH_OK;
    if (K < 0 || n < 1 || S < 1 || R < 1 || R > S || !plb || !lA || !lbias || !cL || !cU || !keys)
        return GH_EINVAL;
    hipLaunchKernelGGL(child_keys_kernel, dim3((2 * K + 255) / 256), dim3(256), 0, stream, K, n,
                       S, R, plb, lA, lbias, cL, cU, keys);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
