// Branch-and-bound step kernels: widest-variable branching, bisection,
// prune + order-preserving append to the device frontier, strided min.
// The predicates and the arithmetic follow the reference step functions
// cited in include/gearhip.h; the oracle (oracle/crown_qcqp.c) restates the
// same rules and the tests compare the two bit for bit.
#include <mutex>

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
// invalid-box check on a bound row whose entry 0 is L0 and entries s >= 1
// are L[s].  Returns keep; *verified_thr gets thr[0] when the box is
// verified with a finite objective threshold, +inf otherwise.
// box_flags: bit 0 some coordinate non-finite, bit 1 some l > u.
__device__ __forceinline__ bool keep_row(double L0, const double* __restrict__ L, int S,
                                         const double* __restrict__ thr, int box_flags,
                                         double* verified_thr)
{
    bool ver = false;
    for (int s = 0; s < S; ++s) {
        double Ls = s == 0 ? L0 : L[s], Ts = thr[s];
        bool gt = isfinite(Ls) && isfinite(Ts) && (Ls > Ts);
        if (s == 0) gt = gt || (isfinite(Ls) && isfinite(Ts) && (Ls >= Ts));
        ver = ver || gt;
    }
    ver = ver || (isinf(L0) && L0 > 0 && isfinite(thr[0]));
    *verified_thr = (ver && isfinite(thr[0])) ? thr[0] : INFINITY;
    const bool nonfinite = box_flags & 1, invalid = box_flags & 2;
    return !ver && !(invalid && !nonfinite);
}

__device__ __forceinline__ bool keep_box(int b, int S, const double* __restrict__ lb,
                                         const double* __restrict__ thr, int box_flags,
                                         double* verified_thr)
{
    const double* L = lb + (size_t)b * S;
    return keep_row(L[0], L, S, thr, box_flags, verified_thr);
}

// keep_row evaluated by a whole wave, lanes over the rows s = lane,
// lane + 64, ... (the predicate is an OR over rows, so the split changes
// nothing).  Every lane returns the result.
__device__ __forceinline__ bool keep_row_wave(int lane, double L0, const double* __restrict__ L,
                                              int S, const double* __restrict__ thr,
                                              int box_flags, double* verified_thr)
{
    bool any = false;
    for (int s = lane; s < S; s += 64) {
        const double Ls = s == 0 ? L0 : L[s], Ts = thr[s];
        bool gt = isfinite(Ls) && isfinite(Ts) && (Ls > Ts);
        if (s == 0) gt = gt || (isfinite(Ls) && isfinite(Ts) && (Ls >= Ts));
        any = any || gt;
    }
    const double T0 = thr[0];
    const bool ver = __ballot(any) != 0 || (isinf(L0) && L0 > 0 && isfinite(T0));
    *verified_thr = (ver && isfinite(T0)) ? T0 : INFINITY;
    const bool nonfinite = box_flags & 1, invalid = box_flags & 2;
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
    __shared__ int bflags[256];
    // box scan: 16 lanes per box, each reading every 16th coordinate of
    // both bounds (coalesced 128-byte runs), 16 boxes per pass
    const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
    for (int p = 0; p < 16; ++p) {
        const int lb_box = p * 16 + g;
        const int bb = blockIdx.x * 256 + lb_box;
        int fl = 0;
        if (bb < B) {
            const double* l = xL + (size_t)bb * n;
            const double* u = xU + (size_t)bb * n;
            for (int i = gl; i < n; i += 16) {
                const double a = l[i], c = u[i];
                fl |= (!isfinite(a) || !isfinite(c)) ? 1 : 0;
                fl |= (a > c) ? 2 : 0;
            }
        }
        // or over the 16 lanes of the group (xor tree inside a row of 16)
        fl |= __shfl_xor(fl, 8, 16);
        fl |= __shfl_xor(fl, 4, 16);
        fl |= __shfl_xor(fl, 2, 16);
        fl |= __shfl_xor(fl, 1, 16);
        if (gl == 0) bflags[lb_box] = fl;
    }
    __syncthreads();
    int b = blockIdx.x * 256 + threadIdx.x;
    double vt = INFINITY;
    bool k = false;
    if (b < B) {
        k = keep_box(b, S, lb, thr, bflags[threadIdx.x], &vt);
        dst[b] = k ? 1 : 0;
    }
    int cnt = __syncthreads_count(k);
    double m = block_min(vt, red);
    if (threadIdx.x == 0) {
        counts[blockIdx.x] = cnt;
        if (m < INFINITY) atomic_min_f64(disc_min, m);
    }
}

// Pass 2: exclusive scan of the N block counts in place (one workgroup of
// 16 waves, 4096 counts per round: 4 per thread, wave shuffles, then the
// wave totals), total survivors.  Counts and their sums fit in int32 (at
// most 2^31 - 1 boxes).
__global__ __launch_bounds__(1024) void filter_scan_kernel(int N, int32_t* __restrict__ counts,
                                                           int64_t* __restrict__ total)
{
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int base = 0; base < N; base += 4096) {
        const int i0 = base + 4 * t;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (i0 + k < N) ? counts[i0 + k] : 0;
        const int s4 = v[0] + v[1] + v[2] + v[3];
        int x = s4;    // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int wpre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int ws = wsum[j];
            wpre += j < w ? ws : 0;
            tot += ws;
        }
        int run = carry + wpre + x - s4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < N) counts[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
        __syncthreads();   // wsum is rewritten next round
    }
    if (t == 0) *total = carry;
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

// Sort key of a bound row: max_s (r[s] - thr[s]), NaN if any term is NaN
// (torch's amax).
__device__ __forceinline__ double frontier_key(const double* __restrict__ r, int S,
                                               const double* __restrict__ thr)
{
    double m = -INFINITY;
    bool nan = false;
    for (int s = 0; s < S; ++s) {
        const double v = r[s] - thr[s];
        nan = nan || isnan(v);
        m = v > m ? v : m;
    }
    return nan ? NAN : m;
}

// Sort key of logical row i.
__global__ void frontier_keys_kernel(int64_t N, int S, const int64_t* __restrict__ order,
                                     const double* __restrict__ lb, const double* __restrict__ thr,
                                     double* __restrict__ keys)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    keys[i] = frontier_key(lb + order[i] * S, S, thr);
}

// Descending-order rank key of a sort key: every NaN first (torch's
// descending sort), then numbers from +inf down; -0 equals +0.
__device__ __forceinline__ unsigned long long desc_key(double x)
{
    if (isnan(x)) return ~0ull;
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x + 0.0);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// Number of keys of run[0 .. len) that come before x (prefix length), by a
// 32-ary search on the calling half-wavefront (hl = lane in the half; the
// half's 32 lanes must all call it).
__device__ __forceinline__ int64_t rank32(const double* __restrict__ run, int64_t len,
                                          unsigned long long x, bool tie, int hl)
{
    int64_t a = 0, b = len;  // keys [0, a) come before x, keys [b, len) do not
    while (a < b) {
        const int64_t step = (b - a + 31) / 32;
        const int64_t p = a + (int64_t)(hl + 1) * step - 1;
        bool before = false;
        if (p < b) {
            const unsigned long long y = desc_key(run[p]);
            before = y > x || (tie && y == x);
        }
        const unsigned long long mask = __ballot(before);
        const int cnt = __popcll((threadIdx.x & 32) ? (mask >> 32) : (mask & 0xffffffffull));
        const int64_t na = a + (int64_t)cnt * step;
        b = min(b, na + step - 1);
        a = na;
        if (a > b) a = b;
    }
    return a;
}

constexpr int kDescSegCap = 2048;

// Stable merge of two descending runs (A before B on ties): element i of A
// goes to i + #{B strictly before it}, element j of B to j + #{A before or
// tied with it}.  A workgroup takes 256 consecutive elements of A or of B:
// wavefront 0 finds the counts of its first and last element in the other
// run (rank32), the keys between them are copied into LDS (kDescSegCap at
// most, else searched in global memory) and every element's count is a
// binary search there.
__global__ __launch_bounds__(256) void merge_desc_kernel(int64_t nA, const double* __restrict__ kA,
                                                        const int64_t* __restrict__ vA, int64_t nB,
                                                        const double* __restrict__ kB,
                                                        const int64_t* __restrict__ vB,
                                                        double* __restrict__ kO,
                                                        int64_t* __restrict__ vO)
{
    __shared__ unsigned long long seg[kDescSegCap];
    __shared__ int64_t bnd[2];
    const int64_t bA = (nA + 255) / 256;
    const bool inA = (int64_t)blockIdx.x < bA;
    const int64_t base = (inA ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - bA) * 256;
    const int64_t nS = inA ? nA : nB, nT = inA ? nB : nA;
    const double* kS = inA ? kA : kB;
    const double* kT = inA ? kB : kA;
    const bool tie = !inA;   // B's elements count A's equal keys as before them
    if (threadIdx.x < 64) {
        // ranks of the block's first (lanes 0..31) and last (32..63) element
        const int lane = threadIdx.x;
        const int64_t te = (lane & 32) ? min(base + 255, nS - 1) : base;
        const int64_t rk = rank32(kT, nT, desc_key(kS[te]), tie, lane & 31);
        if ((lane & 31) == 0) bnd[lane >> 5] = rk;
    }
    __syncthreads();
    const int64_t s0 = bnd[0], sl = bnd[1] - s0;
    if (sl <= kDescSegCap)
        for (int i = threadIdx.x; i < sl; i += 256) seg[i] = desc_key(kT[s0 + i]);
    __syncthreads();
    const int64_t t = base + threadIdx.x;
    if (t >= nS) return;
    const unsigned long long x = desc_key(kS[t]);
    int64_t a = 0, b = sl;
    if (sl <= kDescSegCap) {
        while (a < b) {
            const int64_t mid = (a + b) >> 1;
            const unsigned long long y = seg[mid];
            if (y > x || (tie && y == x)) a = mid + 1; else b = mid;
        }
    } else {
        while (a < b) {
            const int64_t mid = (a + b) >> 1;
            const unsigned long long y = desc_key(kT[s0 + mid]);
            if (y > x || (tie && y == x)) a = mid + 1; else b = mid;
        }
    }
    kO[t + s0 + a] = kS[t];
    vO[t + s0 + a] = inA ? vA[t] : vB[t];
}

// Inverse of desc_key on its image: the key (every NaN as one NaN, -0 as
// +0 -- the order desc_key gives is unchanged).
__device__ __forceinline__ double from_desc_key(unsigned long long dk)
{
    const unsigned long long u = (dk >> 63) ? (dk & 0x7FFFFFFFFFFFFFFFull) : ~dk;
    return __builtin_bit_cast(double, u);
}

constexpr int kSortTile = 1024;     // rows per tile-sort workgroup
constexpr int kSortThreads = 256;   // its threads
constexpr int kSortTileLog = 10;

// Stable descending sort of one tile of kSortTile keys (keys[i] the key of
// slot order[i]): a bitonic network in LDS on (~desc_key, position)
// ascending -- descending key, ties in input order; positions past N pad
// the tile after every key.  Out: the tile's keys and slot ids in sorted
// order.  Thread t handles pairs t and t + 256 of each step; with a stride
// of at most 64 the pairs of wave w lie in its own segments [128 w, +128)
// and [512 + 128 w, +128), so those steps need no workgroup barrier (the
// wave's LDS operations are ordered), only the steps of larger strides.
__device__ __forceinline__ void sort_pair(unsigned long long* sk, int* si, int j, int stride,
                                          int size)
{
    const int a = 2 * j - (j & (stride - 1)), b = a + stride;
    const unsigned long long ka = sk[a], kb = sk[b];
    const int ia = si[a], ib = si[b];
    const bool a_first = ka < kb || (ka == kb && ia < ib);
    if (a_first != ((a & size) == 0)) {
        sk[a] = kb;
        sk[b] = ka;
        si[a] = ib;
        si[b] = ia;
    }
}

__global__ __launch_bounds__(kSortThreads) void sort_tile_kernel(
    int64_t N, const double* __restrict__ keys, const int64_t* __restrict__ order,
    double* __restrict__ kOut, int64_t* __restrict__ vOut)
{
    static_assert(kSortTile == 4 * kSortThreads && kSortThreads == 256, "pair layout");
    __shared__ unsigned long long sk[kSortTile];
    __shared__ int si[kSortTile];
    const int64_t t0 = (int64_t)blockIdx.x * kSortTile;
    const int cnt = (int)min((int64_t)kSortTile, N - t0);
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = t + k * kSortThreads;
        // desc_key is never 0 on a key, so a real rank is below the padding's
        sk[j] = j < cnt ? ~desc_key(keys[t0 + j]) : ~0ull;
        si[j] = j;
    }
    __syncthreads();
    for (int size = 2; size <= kSortTile; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const bool wide = stride > 64;
            if (wide) __syncthreads();
            sort_pair(sk, si, t, stride, size);
            sort_pair(sk, si, t + kSortThreads, stride, size);
            if (wide) {
                __syncthreads();
            } else {
                // the next step's LDS reads stay after this step's writes
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = t + k * kSortThreads;
        if (j < cnt) {
            kOut[t0 + j] = from_desc_key(~sk[j]);
            vOut[t0 + j] = order[t0 + si[j]];
        }
    }
}

// One pass of the stable merge sort: runs of 2^lg sorted descending, runs
// WAYS g .. WAYS g + WAYS - 1 merged into one (ties: the earlier run first).
// Each element goes to its group's start + its index in its run + its rank
// in each other run of the group: the length of the run's prefix whose keys
// come before it (strictly, or tied as well for the runs before its own).
// A workgroup's 256 elements are consecutive in one run (runs >= 1,024), so
// their ranks in another run lie between the ranks of its first and last
// element.  Wavefront u finds those two ranks in the u-th other run (a
// 32-ary search per half-wavefront: 32 probes, a ballot, about log32 of the
// run length steps), the workgroup copies the keys between them into LDS
// (contiguous loads; kSegCap keys at most, else that run is searched in
// global memory between the two ranks), and every element finds its rank
// by a binary search there.
constexpr int kMergeLogWays = 2;   // four-way: 1,024-row tiles to 65,536 rows in three passes
constexpr int kSegCap = 1024;

template <int LOGW>
__global__ __launch_bounds__(256) void merge_pass_kernel(int64_t N, int lg,
                                                         const double* __restrict__ kIn,
                                                         const int64_t* __restrict__ vIn,
                                                         double* __restrict__ kOut,
                                                         int64_t* __restrict__ vOut)
{
    constexpr int WAYS = 1 << LOGW, O = WAYS - 1;
    static_assert(O <= 4, "one wavefront per other run");
    __shared__ unsigned long long seg[O][kSegCap];
    __shared__ int64_t bnd[O][2];
    const int64_t t0 = (int64_t)blockIdx.x * 256;
    const int64_t run = (int64_t)1 << lg;
    const int64_t r = t0 >> lg;                      // the block's run
    const int q = (int)(r & (WAYS - 1));
    const int64_t g0 = (r - q) << lg;                // the group's start
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int64_t lo[O], len[O];
    bool tie[O];
#pragma unroll
    for (int u = 0; u < O; ++u) {
        const int o = u < q ? u : u + 1;             // the other runs, in order
        lo[u] = g0 + o * run;
        len[u] = lo[u] < N ? min(run, N - lo[u]) : 0;
        tie[u] = o < q;
    }
    if (w < O) {
        // ranks of the block's first (lanes 0..31) and last (32..63) element
        const int64_t te = (lane & 32) ? min(t0 + 255, N - 1) : t0;
        const unsigned long long xe = desc_key(kIn[te]);
        int64_t ln = 0, lw = 0;
        bool tw = false;
#pragma unroll
        for (int u = 0; u < O; ++u)
            if (u == w) {
                ln = len[u];
                lw = lo[u];
                tw = tie[u];
            }
        const int64_t rk = rank32(kIn + lw, ln, xe, tw, lane & 31);
        if ((lane & 31) == 0) bnd[w][lane >> 5] = rk;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < O; ++u) {
        const int64_t s0 = bnd[u][0], sl = bnd[u][1] - s0;
        if (sl <= kSegCap)
            for (int i = threadIdx.x; i < sl; i += 256) seg[u][i] = desc_key(kIn[lo[u] + s0 + i]);
    }
    __syncthreads();
    const int64_t t = t0 + threadIdx.x;
    if (t >= N) return;
    const unsigned long long x = desc_key(kIn[t]);
    int64_t dst = g0 + (t - (r << lg));
#pragma unroll
    for (int u = 0; u < O; ++u) {
        const int64_t s0 = bnd[u][0], sl = bnd[u][1] - s0;
        int64_t a = 0, b = sl;   // rank inside [s0, s0 + sl)
        if (sl <= kSegCap) {
            while (a < b) {
                const int64_t mid = (a + b) >> 1;
                const unsigned long long y = seg[u][mid];
                if (y > x || (tie[u] && y == x)) a = mid + 1; else b = mid;
            }
        } else {
            while (a < b) {
                const int64_t mid = (a + b) >> 1;
                const unsigned long long y = desc_key(kIn[lo[u] + s0 + mid]);
                if (y > x || (tie[u] && y == x)) a = mid + 1; else b = mid;
            }
        }
        dst += s0 + a;
    }
    kOut[dst] = kIn[t];
    vOut[dst] = vIn[t];
}

// Slot ids of the children just appended: child t < *total (a survivor)
// sits in physical row hw + t; the others get the dead record's slot.
__global__ void fill_order_kernel(int B, const int64_t* __restrict__ total, int64_t hw,
                                  int64_t dead, int64_t* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < B) out[t] = t < *total ? hw + t : dead;
}

// Rows idx[i] of up to three row-major arrays (widths w0, w1, w2) into
// contiguous outputs: one wave per row, lanes over the row's w0 + w1 + w2
// entries, up to four per lane loaded before any is stored (no per-element
// index division).
__global__ void __launch_bounds__(256)
gather_rows_kernel(int64_t N, const int64_t* __restrict__ idx, int w0,
                   const double* __restrict__ s0, double* __restrict__ d0, int w1,
                   const double* __restrict__ s1, double* __restrict__ d1, int w2,
                   const double* __restrict__ s2, double* __restrict__ d2)
{
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= N) return;
    const int64_t r = idx[i];
    const int W = w0 + w1 + w2;
    for (int c0 = 0; c0 < W; c0 += 256) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = c0 + 64 * k + lane;
            v[k] = 0.0;
            if (c < w0) v[k] = s0[r * w0 + c];
            else if (c < w0 + w1) v[k] = s1[r * w1 + (c - w0)];
            else if (c < W) v[k] = s2[r * w2 + (c - w0 - w1)];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = c0 + 64 * k + lane;
            if (c < w0) d0[i * w0 + c] = v[k];
            else if (c < w0 + w1) d1[i * w1 + (c - w0)] = v[k];
            else if (c < W) d2[i * w2 + (c - w0 - w1)] = v[k];
        }
    }
}

// gather_rows_kernel for rows of W = w0 + w1 + w2 <= LPR entries: 64 / LPR
// rows per wave, one entry per lane.
template <int LPR>
__global__ void __launch_bounds__(256)
gather_rows_lanes_kernel(int64_t N, const int64_t* __restrict__ idx, int w0,
                         const double* __restrict__ s0, double* __restrict__ d0, int w1,
                         const double* __restrict__ s1, double* __restrict__ d1, int w2,
                         const double* __restrict__ s2, double* __restrict__ d2)
{
    const int lane = threadIdx.x & 63, c = lane % LPR;
    const int64_t i = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / LPR) + lane / LPR;
    const int W = w0 + w1 + w2;
    if (i >= N || c >= W) return;
    const int64_t r = idx[i];
    if (c < w0) d0[i * w0 + c] = s0[r * w0 + c];
    else if (c < w0 + w1) d1[i * w1 + (c - w0)] = s1[r * w1 + (c - w0)];
    else d2[i * w2 + (c - w0 - w1)] = s2[r * w2 + (c - w0 - w1)];
}

// min over i < N of x[idx[i] * stride] into *out (atomic min).
__global__ __launch_bounds__(256) void min_rows_kernel(int64_t N, int64_t stride,
                                                       const int64_t* __restrict__ idx,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ out)
{
    __shared__ double red[4];
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * 256)
        m = fmin(m, x[idx[i] * stride]);
    m = block_min(m, red);
    if (threadIdx.x == 0 && m < INFINITY) atomic_min_f64(out, m);
}

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

// Keys of N rows, one thread per row (the row's S loads in flight together
// beat lane groups, which need a reduction).
void launch_frontier_keys(int64_t N, int S, const int64_t* order, const double* lb,
                          const double* thr, double* keys, hipStream_t stream)
{
    hipLaunchKernelGGL(frontier_keys_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       stream, N, S, order, lb, thr, keys);
}

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

extern "C" int gh_frontier_keys_f64(int64_t N, int S, const int64_t* order, const double* lb,
                                    const double* thr, double* keys, hipStream_t stream)
{
    if (N < 0 || S <= 0 || (N > 0 && (!order || !lb || !thr || !keys))) return GH_EINVAL;
    if (N == 0) return GH_OK;
    launch_frontier_keys(N, S, order, lb, thr, keys, stream);
    return status();
}

extern "C" int gh_merge_desc_f64(int64_t nA, const double* kA, const int64_t* vA, int64_t nB,
                                 const double* kB, const int64_t* vB, double* kO, int64_t* vO,
                                 hipStream_t stream)
{
    if (nA < 0 || nB < 0) return GH_EINVAL;
    const int64_t N = nA + nB;
    if (N == 0) return GH_OK;
    if ((nA && (!kA || !vA)) || (nB && (!kB || !vB)) || !kO || !vO) return GH_EINVAL;
    hipLaunchKernelGGL(merge_desc_kernel,
                       dim3((unsigned)((nA + 255) / 256 + (nB + 255) / 256)), dim3(256), 0, stream,
                       nA, kA, vA, nB, kB, vB, kO, vO);
    return status();
}

extern "C" int gh_frontier_sort_f64(int64_t N, int S, const int64_t* order, const double* lb,
                                    const double* thr, double* kOut, int64_t* vOut, double* kTmp,
                                    int64_t* vTmp, hipStream_t stream)
{
    if (N < 0 || S <= 0) return GH_EINVAL;
    if (N == 0) return GH_OK;
    if (!order || !lb || !thr || !kOut || !vOut || !kTmp || !vTmp) return GH_EINVAL;
    int passes = 0;
    for (int lg = kSortTileLog; ((int64_t)1 << lg) < N; lg += kMergeLogWays) ++passes;
    // keys first (into the buffer the tile sort does not write), the tile
    // sort where the last merge pass must end in kOut
    double* k0 = (passes & 1) ? kTmp : kOut;
    int64_t* v0 = (passes & 1) ? vTmp : vOut;
    double* k1 = (passes & 1) ? kOut : kTmp;
    int64_t* v1 = (passes & 1) ? vOut : vTmp;
    const unsigned blocks = (unsigned)((N + 255) / 256);
    launch_frontier_keys(N, S, order, lb, thr, k1, stream);
    const int64_t tiles = (N + kSortTile - 1) / kSortTile;
    hipLaunchKernelGGL(sort_tile_kernel, dim3((unsigned)tiles), dim3(kSortThreads), 0, stream, N,
                       (const double*)k1, order, k0, v0);
    for (int p = 0, lg = kSortTileLog; p < passes; ++p, lg += kMergeLogWays) {
        hipLaunchKernelGGL(merge_pass_kernel<kMergeLogWays>, dim3(blocks), dim3(256), 0, stream,
                           N, lg, k0, v0, k1, v1);
        double* kt = k0;
        k0 = k1;
        k1 = kt;
        int64_t* vt = v0;
        v0 = v1;
        v1 = vt;
    }
    return status();
}

extern "C" int gh_fill_order_i64(int B, const int64_t* total, int64_t hw, int64_t dead,
                                 int64_t* out, hipStream_t stream)
{
    if (B < 0 || (B > 0 && (!total || !out))) return GH_EINVAL;
    if (B == 0) return GH_OK;
    hipLaunchKernelGGL(fill_order_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, B, total,
                       hw, dead, out);
    return status();
}

extern "C" int gh_gather_rows_f64(int64_t N, const int64_t* idx, int w0, const double* s0,
                                  double* d0, int w1, const double* s1, double* d1, int w2,
                                  const double* s2, double* d2, hipStream_t stream)
{
    if (N < 0 || w0 < 0 || w1 < 0 || w2 < 0) return GH_EINVAL;
    if (N == 0 || w0 + w1 + w2 == 0) return GH_OK;
    if (!idx || (w0 && (!s0 || !d0)) || (w1 && (!s1 || !d1)) || (w2 && (!s2 || !d2)))
        return GH_EINVAL;
    const int W = w0 + w1 + w2;
    if (W <= 16)
        hipLaunchKernelGGL(gather_rows_lanes_kernel<16>, dim3((unsigned)((N + 15) / 16)),
                           dim3(256), 0, stream, N, idx, w0, s0, d0, w1, s1, d1, w2, s2, d2);
    else if (W <= 32)
        hipLaunchKernelGGL(gather_rows_lanes_kernel<32>, dim3((unsigned)((N + 7) / 8)), dim3(256),
                           0, stream, N, idx, w0, s0, d0, w1, s1, d1, w2, s2, d2);
    else
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0,
                           stream, N, idx, w0, s0, d0, w1, s1, d1, w2, s2, d2);
    return status();
}

extern "C" int gh_min_rows_f64(int64_t N, int64_t stride, const int64_t* idx, const double* x,
                               double* out, hipStream_t stream)
{
    if (N < 0 || stride <= 0 || !out || (N > 0 && (!x || !idx))) return GH_EINVAL;
    if (N == 0) return GH_OK;
    int64_t blocks = (N + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(min_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, N, stride,
                       idx, x, out);
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

// "red64s": lane l sums the terms i = l, l + 64, ... in order from +0.0,
// then p[l] + p[l ^ o] for o = 32, 16, 8, 4, 2, 1 (every lane ends with
// the sum).  The oracle's plane_minimum_down states the same order.
__device__ __forceinline__ double red64x(double v)
{
    v = v + gh::xor32_f64(v);
    v = v + gh::xor16_f64(v);
    return gh::tree16(v);
}

constexpr double kU53 = 1.1102230246251565404236316680908203125e-16;

// Wave-wide tail of a child key: the red64s sums of the per-lane partials
// pm (plane value at the centre minus its radius term) and mg (magnitude),
// the margin and the step toward -inf.
__device__ __forceinline__ double child_key_final(int n, double pm_sum, double mg_sum, bool bad,
                                                  double b, double p0)
{
    const double pm = pm_sum + b;
    const double margin = (n + 4) * kU53 * (mg_sum + fabs(b));
    const double key = nextafter(pm - margin, -INFINITY);
    return (!bad && isfinite(key)) ? fmax(p0, key) : p0;
}

__device__ __forceinline__ double child_key_tail(int n, double pm, double mg, bool bad, double b,
                                                 double p0)
{
    return child_key_final(n, red64x(pm), red64x(mg), __ballot(bad) != 0, b, p0);
}

// Key of one child box [xl, xu] from its parent's objective plane (a, b)
// and bound row plb: max(plb[0], the plane's minimum over the box lowered
// by (n + 4) u (sum |a| (|c| + r) + |b|) and one step toward -inf);
// plb[0] where the box is empty or the minimum is not finite
// (batch_branch_and_bound.py:189-225 _plane_minimum_down /
// _raise_child_keys).  Wave-wide; every lane returns the key.
__device__ __forceinline__ double child_key(int n, int lane, const double* __restrict__ a,
                                            double b, const double* __restrict__ xl,
                                            const double* __restrict__ xu, double p0)
{
    double pm = 0.0, mg = 0.0;
    bool bad = false;
    for (int i = lane; i < n; i += 64) {
        const double l = xl[i], u = xu[i];
        const double c = (l + u) / 2, r = (u - l) / 2;
        const double ai = a[i];
        pm = pm + (ai * c - fabs(ai) * r);
        mg = mg + fabs(ai) * (fabs(c) + r);
        bad = bad || !(l <= u);
    }
    return child_key_tail(n, pm, mg, bad, b, p0);
}

// First of the widest coordinates of [l, u] as branch_widest_kernel's
// sequential scan finds it (a NaN width at 0 keeps 0, later NaN widths
// never win).  Wave-wide; every lane returns the index.
__device__ __forceinline__ int first_widest(int n, int lane, const double* __restrict__ l,
                                            const double* __restrict__ u)
{
    // per-lane candidates in index order, then the largest width with the
    // lowest index
    double bw = -INFINITY;
    int bi = -1;
    const double w0 = u[0] - l[0];
    for (int i = lane; i < n; i += 64) {
        const double w = u[i] - l[i];
        if (bi < 0 ? !isnan(w) : w > bw) {
            bw = w;
            bi = i;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const double w2 = __shfl_xor(bw, o);
        const int i2 = __shfl_xor(bi, o);
        if (i2 >= 0 && (bi < 0 || w2 > bw || (w2 == bw && i2 < bi))) {
            bw = w2;
            bi = i2;
        }
    }
    return (isnan(w0) || bi < 0) ? 0 : bi;
}

// Coordinate i of child c of a box split at idx: c = 0 the upper half,
// c = 1 the lower half (split.py:74-107 with P = 2, split_kernel's).
__device__ __forceinline__ void split_coord(double& lo, double& hi, int i, int idx, int c)
{
    if (i == idx) {
        const double a = (c + 1) * lo + (2 - c - 1) * hi;
        const double b = c * lo + (2 - c) * hi;
        lo = a / 2.0;
        hi = b / 2.0;
    }
}

// child_key of child c of the parent box [l, u] split at idx, the child's
// coordinates formed in registers (the same values split_coord writes);
// a == NULL gives p0.  *flags gets the child box's bits (bit 0 some
// coordinate non-finite, bit 1 some l > u).  Wave-wide.
__device__ __forceinline__ double split_child_key(int n, int lane, const double* __restrict__ a,
                                                  double b, const double* __restrict__ l,
                                                  const double* __restrict__ u, int idx, int c,
                                                  double p0, int* flags)
{
    double pm = 0.0, mg = 0.0;
    bool bad = false, nonfin = false, inval = false;
    for (int i = lane; i < n; i += 64) {
        double lo = l[i], hi = u[i];
        split_coord(lo, hi, i, idx, c);
        nonfin = nonfin || !isfinite(lo) || !isfinite(hi);
        inval = inval || lo > hi;
        if (a) {
            const double cc = (lo + hi) / 2, r = (hi - lo) / 2;
            const double ai = a[i];
            pm = pm + (ai * cc - fabs(ai) * r);
            mg = mg + fabs(ai) * (fabs(cc) + r);
            bad = bad || !(lo <= hi);
        }
    }
    *flags = (__ballot(nonfin) ? 1 : 0) | (__ballot(inval) ? 2 : 0);
    return a ? child_key_tail(n, pm, mg, bad, b, p0) : p0;
}

// One wave per child k of parent k % K (children [upper halves; lower
// halves], the layout of gh_split_f64).
__global__ void __launch_bounds__(256)
child_keys_kernel(int K, int n, int S, int R, const double *__restrict__ plb,
                  const double *__restrict__ lA, const double *__restrict__ lbias,
                  const double *__restrict__ cL, const double *__restrict__ cU,
                  double *__restrict__ keys)
{
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= 2 * K) return;
    const int p = k % K;
    const double key = child_key(n, lane, lA + (size_t)p * R * n, lbias[(size_t)p * R],
                                 cL + (size_t)k * n, cU + (size_t)k * n, plb[(size_t)p * S]);
    for (int s = lane; s < S; s += 64)
        keys[(size_t)k * S + s] = s == 0 ? key : plb[(size_t)p * S + s];
}

// Widest branching, bisection and both children's keys of parent p in one
// pass (one wave per parent).  The index is the first of the widest
// coordinates as branch_widest_kernel's sequential scan finds it (a NaN
// width at 0 keeps 0, later NaN widths never win); the children are
// split_kernel's; their keys are child_keys_kernel's (plb copied when
// there are no planes).
__global__ void __launch_bounds__(256)
branch_split_keys_kernel(int K, int n, int S, int R, const double *__restrict__ pL,
                         const double *__restrict__ pU, const double *__restrict__ plb,
                         const double *__restrict__ lA, const double *__restrict__ lbias,
                         double *__restrict__ cL, double *__restrict__ cU,
                         double *__restrict__ keys)
{
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= K) return;
    const double *l = pL + (size_t)p * n, *u = pU + (size_t)p * n;
    const int idx = first_widest(n, lane, l, u);
    // children: c = 0 the upper half (row p), c = 1 the lower half (row p + K)
    for (int c = 0; c < 2; ++c) {
        double *ol = cL + (size_t)(p + c * K) * n, *ou = cU + (size_t)(p + c * K) * n;
        for (int i = lane; i < n; i += 64) {
            double lo = l[i], hi = u[i];
            split_coord(lo, hi, i, idx, c);
            ol[i] = lo;
            ou[i] = hi;
        }
    }
    // each lane reads back only the coordinates it wrote (same strided
    // mapping), so program order makes them visible
    const double p0 = plb[(size_t)p * S];
    for (int c = 0; c < 2; ++c) {
        const size_t k = (size_t)p + (size_t)c * K;
        const double key = lA ? child_key(n, lane, lA + (size_t)p * R * n, lbias[(size_t)p * R],
                                          cL + k * n, cU + k * n, p0)
                              : p0;
        for (int s = lane; s < S; s += 64) keys[k * S + s] = s == 0 ? key : plb[(size_t)p * S + s];
    }
}

// The reordered step's decision, split, child keys, prune and append
// without materializing the children.  Pass 1, per parent p (workgroups
// of kSplitBlock parents): widest index, both children's keys, box flags
// and keep predicates (branch_split_keys_kernel's and filter_count_kernel's
// values); per workgroup and side the kept count, and the dropped
// thresholds' minimum into disc_min.
constexpr int kSplitBlock = 16;

// The workgroup's kept counts per side (nw waves' counts) and the dropped
// thresholds' minimum m, written by thread 0.
__device__ __forceinline__ void split_block_counts(int nw, const int* kc0, const int* kc1,
                                                   int32_t* __restrict__ cnt, double m,
                                                   double* __restrict__ disc_min)
{
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int w = 0; w < nw; ++w) {
            a += kc0[w];
            b += kc1[w];
        }
        cnt[blockIdx.x] = a;
        cnt[gridDim.x + blockIdx.x] = b;
        if (m < INFINITY) atomic_min_f64(disc_min, m);
    }
}

// Any n and S: one wave per parent, 16 waves per workgroup.
__global__ void __launch_bounds__(1024)
split_keep_wave_kernel(int K, int n, int S, int R, const double *__restrict__ pL,
                       const double *__restrict__ pU, const int64_t *__restrict__ prows,
                  const double *__restrict__ plb,
                       const double *__restrict__ lA, const double *__restrict__ lbias,
                       const double *__restrict__ thr, const int32_t *__restrict__ cidx,
                       double *__restrict__ ckey,
                       int32_t *__restrict__ sidx, int32_t *__restrict__ keep,
                       int32_t *__restrict__ cnt, double *__restrict__ disc_min)
{
    __shared__ int kc[2][16];
    __shared__ double red[16];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = blockIdx.x * kSplitBlock + wv;
    int kept[2] = {0, 0};
    double vt = INFINITY;
    if (p < K) {
        const size_t br = prows ? (size_t)prows[p] : (size_t)p;   // the parent's box row
        const double *l = pL + br * n, *u = pU + br * n;
        const double *prow = plb + (size_t)p * S;
        const double *a = lA ? lA + (size_t)p * R * n : nullptr;
        const double b = lA ? lbias[(size_t)p * R] : 0.0;
        const double p0 = prow[0];
        const int idx = cidx ? cidx[p] : first_widest(n, lane, l, u);
        for (int c = 0; c < 2; ++c) {
            int fl;
            double v;
            const double key = split_child_key(n, lane, a, b, l, u, idx, c, p0, &fl);
            const bool k = keep_row_wave(lane, key, prow, S, thr, fl, &v);
            const size_t ch = (size_t)p + (size_t)c * K;
            if (lane == 0) {
                ckey[ch] = key;
                keep[ch] = k ? 1 : 0;
            }
            kept[c] = k ? 1 : 0;
            vt = fmin(vt, v);
        }
        if (lane == 0) sidx[p] = idx;
    }
    if (lane == 0) {
        kc[0][wv] = kept[0];
        kc[1][wv] = kept[1];
    }
    const double m = block_min(vt, red);   // synchronizes the workgroup first
    split_block_counts(16, kc[0], kc[1], cnt, m, disc_min);
}

// OR of a predicate over the 16-lane group (lanes 16g .. 16g + 15) of lane.
__device__ __forceinline__ bool group_any(bool x, int lane)
{
    return ((__ballot(x) >> (lane & 48)) & 0xFFFFull) != 0;
}

// n <= 64 and S <= 64: four parents per wave, 16 lanes each; lane gl of a
// group holds coordinates and bound rows gl + 16k (k < 4).  The child-key
// sums are red64s: lane gl's four partials are the 64-lane form's lanes
// gl + 16k, so levels l^32 and l^16 are register pairs (k, k ^ 2) and
// (k, k ^ 1) and levels l^8 .. l^1 the 16-lane DPP tree -- the same operand
// pairs.  The widest index is the unique (largest width, lowest index), so
// any reduction order finds it.
__global__ void __launch_bounds__(256)
split_keep16_kernel(int K, int n, int S, int R, const double *__restrict__ pL,
                    const double *__restrict__ pU, const int64_t *__restrict__ prows,
                  const double *__restrict__ plb,
                    const double *__restrict__ lA, const double *__restrict__ lbias,
                    const double *__restrict__ thr, const int32_t *__restrict__ cidx,
                    double *__restrict__ ckey,
                    int32_t *__restrict__ sidx, int32_t *__restrict__ keep,
                    int32_t *__restrict__ cnt, double *__restrict__ disc_min)
{
    __shared__ int kc[2][4];
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gl = lane & 15;
    const int p = blockIdx.x * kSplitBlock + wv * 4 + (lane >> 4);
    const bool valid = p < K;
    const bool planes = lA != nullptr;
    double li[4], ui[4], ai[4], pr[4], th[4];
    double b = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) li[k] = ui[k] = ai[k] = pr[k] = th[k] = 0.0;
    if (valid) {
        const size_t br = prows ? (size_t)prows[p] : (size_t)p;   // the parent's box row
        const double *l = pL + br * n, *u = pU + br * n;
        const double *prow = plb + (size_t)p * S;
        const double *a = planes ? lA + (size_t)p * R * n : nullptr;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = gl + 16 * k;
            if (i < n) {
                li[k] = l[i];
                ui[k] = u[i];
                if (planes) ai[k] = a[i];
            }
            if (i < S) {
                pr[k] = prow[i];
                th[k] = thr[i];
            }
        }
        if (planes) b = lbias[(size_t)p * R];
    }
    const int lead = lane & 48;
    const double p0 = __shfl(pr[0], lead), T0 = __shfl(th[0], lead);
    // first widest coordinate: per-lane candidates in index order, then the
    // 16-lane reduction
    double bw = -INFINITY;
    int bi = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = gl + 16 * k;
        const double w = ui[k] - li[k];
        if (i < n && (bi < 0 ? !isnan(w) : w > bw)) {
            bw = w;
            bi = i;
        }
    }
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) {
        const double w2 = __shfl_xor(bw, o);
        const int i2 = __shfl_xor(bi, o);
        if (i2 >= 0 && (bi < 0 || w2 > bw || (w2 == bw && i2 < bi))) {
            bw = w2;
            bi = i2;
        }
    }
    const double w0 = __shfl(ui[0] - li[0], lead);
    // a given choice (smart branching) replaces the widest input
    const int idx = (cidx && valid) ? cidx[p] : ((isnan(w0) || bi < 0) ? 0 : bi);
    int kept[2];
    double vt = INFINITY;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double pm[4], mg[4];
        bool bad = false, nonfin = false, inval = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = gl + 16 * k;
            double lo = li[k], hi = ui[k];
            split_coord(lo, hi, i, idx, c);
            pm[k] = 0.0;
            mg[k] = 0.0;
            if (i < n) {
                nonfin = nonfin || !isfinite(lo) || !isfinite(hi);
                inval = inval || lo > hi;
                const double cc = (lo + hi) / 2, r = (hi - lo) / 2;
                pm[k] = pm[k] + (ai[k] * cc - fabs(ai[k]) * r);
                mg[k] = mg[k] + fabs(ai[k]) * (fabs(cc) + r);
                bad = bad || !(lo <= hi);
            }
        }
        const int fl = (group_any(nonfin, lane) ? 1 : 0) | (group_any(inval, lane) ? 2 : 0);
        double key = p0;
        if (planes) {
            const double ps = gh::tree16((pm[0] + pm[2]) + (pm[1] + pm[3]));
            const double ms = gh::tree16((mg[0] + mg[2]) + (mg[1] + mg[3]));
            key = child_key_final(n, ps, ms, group_any(bad, lane), b, p0);
        }
        // keep predicate on the row [key, plb[p][1:]]
        bool gt_any = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int s = gl + 16 * k;
            if (s < S) {
                const double Ls = s == 0 ? key : pr[k], Ts = th[k];
                bool gt = isfinite(Ls) && isfinite(Ts) && (Ls > Ts);
                if (s == 0) gt = gt || (isfinite(Ls) && isfinite(Ts) && (Ls >= Ts));
                gt_any = gt_any || gt;
            }
        }
        const bool ver = group_any(gt_any, lane) || (isinf(key) && key > 0 && isfinite(T0));
        const double v = (ver && isfinite(T0)) ? T0 : INFINITY;
        const bool kp = !ver && !((fl & 2) && !(fl & 1));
        const size_t ch = (size_t)p + (size_t)c * K;
        if (valid && gl == 0) {
            ckey[ch] = key;
            keep[ch] = kp ? 1 : 0;
        }
        kept[c] = __popcll(__ballot(valid && gl == 0 && kp));
        if (valid) vt = fmin(vt, v);
    }
    if (valid && gl == 0) sidx[p] = idx;
    if (lane == 0) {
        kc[0][wv] = kept[0];
        kc[1][wv] = kept[1];
    }
    const double m = block_min(vt, red);   // synchronizes the workgroup first
    split_block_counts(4, kc[0], kc[1], cnt, m, disc_min);
}

// Pass 3 (after filter_scan_kernel turned the counts, side 0 blocks then
// side 1 blocks, into exclusive offsets): one wave per parent p (16 per
// workgroup, the blocks of pass 1) writes each kept child's box, rebuilt
// from the parent, and bound row [key, plb[p][1:]] to frontier row
// base + rank, and both children's slot ids (base + t for the first *total
// positions t, the dead record after them).  Every load is issued before
// the first store.  prows may be the first K entries of order_out itself
// (the frontier's picks, read in place): the wave of parent p reads only
// prows[p], before it writes order_out[p] and order_out[p + K], and no
// other wave reads either.
__global__ void __launch_bounds__(1024)
split_copy_kernel(int K, int n, int S, const double *__restrict__ pL,
                  const double *__restrict__ pU, const int64_t *__restrict__ prows,
                  const double *__restrict__ plb,
                  const double *__restrict__ ckey, const int32_t *__restrict__ sidx,
                  const int32_t *__restrict__ keep, const int32_t *__restrict__ offs,
                  const int64_t *__restrict__ total, double *__restrict__ FL,
                  double *__restrict__ FU, double *__restrict__ Flb, int64_t base,
                  int64_t *__restrict__ order_out, int64_t dead)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const int p = blk * kSplitBlock + wv;
    if (p >= K) return;
    // keep flags of the block's parents: lanes 0..15 side 0, 16..31 side 1
    const int q = blk * kSplitBlock + (lane & 15);
    const bool kq = lane < 32 && q < K && keep[(size_t)q + (size_t)(lane >> 4) * K];
    const int32_t off0 = offs[blk], off1 = offs[nblk + blk];
    const int idx = sidx[p];
    const int64_t T = *total;
    const double key0 = ckey[p], key1 = ckey[(size_t)p + K];
    const size_t br = prows ? (size_t)prows[p] : (size_t)p;   // the parent's box row
    const double *l = pL + br * n, *u = pU + br * n;
    const double *prow = plb + (size_t)p * S;
    const double li = lane < n ? l[lane] : 0.0, ui = lane < n ? u[lane] : 0.0;
    const double pb = lane < S ? prow[lane] : 0.0;
    const unsigned long long km = __ballot(kq);
    for (int c = 0; c < 2; ++c) {
        const unsigned long long side = (km >> (16 * c)) & 0xFFFFull;
        if (!((side >> wv) & 1)) continue;
        const int64_t row =
            base + (c ? off1 : off0) + __popcll(side & ((1ull << wv) - 1ull));
        double *fl = FL + (size_t)row * n, *fu = FU + (size_t)row * n;
        if (lane < n) {
            double lo = li, hi = ui;
            split_coord(lo, hi, lane, idx, c);
            fl[lane] = lo;
            fu[lane] = hi;
        }
        for (int i = lane + 64; i < n; i += 64) {
            double lo = l[i], hi = u[i];
            split_coord(lo, hi, i, idx, c);
            fl[i] = lo;
            fu[i] = hi;
        }
        double *fb = Flb + (size_t)row * S;
        if (lane < S) fb[lane] = lane == 0 ? (c ? key1 : key0) : pb;
        for (int s = lane + 64; s < S; s += 64) fb[s] = prow[s];
    }
    if (lane < 2) {
        const int64_t ch = (int64_t)p + (int64_t)lane * K;
        order_out[ch] = ch < T ? base + ch : dead;
    }
}

// split_copy_kernel for n <= 64 and S <= 64: four parents per wave, 16
// lanes each (lane gl of a group holds coordinates and bound entries
// gl + 16k), the same writes.  prows may alias order_out as there: the
// group of parent p reads only prows[p] before it writes order_out[p] and
// order_out[p + K].
__global__ void __launch_bounds__(256)
split_copy16_kernel(int K, int n, int S, const double *__restrict__ pL,
                    const double *__restrict__ pU, const int64_t *__restrict__ prows,
                    const double *__restrict__ plb, const double *__restrict__ ckey,
                    const int32_t *__restrict__ sidx, const int32_t *__restrict__ keep,
                    const int32_t *__restrict__ offs, const int64_t *__restrict__ total,
                    double *__restrict__ FL, double *__restrict__ FU, double *__restrict__ Flb,
                    int64_t base, int64_t *__restrict__ order_out, int64_t dead)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, gl = lane & 15;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const int j = wv * 4 + (lane >> 4);           // the parent's place in the block
    const int p = blk * kSplitBlock + j;
    // keep flags of the block's parents: lanes 0..15 side 0, 16..31 side 1
    const int q = blk * kSplitBlock + (lane & 15);
    const bool kq = lane < 32 && q < K && keep[(size_t)q + (size_t)(lane >> 4) * K];
    const unsigned long long km = __ballot(kq);
    if (p >= K) return;
    const int32_t off0 = offs[blk], off1 = offs[nblk + blk];
    const int idx = sidx[p];
    const int64_t T = *total;
    const double key0 = ckey[p], key1 = ckey[(size_t)p + K];
    const size_t br = prows ? (size_t)prows[p] : (size_t)p;   // the parent's box row
    const double *l = pL + br * n, *u = pU + br * n;
    const double *prow = plb + (size_t)p * S;
    double li[4], ui[4], pb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = gl + 16 * k;
        li[k] = i < n ? l[i] : 0.0;
        ui[k] = i < n ? u[i] : 0.0;
        pb[k] = i < S ? prow[i] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const unsigned long long side = (km >> (16 * c)) & 0xFFFFull;
        if (!((side >> j) & 1)) continue;
        const int64_t row = base + (c ? off1 : off0) + __popcll(side & ((1ull << j) - 1ull));
        double *fl = FL + (size_t)row * n, *fu = FU + (size_t)row * n;
        double *fb = Flb + (size_t)row * S;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = gl + 16 * k;
            if (i < n) {
                double lo = li[k], hi = ui[k];
                split_coord(lo, hi, i, idx, c);
                fl[i] = lo;
                fu[i] = hi;
            }
            if (i < S) fb[i] = i == 0 ? (c ? key1 : key0) : pb[k];
        }
    }
    if (gl < 2) {
        const int64_t ch = (int64_t)p + (int64_t)gl * K;
        order_out[ch] = ch < T ? base + ch : dead;
    }
}

}  // namespace

extern "C" int gh_branch_split_keys_f64(int K, int n, int S, int R, const double *pL,
                                        const double *pU, const double *plb, const double *lA,
                                        const double *lbias, double *cL, double *cU,
                                        double *keys, hipStream_t stream)
{
    if (K == 0) return GH_OK;
    if (K < 0 || n < 1 || S < 1 || !pL || !pU || !plb || !cL || !cU || !keys) return GH_EINVAL;
    if (lA && (R < 1 || R > S || !lbias)) return GH_EINVAL;
    hipLaunchKernelGGL(branch_split_keys_kernel, dim3((K + 3) / 4), dim3(256), 0, stream, K, n, S,
                       R, pL, pU, plb, lA, lbias, cL, cU, keys);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" int gh_child_keys_f64(int K, int n, int S, int R, const double *plb,
                                 const double *lA, const double *lbias, const double *cL,
                                 const double *cU, double *keys, hipStream_t stream)
{
    if (K == 0) return GH_OK;
    if (K < 0 || n < 1 || S < 1 || R < 1 || R > S || !plb || !lA || !lbias || !cL || !cU || !keys)
        return GH_EINVAL;
    hipLaunchKernelGGL(child_keys_kernel, dim3((2 * K + 3) / 4), dim3(256), 0, stream, K, n, S,
                       R, plb, lA, lbias, cL, cU, keys);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

extern "C" size_t gh_split_append_workspace_bytes(int K)
{
    if (K < 0) return 0;
    const size_t nblk = ((size_t)K + kSplitBlock - 1) / kSplitBlock;
    return sizeof(double) * 2 * (size_t)K + sizeof(int32_t) * (3 * (size_t)K + 2 * nblk) + 256;
}

namespace {

int split_append_impl(int K, int n, int S, int R, const double *pL, const double *pU,
                      const int64_t *prows, const double *plb, const double *lA,
                      const double *lbias, const double *thr, const int32_t *cidx, double *FL,
                      double *FU, double *Flb, int64_t base, int64_t cap, int64_t *order_out,
                      int64_t dead, int64_t *total, double *disc_min, void *ws,
                      hipStream_t stream)
{
    if (K < 0 || n < 1 || S < 1 || base < 0 || !total || !disc_min) return GH_EINVAL;
    if (lA && (R < 1 || R > S || !lbias)) return GH_EINVAL;
    if (K == 0) {
        hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, stream, total, (int64_t)0);
        return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
    }
    if (!pL || !pU || !plb || !thr || !FL || !FU || !Flb || !order_out || !ws) return GH_EINVAL;
    if (base + 2 * (int64_t)K > cap) return GH_EINVAL;
    const int nblk = (K + kSplitBlock - 1) / kSplitBlock;
    double *ckey = (double *)ws;
    int32_t *sidx = (int32_t *)(ckey + 2 * (size_t)K);
    int32_t *keep = sidx + K;
    int32_t *cnt = keep + 2 * (size_t)K;
    if (n <= 64 && S <= 64)
        hipLaunchKernelGGL(split_keep16_kernel, dim3(nblk), dim3(256), 0, stream, K, n, S, R, pL,
                           pU, prows, plb, lA, lbias, thr, cidx, ckey, sidx, keep, cnt, disc_min);
    else
        hipLaunchKernelGGL(split_keep_wave_kernel, dim3(nblk), dim3(1024), 0, stream, K, n, S, R,
                           pL, pU, prows, plb, lA, lbias, thr, cidx, ckey, sidx, keep, cnt,
                           disc_min);
    hipLaunchKernelGGL(filter_scan_kernel, dim3(1), dim3(1024), 0, stream, 2 * nblk, cnt, total);
    if (n <= 64 && S <= 64)
        hipLaunchKernelGGL(split_copy16_kernel, dim3(nblk), dim3(256), 0, stream, K, n, S, pL, pU,
                           prows, plb, ckey, sidx, keep, cnt, total, FL, FU, Flb, base, order_out,
                           dead);
    else
        hipLaunchKernelGGL(split_copy_kernel, dim3(nblk), dim3(1024), 0, stream, K, n, S, pL, pU,
                           prows, plb, ckey, sidx, keep, cnt, total, FL, FU, Flb, base, order_out,
                           dead);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

// Scratch of the allocation-free split-append: the workspace layout above
// for the largest K it accepts, in a file-scope device array (one copy per
// GPU, held until the process ends).  Host code never names the array: it
// uses the address hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocMaxParents = 1 << 17;
constexpr size_t kNoAllocBytes =
    sizeof(double) * 2 * (size_t)kNoAllocMaxParents +
    sizeof(int32_t) * (3 * (size_t)kNoAllocMaxParents +
                       2 * (((size_t)kNoAllocMaxParents + kSplitBlock - 1) / kSplitBlock)) +
    256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// The current device's address of `scratch`, looked up once per device.
int scratch_address(void **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return GH_ELAUNCH;
    if (dev < 0 || dev >= kMaxDevices) return GH_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GH_ELAUNCH;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GH_OK;
}

}  // namespace

extern "C" int gh_split_append_f64(int K, int n, int S, int R, const double *pL, const double *pU,
                                   const int64_t *prows, const double *plb, const double *lA,
                                   const double *lbias, const double *thr, double *FL, double *FU,
                                   double *Flb, int64_t base, int64_t cap, int64_t *order_out,
                                   int64_t dead, int64_t *total, double *disc_min, void *ws,
                                   hipStream_t stream)
{
    return split_append_impl(K, n, S, R, pL, pU, prows, plb, lA, lbias, thr, nullptr, FL, FU,
                             Flb, base, cap, order_out, dead, total, disc_min, ws, stream);
}

extern "C" int gh_split_append_idx_f64(int K, int n, int S, int R, const double *pL,
                                       const double *pU, const int64_t *prows, const double *plb,
                                       const double *lA, const double *lbias, const double *thr,
                                       const int32_t *idx, double *FL, double *FU, double *Flb,
                                       int64_t base, int64_t cap, int64_t *order_out,
                                       int64_t dead, int64_t *total, double *disc_min, void *ws,
                                       hipStream_t stream)
{
    if (K > 0 && !idx) return GH_EINVAL;
    return split_append_impl(K, n, S, R, pL, pU, prows, plb, lA, lbias, thr, idx, FL, FU, Flb,
                             base, cap, order_out, dead, total, disc_min, ws, stream);
}

extern "C" int gh_split_append_f64_noalloc(int K, int n, int S, int R, const double *pL,
                                           const double *pU, const int64_t *prows,
                                           const double *plb, const double *lA,
                                           const double *lbias, const double *thr, double *FL,
                                           double *FU, double *Flb, int64_t base, int64_t cap,
                                           int64_t *order_out, int64_t dead, int64_t *total,
                                           double *disc_min, hipStream_t stream)
{
    if (K > kNoAllocMaxParents) return GH_EUNSUPPORTED;
    if (K < 0) return GH_EINVAL;
    void *ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc != GH_OK) return rc;
    return split_append_impl(K, n, S, R, pL, pU, prows, plb, lA, lbias, thr, nullptr, FL, FU,
                             Flb, base, cap, order_out, dead, total, disc_min, ws, stream);
}

extern "C" int gh_split_append_noalloc_max_parents(void) { return kNoAllocMaxParents; }

namespace {

// Any non-finite plane coefficient, bound or threshold (+inf thresholds
// count as the box's bound) sets *flag: the batch then takes the widest
// input (branching_heuristics.py:68-87).
__global__ void __launch_bounds__(256)
sb_finite_kernel(int B, int n, int S, const double *__restrict__ lb,
                 const double *__restrict__ thr, const double *__restrict__ lA,
                 int32_t *__restrict__ flag)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t na = (size_t)B * S * n, nb = (size_t)B * S;
    bool bad = false;
    if (t < na) bad = !isfinite(lA[t]);
    if (t < nb) {
        const double l = lb[t];
        const double tv = thr[t % S];
        const double te = (isinf(tv) && tv > 0) ? l : tv;
        bad = bad || !isfinite(l) || !isfinite(te);
    }
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void set_i32_kernel(int32_t *p, int32_t v) { *p = v; }

// One wave per box: lane i's score sum_s max(|a_si|, t) w_i / 2 summed over
// the rows in order, the largest score with the lowest index, then the
// exhaustive guard against the widest input relative to the root.
__global__ void __launch_bounds__(256)
sb_choice_kernel(int B, int n, int S, const double *__restrict__ xL,
                 const double *__restrict__ xU, const int64_t *__restrict__ rows,
                 const double *__restrict__ lA, double coeff_thresh,
                 const double *__restrict__ root_w, double ratio,
                 const int32_t *__restrict__ flag, int32_t *__restrict__ idx)
{
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= B) return;
    const size_t br = rows ? (size_t)rows[b] : (size_t)b;
    const double *l = xL + br * n, *u = xU + br * n;
    if (*flag) {
        const int w = first_widest(n, lane, l, u);
        if (lane == 0) idx[b] = w;
        return;
    }
    const double *a = lA + (size_t)b * S * n;
    double bs = -INFINITY;
    int bi = -1;
    for (int i = lane; i < n; i += 64) {
        const double w = u[i] - l[i];
        double sc = 0.0;
        for (int s = 0; s < S; ++s) sc = sc + (fmax(fabs(a[(size_t)s * n + i]), coeff_thresh) * w) / 2;
        if (bi < 0 || sc > bs) {
            bs = sc;
            bi = i;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const double s2 = __shfl_xor(bs, o);
        const int i2 = __shfl_xor(bi, o);
        if (i2 >= 0 && (bi < 0 || s2 > bs || (s2 == bs && i2 < bi))) {
            bs = s2;
            bi = i2;
        }
    }
    int best = bi < 0 ? 0 : bi;
    if (root_w && ratio > 0) {
        double rm = -INFINITY, rc = 0.0;
        int wi = -1;
        for (int i = lane; i < n; i += 64) {
            const double r = root_w[i] > 0 ? (u[i] - l[i]) / fmax(root_w[i], 1e-300) : 0.0;
            if (wi < 0 || r > rm) {
                rm = r;
                wi = i;
            }
            if (i == best) rc = r;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const double r2 = __shfl_xor(rm, o);
            const int i2 = __shfl_xor(wi, o);
            if (i2 >= 0 && (wi < 0 || r2 > rm || (r2 == rm && i2 < wi))) {
                rm = r2;
                wi = i2;
            }
        }
        rc = __shfl(rc, best & 63);
        if (rc < ratio * rm) best = wi;
    }
    if (lane == 0) idx[b] = best;
}

}  // namespace

extern "C" int gh_branch_sb_f64(int B, int n, int S, const double *xL, const double *xU,
                                const int64_t *rows, const double *lb, const double *thr,
                                const double *lA, double coeff_thresh, const double *root_w,
                                double min_width_ratio, int32_t *idx, int32_t *flag,
                                hipStream_t stream)
{
    if (B < 0 || n < 1 || S < 1) return GH_EINVAL;
    if (B == 0) return GH_OK;
    if (!xL || !xU || !lb || !thr || !lA || !idx || !flag) return GH_EINVAL;
    hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(1), 0, stream, flag, 0);
    const size_t na = (size_t)B * S * n;
    hipLaunchKernelGGL(sb_finite_kernel, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, stream,
                       B, n, S, lb, thr, lA, flag);
    hipLaunchKernelGGL(sb_choice_kernel, dim3((B + 3) / 4), dim3(256), 0, stream, B, n, S, xL,
                       xU, rows, lA, coeff_thresh, root_w, min_width_ratio, flag, idx);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
