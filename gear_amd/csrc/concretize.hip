// Constrained concretization: lower bounds of min c_h . x over a box
// intersected with linear cuts A_k x + b_k <= 0, by the sequential
// one-row dual sweep (auto_LiRPA/concretize_func.py:143-227 _solve_dual_var,
// :269-310 drop_redundant_constraints, :312-400 sort_out_constr_batches,
// :431-850 constraints_solving; the fused Triton sweep of
// concretize_triton.py:159-217 is the reference's GPU form of the same).
//
// One workgroup per box, lane i = input coordinate i (X <= 64).  The waves
// first prepare the box's cut rows (validity, the sort-out test on the box,
// the centred offset d widened down, redundancy and the row cap) into LDS,
// then take the objective rows h = wave, wave + W, ...  For each kept cut
// the multiplier is the breakpoint where the slope of the concave one-row
// dual changes sign: breakpoints q_i = -(c_i / a_i) are sorted by a
// wave-wide bitonic network on (q, index) (integer keys, VALU lane moves),
// the weights |a_i eps_i| are gathered in that order and prefix-summed,
// and a ballot finds the first sign change.
// Every sum has a fixed order (xor trees over 64 lanes, a Hillis-Steele
// scan, row order over cuts) that oracle/concretize.c states too, so the
// results are bit-identical to the oracle.
#include "common.h"
#include "gearhip_internal.h"

namespace {

constexpr int kWaves = 4;
constexpr double kEps = 2.220446049250313080847e-16;

__device__ __forceinline__ double k_round(int n) { return (double)(n + 16) * kEps * 4.0; }

// p[l] + p[l ^ o] for o = 32, 16, 8, 4, 2, 1; every lane ends with the sum.
__device__ __forceinline__ double red64(double v)
{
    v = v + gh::xor32_f64(v);
    v = v + gh::xor16_f64(v);
    return gh::tree16(v);
}

// Sort key of a breakpoint: an unsigned integer in the order of the
// doubles (-0 and +0 equal), every NaN after every number.  Ordering
// (key, index) is ordering (q, index) ascending with NaN last.
__device__ __forceinline__ unsigned long long sort_key(double q)
{
    if (isnan(q)) return ~0ull;
    const unsigned long long u = __builtin_bit_cast(unsigned long long, q + 0.0);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// Value of lane l ^ STRIDE, moved by VALU cross-lane operations (no LDS):
// DPP quad_perm for 1 and 2; for 4, row_ror:4 or row_ror:12 (inside a
// 16-lane row one of them lands on l ^ 4, by bit 2 of l); row_ror:8 for 8;
// the gfx950 permlane swaps for 16 and 32.
template <int STRIDE>
__device__ __forceinline__ int xor_lane_i32(int v)
{
    if constexpr (STRIDE == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    if constexpr (STRIDE == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    if constexpr (STRIDE == 4) {
        // one of the two rotations brings lane l ^ 4; the rotated lane id
        // says which, whatever the rotation's direction
        const int lane = (int)(threadIdx.x & 63);
        const int a = __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false);
        const int b = __builtin_amdgcn_update_dpp(0, v, 0x12C, 0xF, 0xF, false);
        const int from_a = __builtin_amdgcn_update_dpp(0, lane, 0x124, 0xF, 0xF, false);
        return from_a == (lane ^ 4) ? a : b;
    }
    if constexpr (STRIDE == 8) return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
    if constexpr (STRIDE == 16) {
        auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
        return (int)((((threadIdx.x >> 4) & 1) == 0) ? r[1] : r[0]);
    }
    if constexpr (STRIDE == 32) {
        auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
        return (int)(((threadIdx.x & 63) < 32) ? r[1] : r[0]);
    }
    return v;
}

template <int STRIDE>
__device__ __forceinline__ unsigned long long xor_lane_u64(unsigned long long v)
{
    int2 h = __builtin_bit_cast(int2, v);
    h.x = xor_lane_i32<STRIDE>(h.x);
    h.y = xor_lane_i32<STRIDE>(h.y);
    return __builtin_bit_cast(unsigned long long, h);
}

// One compare-exchange level of the bitonic network on (key, index).
template <int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_level(int lane, unsigned long long& key, int& id)
{
    const unsigned long long k2 = xor_lane_u64<STRIDE>(key);
    const int id2 = xor_lane_i32<STRIDE>(id);
    const bool up = (lane & SIZE) == 0;
    const bool lower = (lane & STRIDE) == 0;
    const bool mine_first = key < k2 || (key == k2 && id < id2);
    if ((lower == up) ? !mine_first : mine_first) {
        key = k2;
        id = id2;
    }
    if constexpr (STRIDE > 1) bitonic_level<SIZE, STRIDE / 2>(lane, key, id);
}

template <int SIZE>
__device__ __forceinline__ void bitonic_sort(int lane, unsigned long long& key, int& id)
{
    if constexpr (SIZE > 2) bitonic_sort<SIZE / 2>(lane, key, id);
    bitonic_level<SIZE, SIZE / 2>(lane, key, id);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    int2 h = __builtin_bit_cast(int2, v);
    h.x = __builtin_amdgcn_readlane(h.x, l);
    h.y = __builtin_amdgcn_readlane(h.y, l);
    return __builtin_bit_cast(double, h);
}

// Multiplier of one centred cut (a, d) for objective c (lane-distributed).
__device__ double solve_beta(int X, int lane, double c, double a, double d, double ep)
{
    double q = INFINITY, w = 0.0;
    if (lane < X) {
        q = -(c / a);
        w = fabs(a * ep);
    }
    unsigned long long key = sort_key(q);
    int id = lane;
    bitonic_sort<64>(lane, key, id);
    double s = __shfl(w, id);             // weights in sorted order
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(s, o);
        if (lane >= o) s = s + t;
    }
    const double T = readlane_f64(s, X - 1);
    double cum = __shfl_up(s, 1);
    if (lane == 0) cum = 0.0;
    const double g = (cum - T) + cum - d;
    const unsigned long long m = __ballot(lane <= X && g >= 0);
    int pick;
    if (m) pick = __ffsll((long long)m) - 1;
    else if (X == 64 && ((T - T) + T - d) >= 0) pick = 64;
    else pick = X + 1;
    const int src = __builtin_amdgcn_readlane(id, (pick > 0 && pick <= X) ? pick - 1 : 0);
    const double qs = readlane_f64(q, src);
    const double beta = pick == 0 ? 0.0 : (pick <= X ? qs : INFINITY);
    return beta < 0 ? 0.0 : beta;
}

__global__ __launch_bounds__(64 * kWaves) void constrained_concretize_kernel(
    int H, int K, int X, const double* __restrict__ obj, const double* __restrict__ A,
    const double* __restrict__ b, const double* __restrict__ xL, const double* __restrict__ xU,
    int64_t lda, int64_t ldb, int max_rows, int Kc, double* __restrict__ root,
    int32_t* __restrict__ flags, int32_t* __restrict__ nkept, const double* __restrict__ bias,
    double* __restrict__ lb_io)
{
    extern __shared__ double lds[];
    double* rows = lds;                    // Kc x 64
    double* dk = rows + (size_t)Kc * 64;   // K
    double* score = dk + K;                // K
    int* cand = (int*)(score + K);         // K
    int* keep = cand + K;                  // Kc
    int* sh = keep + Kc;                   // infeasible, not covered, kept rows

    const int bx = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) sh[0] = sh[1] = sh[2] = 0;
    const double lo_i = lane < X ? xL[(size_t)bx * X + lane] : 0.0;
    const double hi_i = lane < X ? xU[(size_t)bx * X + lane] : 0.0;
    const double ep = (hi_i - lo_i) / 2.0, x0 = (hi_i + lo_i) / 2.0;
    const double mxv = lane < X ? gh::finite_abs(fmax(fabs(lo_i), fabs(hi_i))) : 0.0;
    __syncthreads();

    // ---- cut rows: validity, sort-out on the box, centred offset, redundancy
    for (int k = wv; k < K; k += kWaves) {
        const double ai = lane < X ? A[(size_t)bx * lda + (size_t)k * X + lane] : 0.0;
        double bk = b[(size_t)bx * ldb + k];
        const bool valid = isfinite(bk) && __ballot(lane < X && !isfinite(ai)) == 0;
        const double row = valid ? ai : 0.0;
        if (!valid) bk = -1.0;
        const double xt = red64(lane < X ? row * x0 : 0.0) + bk;
        double et = red64(lane < X ? fabs(row) * ep : 0.0);
        const double blo = x0 - ep, bhi = x0 + ep;
        const double err =
            (red64(lane < X ? gh::finite_abs(row) * gh::finite_abs(fmax(fabs(blo), fabs(bhi)))
                            : 0.0) + fabs(bk)) * k_round(X);
        const double mn = gh::widen(xt - et, err, -1.0), mx = gh::widen(xt + et, err, 1.0);
        const double derr =
            (red64(lane < X ? gh::finite_abs(row) * gh::finite_abs(fabs(x0)) : 0.0) + fabs(bk)) *
            k_round(X);
        double d = gh::widen(xt, derr, -1.0);
        if (!valid || !isfinite(d)) {
            d = -1.0;
            et = 0.0;
        }
        const double mv = d + et;
        if (lane == 0) {
            if (mn > 0) atomicOr(&sh[0], 1);
            if (!(mx <= 0)) atomicOr(&sh[1], 1);
            dk[k] = d;
            score[k] = isnan(mv) ? INFINITY : mv;
            cand[k] = !(mv <= 0);
        }
    }
    __syncthreads();

    // ---- kept rows in row order; at most max_rows, the largest scores
    // (ties: lower row first)
    if (wv == 0) {
        int nk = 0;
        for (int k0 = 0; k0 < K; k0 += 64) {
            const int k = k0 + lane;
            bool c = k < K && cand[k];
            if (c && max_rows > 0) {
                const double sc = score[k];
                int rank = 0;
                for (int j = 0; j < K; ++j)
                    rank += (cand[j] && (score[j] > sc || (score[j] == sc && j < k))) ? 1 : 0;
                c = rank < max_rows;
            }
            const unsigned long long m = __ballot(c);
            const int pos = nk + __popcll(m & ((1ull << lane) - 1ull));
            if (c) keep[pos] = k;
            nk += __popcll(m);
        }
        if (lane == 0) sh[2] = nk;
    }
    __syncthreads();
    const int nk = sh[2];
    for (int idx = threadIdx.x; idx < nk * 64; idx += 64 * kWaves) {
        const int t = idx >> 6, i = idx & 63;
        rows[idx] = i < X ? A[(size_t)bx * lda + (size_t)keep[t] * X + i] : 0.0;
    }
    const bool infeasible = sh[0] != 0;
    const bool covered = K == 0 || sh[1] == 0;
    const bool active = !covered && !infeasible;
    if (threadIdx.x == 0) {
        flags[bx] = (infeasible ? 1 : 0) | (covered ? 2 : 0);
        if (nkept) nkept[bx] = active ? nk : 0;
    }
    __syncthreads();

    // ---- the sweep, one objective row per wave at a time
    for (int h = wv; h < H; h += kWaves) {
        double c = lane < X ? obj[((size_t)bx * H + h) * X + lane] : 0.0;
        const double base = red64(lane < X ? c * x0 : 0.0);
        const double et0 = red64(lane < X ? fabs(c) * ep : 0.0);
        const double objerr = red64(lane < X ? gh::finite_abs(c) * mxv : 0.0) * k_round(X);
        double naive = gh::widen(base - et0, objerr, -1.0);
        if (!isfinite(naive)) naive = -INFINITY;
        double out = naive;
        if (active) {
            double cm = gh::finite_abs(c);
            double dual = 0.0, dmag = 0.0;
            for (int t = 0; t < nk; ++t) {
                const double a = rows[t * 64 + lane];
                const double d = dk[keep[t]];
                const double beta = solve_beta(X, lane, c, a, d, ep);
                const double upd = beta * a;
                if (lane < X) {
                    c = c + upd;
                    cm = cm + gh::finite_abs(upd);
                }
                const double bd = beta * d;
                dual = dual + bd;
                dmag = dmag + gh::finite_abs(bd);
            }
            const double et = red64(lane < X ? fabs(c) * ep : 0.0);
            dual = dual - et;
            dmag = dmag + gh::finite_abs(et);
            dmag = dmag + red64(lane < X ? cm * gh::finite_abs(ep) : 0.0);
            double fin = base + dual;
            const double err = objerr + (dmag + gh::finite_abs(base)) * k_round(X + nk);
            fin = gh::widen(fin, err, -1.0);
            if (!isfinite(fin)) fin = naive;
            out = fmax(fin, naive);
        }
        if (lane == 0) {
            const size_t bh = (size_t)bx * H + h;
            if (bias) {
                // the bound: lbias widened and added as the box concretization
                // adds it, never below the given bound; +inf where a cut
                // misses the box (the domain is infeasible)
                double v = INFINITY;
                if (!infeasible) {
                    const double bb = bias[bh];
                    const double bw = gh::widen(bb, gh::finite_abs(bb) * k_round(X), -1.0);
                    v = fmax(gh::add_enclosed(bw, out, -1.0, k_round(1)), lb_io[bh]);
                }
                lb_io[bh] = v;
            } else {
                root[bh] = out;
            }
        }
    }
}

int launch_constrained(int B, int H, int K, int X, const double* obj, const double* A,
                       int64_t lda, const double* b, int64_t ldb, const double* xL,
                       const double* xU, int max_rows, double* root, int32_t* flags,
                       int32_t* nkept, const double* bias, double* lb_io, hipStream_t stream)
{
    if (B < 0 || H < 1 || K < 0 || X < 1) return GH_EINVAL;
    if (K > 0 && (lda < (int64_t)K * X || ldb < K)) return GH_EINVAL;
    if (X > 64 || K > 256) return GH_EUNSUPPORTED;
    const int Kc = max_rows > 0 ? (K < max_rows ? K : max_rows) : K;
    if (Kc > 64) return GH_EUNSUPPORTED;
    if (B == 0) return GH_OK;
    if (!obj || !xL || !xU || !flags || (K > 0 && (!A || !b))) return GH_EINVAL;
    const size_t lds = (size_t)Kc * 64 * sizeof(double) + 2 * (size_t)K * sizeof(double) +
                       ((size_t)K + Kc + 3) * sizeof(int);
    hipLaunchKernelGGL(constrained_concretize_kernel, dim3(B), dim3(64 * kWaves), lds, stream, H,
                       K, X, obj, A, b, xL, xU, lda, ldb, max_rows, Kc, root, flags, nkept, bias,
                       lb_io);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

}  // namespace

extern "C" int gh_constrained_concretize_f64(int B, int H, int K, int X, const double* obj,
                                             const double* A, const double* b, const double* xL,
                                             const double* xU, int max_rows, double* root,
                                             int32_t* flags, int32_t* nkept, hipStream_t stream)
{
    if (B > 0 && !root) return GH_EINVAL;
    return launch_constrained(B, H, K, X, obj, A, (int64_t)K * X, b, K, xL, xU, max_rows, root,
                              flags, nkept, nullptr, nullptr, stream);
}

extern "C" int gh_constrained_bound_f64(int B, int H, int K, int X, const double* lA,
                                        const double* lbias, const double* A, int64_t lda,
                                        const double* b, int64_t ldb, const double* xL,
                                        const double* xU, int max_rows, double* lb,
                                        int32_t* flags, int32_t* nkept, hipStream_t stream)
{
    if (B > 0 && (!lbias || !lb)) return GH_EINVAL;
    return launch_constrained(B, H, K, X, lA, A, lda, b, ldb, xL, xU, max_rows, nullptr, flags,
                              nkept, lbias, lb, stream);
}

// ------------------------------------------------------------ clipping
namespace {

__device__ __forceinline__ double nan_min(double a, double b)
{
    return (isnan(a) || isnan(b)) ? NAN : (a < b ? a : b);
}
__device__ __forceinline__ double nan_max(double a, double b)
{
    return (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b);
}

// One wave per box, lane i = coordinate i.  Box b takes plane set b % P.
// Every value a box's lanes need is read before its outputs are written,
// so oL, oU may alias xL, xU.
__global__ __launch_bounds__(256) void clip_domains_kernel(
    int B, int P, int n, int S, const double* __restrict__ lA, const double* __restrict__ lbias,
    const double* __restrict__ thr, const double* xL, const double* xU, double* oL, double* oU,
    int32_t* __restrict__ state, double und, double cfac)
{
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= B) return;
    const bool on = lane < n;
    const double l = on ? xL[(size_t)b * n + lane] : 0.0;
    const double u = on ? xU[(size_t)b * n + lane] : 0.0;
    const double xh = (u + l) / 2, ep = (u - l) / 2;
    const double mx = gh::finite_abs(fmax(fabs(l), fabs(u)));
    const double* A = lA + (size_t)(b % P) * S * n;
    const double* c = lbias + (size_t)(b % P) * S;
    const double kn = (double)(n + 16) * kEps * 4.0;
    double accU = INFINITY, accL = -INFINITY;
    bool fin = true, cert = false;
    for (int s = 0; s < S; ++s) {
        const double a = on ? A[(size_t)s * n + lane] : 0.0;
        const double cs = c[s], ts = thr[s];
        fin = fin && isfinite(a) && isfinite(cs) && (isfinite(ts) || (isinf(ts) && ts > 0));
        const double t1 = red64(on ? a * xh : 0.0);
        const double t2 = red64(on ? fabs(a) * ep : 0.0);
        const double dm = (t1 - t2) + cs;
        const double bre = (red64(on ? gh::finite_abs(a) * mx : 0.0) + gh::finite_abs(cs)) * kn;
        cert = cert || ((dm - bre) - und > ts);
        const double derr = bre * cfac + und;
        const double lx = a * xh, le = fabs(a) * ep;
        const double cm1 = (dm - lx) + le;
        const double num = ts - cm1;
        const double mag =
            ((gh::finite_abs(ts) + gh::finite_abs(dm)) + gh::finite_abs(lx)) + gh::finite_abs(le);
        const double nerr = (mag * (4 * kEps) + und) + derr;
        double lo = num, hi = num;
        if (isfinite(num)) {
            lo = num - nerr;
            hi = num + nerr;
        }
        const double qlo = lo / a, qhi = hi / a;
        const double cu = nextafter(nan_max(qlo, qhi), INFINITY);
        const double cl = nextafter(nan_min(qlo, qhi), -INFINITY);
        accU = nan_min(accU, a > 0 ? cu : INFINITY);
        accL = nan_max(accL, a < 0 ? cl : -INFINITY);
    }
    fin = __ballot(on && !fin) == 0;
    double nl = l, nu = u;
    int st = 0;
    if (fin) {
        if (cert) {
            if (lane == 0) nl = nextafter(u, INFINITY);
            st = 2;
        } else {
            nu = nan_min(accU, u);
            nl = nan_max(accL, l);
            st = __ballot(on && !(nl <= nu)) ? 3 : 1;
        }
    }
    if (on) {
        oL[(size_t)b * n + lane] = nl;
        oU[(size_t)b * n + lane] = nu;
    }
    if (lane == 0) state[b] = st;
}

}  // namespace

extern "C" int gh_clip_domains_f64(int B, int P, int n, int S, const double* lA,
                                   const double* lbias, const double* thr, const double* xL,
                                   const double* xU, double* oL, double* oU, int32_t* state,
                                   hipStream_t stream)
{
    if (B < 0 || P < 1 || n < 1 || S < 1) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    if (B == 0) return GH_OK;
    if (!lA || !lbias || !thr || !xL || !xU || !oL || !oU || !state) return GH_EINVAL;
    const double und = (double)(n + 16) * 2.2250738585072014e-308 * kEps;
    const double cfac = (double)(n + 4) / (double)(4 * (n + 16));
    hipLaunchKernelGGL(clip_domains_kernel, dim3((B + 3) / 4), dim3(256), 0, stream, B, P, n, S,
                       lA, lbias, thr, xL, xU, oL, oU, state, und, cfac);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
