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
// Mapping.  One wavefront per 16 boxes; lane 16q + r owns box r and its
// variable indices q + 4m.  Q is packed once per problem into MFMA operand
// order (gh_qcqp_pack_f64): every A operand is one coalesced, L2-resident
// load, and the B operands are the lane's own box vectors, so nothing goes
// through LDS.  Per row s:
//   1. IBP of y_s = Q_s x on FP64 MFMA (16x16x4): D[j][box], contraction
//      over i; centre, radius and magnitude share each Q tile;
//   2. McCormick plane choice and coefficients and the interval of z, per
//      (box, j) in the lane that owns them;
//   3. backward A = beta^T Q_s on MFMA with the lane's beta as B operand;
//   4. merges, A67 charges and concretization: partial sums per lane row q
//      and register v (index class q + 4v mod 16), then the red16 tree
//      (register pairs for l^8, l^4; lane swaps for l^2, l^1).
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

// D = C + A * B on four 4x4x4 FP64 blocks.  Block b: A_b[i][k] in lane
// 16k + 4b + i, B_b[k][j] in lane 16k + 4b + j, D_b[i][j] in lane
// 16i + 4b + j -- so with A replicated over the four blocks (lane 16k + 4b + i
// holding row i) lane l gets D[l / 16][l % 16], the 16x16x4 layout's element
// v = 0, with the lane's own B operand.  Each output is the FMA chain over
// its 4 k in order.
__device__ __forceinline__ double mfma4(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// The last 16-row tile c = NCH - 1 holds V = KT - 4 (NCH - 1) indices per
// lane row (m = 4c + v, rows 16c + 4v .. 16c + 4v + 3 for v < V).  With
// V < 4 each class v is packed as its own tile, replicated over the four
// 4x4 blocks, and run on mfma4 (~1/6 of a 16x16x4 each); tail_v gives V,
// or 0 where the tile stays on 16x16x4 (V = 4, or the packed row of KT = 15
// would not fit the staged kernel's double-buffered LDS).
__host__ __device__ constexpr int tail_v(int kt)
{
    const int nch = (kt + 3) / 4, v = kt - 4 * (nch - 1);
    return (v < 4 && (size_t)(nch - 1 + v) * kt * 64 * 32 <= 160 * 1024) ? v : 0;
}
// Tiles per packed row: the regular 16x16x4 tiles, then the tail classes.
__host__ __device__ constexpr int packed_tiles(int kt)
{
    return (kt + 3) / 4 - (tail_v(kt) ? 1 : 0) + tail_v(kt);
}

// Packed Q operands (built once per problem by gh_qcqp_pack_f64):
//   QB[s][c][t][l] = Q[s][16c + l%16][4t + l/16]   IBP B operand (Q^T tiles)
//   QA[s][c][t][l] = Q[s][4t + l/16][16c + l%16]   backward B operand
// zero outside 0 <= row, col < n; c < NCH, t < KT = ceil(n / 4).  Each MFMA
// B operand is then one coalesced 512-byte load per wave.
__global__ void qcqp_pack_kernel(int n, int S, int nch, int kt, const double* __restrict__ Q,
                                 double* __restrict__ QB, double* __restrict__ QA)
{
    const int nt = packed_tiles(kt), nr = tail_v(kt) ? nch - 1 : nch;
    size_t per = (size_t)nt * kt * 64;
    size_t total = (size_t)S * per;
    size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= total) return;
    int s = (int)(u / per);
    size_t r = u - (size_t)s * per;
    int c = (int)(r / ((size_t)kt * 64));
    r -= (size_t)c * kt * 64;
    int t = (int)(r / 64), l = (int)(r - (size_t)t * 64);
    const double* Qs = Q + (size_t)s * n * n;
    // a tail class v's lane l carries row 16 (NCH - 1) + 4v + l % 4
    // (replicated per block); a regular tile's row 16c + l % 16
    const int row = c < nr ? 16 * c + (l & 15) : 16 * nr + 4 * (c - nr) + (l & 3);
    int j = row, i = 4 * t + (l >> 4);
    QB[u] = (j < n && i < n) ? Qs[(size_t)j * n + i] : 0.0;
    int j2 = 4 * t + (l >> 4), i2 = row;
    QA[u] = (j2 < n && i2 < n) ? Qs[(size_t)j2 * n + i2] : 0.0;
}

// The lane's box values: endpoints xl, xu at indices q + 4m; the centre,
// radius and magnitude of an index are recomputed where used (registers
// are the kernel's scarce resource).
template <int KT>
struct BoxVals {
    double xl[KT], xu[KT];
    __device__ __forceinline__ double x0(int m) const { return (xu[m] + xl[m]) / 2.0; }
    __device__ __forceinline__ double ep(int m) const { return (xu[m] - xl[m]) / 2.0; }
    __device__ __forceinline__ double mx(int m) const
    {
        return gh::finite_abs(fmax(-xl[m], xu[m]));
    }
};

// Per-box state of one row carried from its forward half (IBP and the plane
// choice) to its backward half (backward product, merges, concretization).
template <int KT>
struct RowState {
    double acm[KT], beta[KT];
    double ql, qu, gsum, own;
};

// The plane work of index m (j = q + 4m) of box r from the IBP products
// cc, dv, mg: the interval of y_j, the interval of z_j and the
// coeff-aware McCormick lower plane; partial sums in the lane's class v.
template <int KT>
__device__ __forceinline__ void plane_item(int n, int m, int j, double cc, double dv, double mg,
                                           const double* cs, const BoxVals<KT>& bx,
                                           const QcqpConsts& k, RowState<KT>& st, double& p_ql,
                                           double& p_qu, double& p_g, double& p_own)
{
    st.acm[m] = 0.0;
    st.beta[m] = 0.0;
    if (j < n) {
        double e = mg * k.kn;
        double yl = gh::widen((cc - dv) + 0.0, e, -1.0);
        double yu = gh::widen((cc + dv) + 0.0, e, 1.0);
        double xlj = bx.xl[m], xuj = bx.xu[m];
        // interval of z_j (bivariate.py:563-597)
        double r0 = xlj * yl, r1 = xlj * yu, r2 = xuj * yl, r3 = xuj * yu;
        p_ql = p_ql + fmin(fmin(r0, r1), fmin(r2, r3));
        p_qu = p_qu + fmax(fmax(r0, r1), fmax(r2, r3));
        // coeff-aware McCormick lower plane (bivariate.py:190-261, 40-75)
        double a1 = yl, b1 = xlj, c1 = (-xlj) * yl;
        double a2 = yu, b2 = xuj, c2 = (-xuj) * yu;
        double m1 = a1 * (a1 >= 0 ? xlj : xuj);
        m1 = m1 + b1 * (b1 >= 0 ? yl : yu);
        m1 = m1 + c1;
        double m2 = a2 * (a2 >= 0 ? xlj : xuj);
        m2 = m2 + b2 * (b2 >= 0 ? yl : yu);
        m2 = m2 + c2;
        double rr = (m1 >= m2) ? 1.0 : 0.0;
        double t = (yl - yu) * rr;
        double al = t + yu;
        t = (xlj - xuj) * rr;
        double be = t + xuj;
        double pp = yu * xuj, qq = yl * xlj;
        t = (pp - qq) * rr;
        double ga = t - pp;
        st.acm[m] = cs[j] + al;
        st.beta[m] = be;
        p_g = p_g + ga;
        p_own = fma(fabs(be), k.hsn * mg, p_own);
    }
}

// Forward half of row s.  qb points at the row's packed IBP tiles (+ lane);
// G IBP tiles are accumulated at a time; with a tail tile (tail_tile(KT))
// the last tile runs on mfma4.
template <int KT, int G>
__device__ __forceinline__ void row_forward(int n, const double* qb, const double* cs,
                                            const BoxVals<KT>& bx, int lq,
                                            const QcqpConsts& k, RowState<KT>& st)
{
    constexpr int NCH = (KT + 3) / 4;            // 16-wide tiles of j (and of i)
    constexpr int NM = KT;                       // per-lane indices q + 4m
    constexpr int TV = tail_v(KT);               // tail classes on mfma4
    constexpr int NR = TV ? NCH - 1 : NCH;       // tiles on 16x16x4
    double p_ql[4], p_qu[4], p_g[4], p_own[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) p_ql[v] = p_qu[v] = p_g[v] = p_own[v] = 0.0;

#pragma unroll
    for (int c0 = 0; c0 < NR; c0 += G) {
        // ---- 1. IBP of y_s for tiles c0..c0+G-1: D[j][box] = sum_i
        // Q[j][i] x[box][i] on MFMA (A = Q tile, B = the lane's x at
        // index q + 4t); centre, radius and magnitude as in
        // concretize_func.py:548-551, linear.py:607-653,
        // backward_bound.py:739-741.
        d4 Dc[G], Dd[G], Dm[G];
#pragma unroll
        for (int g = 0; g < G; ++g) Dc[g] = Dd[g] = Dm[g] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            const double x0 = bx.x0(t), ep = bx.ep(t), mx = bx.mx(t);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (c0 + g < NR) {
                    double q = qb[((size_t)(c0 + g) * KT + t) * 64];
                    double qa_ = fabs(q);
                    Dc[g] = mfma16(q, x0, Dc[g]);
                    Dd[g] = mfma16(qa_, ep, Dd[g]);
                    Dm[g] = mfma16(qa_, mx, Dm[g]);
                }
            }
            if (t % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
        // ---- 2. Elementwise: item v of tile c is index j = q + 4m,
        // m = 4c + v, of box r.
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = c0 + g, m = 4 * c + v;
                if (c < NR && m < NM)
                    plane_item<KT>(n, m, lq + 4 * m, Dc[g][v], Dd[g][v], Dm[g][v], cs, bx, k, st,
                                   p_ql[v], p_qu[v], p_g[v], p_own[v]);
            }
        }
    }
    if constexpr (TV > 0) {
        // the last tile: indices m = 4 NR + v, v < TV, one packed class each
        double tc[TV], td[TV], tm[TV];
#pragma unroll
        for (int v = 0; v < TV; ++v) tc[v] = td[v] = tm[v] = 0.0;
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            const double x0 = bx.x0(t), ep = bx.ep(t), mx = bx.mx(t);
#pragma unroll
            for (int v = 0; v < TV; ++v) {
                const double q = qb[((size_t)(NR + v) * KT + t) * 64];
                const double qa_ = fabs(q);
                tc[v] = mfma4(q, x0, tc[v]);
                td[v] = mfma4(qa_, ep, td[v]);
                tm[v] = mfma4(qa_, mx, tm[v]);
            }
            if (t % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int v = 0; v < TV; ++v) {
            const int m = 4 * NR + v;
            plane_item<KT>(n, m, lq + 4 * m, tc[v], td[v], tm[v], cs, bx, k, st, p_ql[v], p_qu[v],
                           p_g[v], p_own[v]);
        }
    }
    st.ql = gh::tree4x4(p_ql);
    st.qu = gh::tree4x4(p_qu);
    st.gsum = gh::tree4x4(p_g);
    st.own = gh::tree4x4(p_own);
}

// Backward half of row s.  qa points at the row's packed backward tiles
// (+ lane).
template <int KT>
__device__ __forceinline__ void row_backward(int n, int S, int s, const double* qa,
                                             const double* cs, double ds,
                                             const BoxVals<KT>& bx, int lq, int box,
                                             bool live, double* lb, double* lA, double* lbias,
                                             int PR, const QcqpConsts& k, const RowState<KT>& st)
{
    constexpr int NCH = (KT + 3) / 4;
    constexpr int NM = KT;
    constexpr int TV = tail_v(KT);               // tail classes on mfma4
    constexpr int NR = TV ? NCH - 1 : NCH;       // tiles on 16x16x4
    // ---- 3. Backward product: D[i][box] = sum_j Q[j][i] beta[box][j]
    // (A = Q^T tile c, B = the lane's beta at index q + 4t); the tail
    // classes (indices m = 4 NR + v) on mfma4.
    d4 Da[NCH];
    double Dt[TV > 0 ? TV : 1];
#pragma unroll
    for (int c = 0; c < NCH; ++c) Da[c] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < TV; ++v) Dt[v] = 0.0;
#pragma unroll
    for (int t = 0; t < KT; ++t) {
#pragma unroll
        for (int c = 0; c < NR; ++c)
            Da[c] = mfma16(qa[((size_t)c * KT + t) * 64], st.beta[t], Da[c]);
#pragma unroll
        for (int v = 0; v < TV; ++v)
            Dt[v] = mfma4(qa[((size_t)(NR + v) * KT + t) * 64], st.beta[t], Dt[v]);
        // keep the operand loads of later steps from being hoisted here
        // (they would hold KT * NCH registers)
        if (t % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int v = 0; v < TV; ++v) Da[NCH - 1][v] = Dt[v];

    // ---- 4. Merges, charges and concretization of box r, in two passes
    // over the lane's indices to bound the live partial sums.
    double p_lc[4], p_ldv[4], p_lmg[4], p_m1[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) p_lc[v] = p_ldv[v] = p_lmg[v] = p_m1[v] = 0.0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = 4 * c + v, i = lq + 4 * m;
            if (m < NM && i < n) {
                const double x0 = bx.x0(m), ep = bx.ep(m), mx = bx.mx(m);
                double csi = cs[i];
                double ca = fabs(csi);
                p_lc[v] = fma(csi, x0, p_lc[v]);
                p_ldv[v] = fma(ca, ep, p_ldv[v]);
                p_lmg[v] = fma(ca, mx, p_lmg[v]);
                p_m1[v] = fma(fabs(st.acm[m]), mx, p_m1[v]);
            }
        }
    }
    const double lc = gh::tree4x4(p_lc), ldv = gh::tree4x4(p_ldv), lmg = gh::tree4x4(p_lmg);
    const double merge1 = gh::tree4x4(p_m1);
    double p_m2[4], p_base[4], p_et[4], p_ma[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) p_m2[v] = p_base[v] = p_et[v] = p_ma[v] = 0.0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = 4 * c + v, i = lq + 4 * m;
            if (m < NM && i < n) {
                const double x0 = bx.x0(m), ep = bx.ep(m), mx = bx.mx(m);
                double a = st.acm[m] + Da[c][v];
                p_m2[v] = fma(fabs(a), mx, p_m2[v]);
                p_base[v] = fma(a, x0, p_base[v]);
                p_et[v] = fma(fabs(a), ep, p_et[v]);
                p_ma[v] = fma(fabs(a), mx, p_ma[v]);
                if (lA && live && s < PR) lA[((size_t)box * PR + s) * n + i] = a;
            }
        }
    }
    const double merge2 = gh::tree4x4(p_m2);
    const double base = gh::tree4x4(p_base), et = gh::tree4x4(p_et), ma = gh::tree4x4(p_ma);
    if (live && lq == 0) {
        double le = (lmg + fabs(ds)) * k.kn;
        double ll = gh::widen((lc - ldv) + ds, le, -1.0);
        double lu = gh::widen((lc + ldv) + ds, le, 1.0);
        double mag_add = (0.0 + gh::finite_abs(fmax(fabs(st.ql), fabs(st.qu))))
                         + gh::finite_abs(fmax(fabs(ll), fabs(lu)));
        double add_err = mag_add * k.k2;
        double a = 0.0;
        a = gh::add_enclosed(a, 0.0 - add_err, -1.0, k.k1);                    // f
        a = gh::add_enclosed(a, 0.0, -1.0, k.k1);                              // q
        double vec = (0.0 + k.hs * lmg) + k.hs * fabs(ds);
        a = gh::add_enclosed(a, ds - vec, -1.0, k.k1);                         // l
        a = gh::add_enclosed(a, st.gsum, -1.0, k.k1);                          // z
        a = gh::add_enclosed(a, 0.0 - k.h1 * merge1, -1.0, k.k1);              // u
        a = gh::add_enclosed(a, 0.0, -1.0, k.k1);                              // v
        a = gh::add_enclosed(a, 0.0 - (st.own + k.h1 * merge2), -1.0, k.k1);   // y
        double bias = gh::widen(a, gh::finite_abs(a) * k.kn, -1.0);
        double root = gh::widen(base - et, ma * k.kn, -1.0);
        if (!isfinite(root)) root = -INFINITY;
        lb[(size_t)box * S + s] = gh::add_enclosed(bias, root, -1.0, k.k1);
        if (lbias && s < PR) lbias[(size_t)box * PR + s] = a;
    }
}

// One wavefront per 16 boxes, one box per lane column: lane (q, r) = lane
// 16q + r owns box r and the variable indices q + 4m, m < KT = ceil(n/4).
// The MFMAs take Q tiles as the A operand and the lane's box vectors as the
// B operand, so every per-box vector (box rows, Cm + alpha, beta) stays in
// the lane's registers.
//
// W wavefronts share a workgroup.  W = 1: Q tiles are read from L2.  W = 8:
// the workgroup's LDS holds the current row's Q tiles while the next row's
// are copied in (double buffer, direct global -> LDS copies), two waves per
// SIMD.  FP64 MFMA and FP64 VALU work of one SIMD do not overlap on gfx950
// (a wave running MFMAs and a wave running FMAs on one SIMD take the sum of
// their separate times), so the second wave hides latency only; the row's
// time is its MFMA cycles plus its VALU cycles plus stalls.  IBP tiles are
// accumulated one or two at a time to bound the live registers.  Each box
// row's arithmetic and its order are the same in every form.
template <int KT, int W>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(2))) void qcqp_crown_f64_kernel(
    int B, int n, int S, const double* __restrict__ QB, const double* __restrict__ QA,
    const double* __restrict__ Cm, const double* __restrict__ d, const double* __restrict__ xL,
    const double* __restrict__ xU, const int64_t* __restrict__ rows, double* __restrict__ lb,
    double* __restrict__ lA, double* __restrict__ lbias, int PR, QcqpConsts k, int rpb)
{
    constexpr int NCH = (KT + 3) / 4;
    constexpr int NM = KT;
    constexpr int G = (KT > 12) ? 1 : 2;  // IBP tiles per group
    extern __shared__ double qsm[];       // W > 1: 2 x [QB row | QA row]
    const int l = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int lr = l & 15, lq = l >> 4;
    const int box = (blockIdx.x * W + wv) * 16 + lr;
    const int bo = min(box, B - 1);      // rows past B compute on box B-1, never stored
    const bool live = box < B;

    // the box's row of xL / xU: rows[bo] when an index is given
    const size_t xr = rows ? (size_t)rows[bo] : (size_t)bo;
    BoxVals<KT> bx;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        int i = lq + 4 * m;
        bx.xl[m] = (i < n) ? xL[xr * n + i] : 0.0;
        bx.xu[m] = (i < n) ? xU[xr * n + i] : 0.0;
    }

    const size_t rowstride = (size_t)packed_tiles(KT) * KT * 64;
    // this workgroup's rows: the spec rows are independent, so a batch too
    // small to fill the GPU is also split over rows (blockIdx.y)
    const int s0 = blockIdx.y * rpb, s1 = min(S, s0 + rpb);
    RowState<KT> st;
    if constexpr (W > 1) {
        gh::stage_row<double, W>(QB, QA, rowstride, s0, qsm + (s0 & 1) * 2 * rowstride, wv, l);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    for (int s = s0; s < s1; ++s) {
        const double* qb;
        const double* qa;
        if constexpr (W > 1) {
            if (s + 1 < s1)
                gh::stage_row<double, W>(QB, QA, rowstride, s + 1,
                                         qsm + ((s + 1) & 1) * 2 * rowstride, wv, l);
            const double* buf = qsm + (s & 1) * 2 * rowstride;
            qb = buf + l;
            qa = buf + rowstride + l;
        } else {
            qb = QB + (size_t)s * rowstride + l;
            qa = QA + (size_t)s * rowstride + l;
        }
        const double* cs = Cm + (size_t)s * n;
        row_forward<KT, G>(n, qb, cs, bx, lq, k, st);
        row_backward<KT>(n, S, s, qa, cs, d[s], bx, lq, box, live, lb, lA, lbias, PR, k,
                              st);
        if constexpr (W > 1) {
            // the next row's tiles have landed, and every wave is done with
            // this row's buffer before it is refilled
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
}

// Batches of at least this many boxes take the LDS-staged form (128 boxes
// per workgroup; one workgroup per CU at 32,768 boxes).  Below one
// workgroup per CU the rows are split over more workgroups.
constexpr int kStagedMinBoxes = 4096;
constexpr int kStagedWaves = 8;

int compute_units()
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
inline int rows_per_block(int S, long nwg, long target)
{
    long chunks = (target + nwg - 1) / nwg;
    if (chunks < 1) chunks = 1;
    if (chunks > S) chunks = S;
    return (int)((S + chunks - 1) / chunks);
}

template <int KT>
int launch_f64(int B, int n, int S, const double* QB, const double* QA, const double* Cm,
               const double* d, const double* xL, const double* xU, const int64_t* rows,
               double* lb, double* lA, double* lbias, int PR, const QcqpConsts& k,
               hipStream_t stream)
{
    if (B >= kStagedMinBoxes) {
        constexpr int W = kStagedWaves;
        const size_t rowstride = (size_t)packed_tiles(KT) * KT * 64;
        const size_t lds = 4 * rowstride * sizeof(double);
        static bool attr = false;   // once per instantiation (same value every call)
        if (!attr) {
            if (hipFuncSetAttribute((const void*)qcqp_crown_f64_kernel<KT, W>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                return GH_ELAUNCH;
            attr = true;
        }
        const long nwg = (B + 16 * W - 1) / (16 * W);
        const int rpb = rows_per_block(S, nwg, compute_units());
        dim3 grid((unsigned)nwg, (unsigned)((S + rpb - 1) / rpb));
        hipLaunchKernelGGL((qcqp_crown_f64_kernel<KT, W>), grid, dim3(64 * W), lds, stream, B, n,
                           S, QB, QA, Cm, d, xL, xU, rows, lb, lA, lbias, PR, k, rpb);
    } else {
        const long nwg = (B + 15) / 16;
        const int rpb = rows_per_block(S, nwg, 4L * compute_units());
        dim3 grid((unsigned)nwg, (unsigned)((S + rpb - 1) / rpb));
        hipLaunchKernelGGL((qcqp_crown_f64_kernel<KT, 1>), grid, dim3(64), 0, stream, B, n, S,
                           QB, QA, Cm, d, xL, xU, rows, lb, lA, lbias, PR, k, rpb);
    }
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

inline int nch_of(int n) { return (n + 15) / 16; }

}  // namespace

extern "C" size_t gh_qcqp_pack_bytes(int n, int S)
{
    if (n <= 0 || S <= 0 || n > 64) return 0;
    const int kt = (n + 3) / 4;
    return sizeof(double) * 2 * (size_t)S * packed_tiles(kt) * kt * 64;
}

extern "C" int gh_qcqp_pack_f64(int n, int S, const double* Q, void* packed, hipStream_t stream)
{
    if (n <= 0 || S <= 0 || !Q || !packed) return GH_EINVAL;
    if (n > 64) return GH_EUNSUPPORTED;
    int nch = nch_of(n), kt = (n + 3) / 4;
    size_t per = (size_t)S * packed_tiles(kt) * kt * 64;
    double* QB = (double*)packed;
    double* QA = QB + per;
    hipLaunchKernelGGL(qcqp_pack_kernel, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, stream,
                       n, S, nch, kt, Q, QB, QA);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}

namespace {

int crown_f64(int B, int n, int S, const void* packed, const double* Cm, const double* d,
              const double* xL, const double* xU, const int64_t* rows, double* lb, double* lA,
              double* lbias, int plane_rows, hipStream_t stream)
{
    if (B < 0 || n <= 0 || S <= 0) return GH_EINVAL;
    if ((lA || lbias) && (plane_rows < 1 || plane_rows > S)) return GH_EINVAL;
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
    int kt = (n + 3) / 4;
    const double* QB = (const double*)packed;
    const double* QA = QB + (size_t)S * packed_tiles(kt) * kt * 64;
#define GH_KT(K) \
    case K: return launch_f64<K>(B, n, S, QB, QA, Cm, d, xL, xU, rows, lb, lA, lbias, \
                                 plane_rows, k, stream);
    switch (kt) {
        GH_KT(1) GH_KT(2) GH_KT(3) GH_KT(4) GH_KT(5) GH_KT(6) GH_KT(7) GH_KT(8)
        GH_KT(9) GH_KT(10) GH_KT(11) GH_KT(12) GH_KT(13) GH_KT(14) GH_KT(15) GH_KT(16)
    default: return GH_EUNSUPPORTED;
    }
#undef GH_KT
}

}  // namespace

extern "C" int gh_qcqp_crown_f64(int B, int n, int S, const void* packed, const double* Cm,
                                 const double* d, const double* xL, const double* xU, double* lb,
                                 double* lA, double* lbias, int plane_rows, hipStream_t stream)
{
    return crown_f64(B, n, S, packed, Cm, d, xL, xU, nullptr, lb, lA, lbias, plane_rows, stream);
}

extern "C" int gh_qcqp_crown_rows_f64(int B, int n, int S, const void* packed, const double* Cm,
                                      const double* d, const double* xL, const double* xU,
                                      const int64_t* rows, double* lb, double* lA, double* lbias,
                                      int plane_rows, hipStream_t stream)
{
    if (B > 0 && !rows) return GH_EINVAL;
    return crown_f64(B, n, S, packed, Cm, d, xL, xU, rows, lb, lA, lbias, plane_rows, stream);
}
