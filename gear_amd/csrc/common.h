// Shared device helpers for the gfx950 kernels.  Every file is built with
// -ffp-contract=off: the only fused multiply-adds are the explicit fma()
// calls, so each arithmetic chain has exactly one meaning.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

#define GH_LANES 16   // lanes that cooperate on one box

namespace gh {

// finite_abs: |v| for finite v, 0 otherwise (infinite terms give vacuous
// bounds elsewhere).
__device__ __forceinline__ double finite_abs(double v) { return isfinite(v) ? fabs(v) : 0.0; }
__device__ __forceinline__ float finite_abs(float v) { return isfinite(v) ? fabsf(v) : 0.0f; }

// widen_bound: bound + sign * err where the bound is finite; an infinite
// error gives the vacuous endpoint.
__device__ __forceinline__ double widen(double bound, double err, double sign)
{
    if (isinf(err) && err > 0) return sign * INFINITY;
    return isfinite(bound) ? bound + sign * err : bound;
}

// One rounded addition of enclosed bounds, widened by u-level |result|
// (k1 = (1 + 16) * 4 * eps).
__device__ __forceinline__ double add_enclosed(double a, double b, double sign, double k1)
{
    double r = a + b;
    return widen(r, finite_abs(r) * k1, sign);
}

// DPP lane move of a double (two 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int2 h = __builtin_bit_cast(int2, v);
    h.x = __builtin_amdgcn_update_dpp(0, h.x, CTRL, 0xF, 0xF, false);
    h.y = __builtin_amdgcn_update_dpp(0, h.y, CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, h);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// xor tree over the 16 lanes of a group: p[l] += p[l^8]; p[l] += p[l^4];
// p[l] += p[l^2]; p[l] += p[l^1].  Every lane ends with the same value.
// The moves are DPP (no LDS): row_ror:8 is exactly lane l^8; after that
// level p[x] == p[x^8], so row_ror:4 delivers a value equal to p[l^4];
// quad_perm [2,3,0,1] and [1,0,3,2] are l^2 and l^1.  Each addition is
// p[l] + p[l^off], the same operands as the oracle's tree.
__device__ __forceinline__ double tree16(double v)
{
    v = v + dpp_f64<0x128>(v);   // row_ror:8
    v = v + dpp_f64<0x124>(v);   // row_ror:4
    v = v + dpp_f64<0x4E>(v);    // quad_perm 2,3,0,1
    v = v + dpp_f64<0xB1>(v);    // quad_perm 1,0,3,2
    return v;
}
__device__ __forceinline__ float tree16(float v)
{
    v = v + dpp_f32<0x128>(v);
    v = v + dpp_f32<0x124>(v);
    v = v + dpp_f32<0x4E>(v);
    v = v + dpp_f32<0xB1>(v);
    return v;
}

// Values of lane l^16 and lane l^32 (gfx950 permlane swaps, no LDS).
__device__ __forceinline__ double xor16_f64(double v)
{
    int2 h = __builtin_bit_cast(int2, v);
    auto lo = __builtin_amdgcn_permlane16_swap((unsigned)h.x, (unsigned)h.x, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((unsigned)h.y, (unsigned)h.y, false, false);
    // after the swap: element 0 holds rows (0,0,2,2), element 1 rows (1,1,3,3)
    bool even_row = ((threadIdx.x >> 4) & 1) == 0;
    int2 r;
    r.x = (int)(even_row ? lo[1] : lo[0]);
    r.y = (int)(even_row ? hi[1] : hi[0]);
    return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double xor32_f64(double v)
{
    int2 h = __builtin_bit_cast(int2, v);
    auto lo = __builtin_amdgcn_permlane32_swap((unsigned)h.x, (unsigned)h.x, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((unsigned)h.y, (unsigned)h.y, false, false);
    // after the swap: element 0 holds halves (lo, lo), element 1 (hi, hi)
    bool low_half = (threadIdx.x & 63) < 32;
    int2 r;
    r.x = (int)(low_half ? lo[1] : lo[0]);
    r.y = (int)(low_half ? hi[1] : hi[0]);
    return __builtin_bit_cast(double, r);
}

// Lane I of this lane's 16-lane row (DPP row_newbcast, one v_mov_b64_dpp).
template <int I>
__device__ __forceinline__ double bcast16_f64(double v)
{
    return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + I, 0xF, 0xF, false);
}

// f(integral_constant<int, I>) for I = 0 .. LIMIT - 1 while I < count (count
// the same in every lane): loops whose lane index must be a compile-time
// constant (DPP controls, register arrays of LIMIT entries).
template <int LIMIT, int I = 0, class F>
__device__ __forceinline__ void for_lanes_upto(int count, F &&f)
{
    if constexpr (I < LIMIT) {
        if (I < count) {
            f(std::integral_constant<int, I>{});
            for_lanes_upto<LIMIT, I + 1>(count, f);
        }
    }
}
template <class F>
__device__ __forceinline__ void for_lanes16(int count, F &&f)
{
    for_lanes_upto<16>(count, f);
}

// The red16 tree when partial l = q + 4v lives in lane row q (lane / 16)
// and register v: levels l^8 and l^4 are register pairs (v^2, v^1), levels
// l^2 and l^1 are lanes ^32 and ^16.  Same operand pairs as the oracle.
__device__ __forceinline__ double tree4x4(const double p[4])
{
    double q0 = p[0] + p[2], q1 = p[1] + p[3];
    double r = q0 + q1;
    r = r + xor32_f64(r);
    r = r + xor16_f64(r);
    return r;
}

// Direct global -> LDS copy of one row's packed tiles (rowstride elements of
// A, then rowstride of B) by the W wavefronts of a workgroup: gfx950's
// global_load_lds writes lane l's 16 bytes at the wave's LDS base + 16 l,
// so no VGPR holds the data.  Asynchronous: the caller waits (vmcnt) and
// synchronizes the workgroup before reading dst.  rowstride * sizeof(T)
// must be a multiple of 16; the row must lie inside A and B.
template <typename T, int W>
__device__ __forceinline__ void stage_row(const T* __restrict__ A, const T* __restrict__ B,
                                          size_t rowstride, int s, T* dst, int wv, int lane)
{
    constexpr int per = 16 / (int)sizeof(T);      // elements per 16-byte piece
    const int chunks = (int)(rowstride / per);
    const T* a = A + (size_t)s * rowstride;
    const T* b = B + (size_t)s * rowstride;
    for (int g = wv; g * 64 < chunks; g += W) {
        const int ch = g * 64 + lane;
        if (ch < chunks) {
            __builtin_amdgcn_global_load_lds(
                (const void*)(a + (size_t)per * ch),
                (__attribute__((address_space(3))) void*)(dst + (size_t)g * 64 * per), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                (const void*)(b + (size_t)per * ch),
                (__attribute__((address_space(3))) void*)(dst + rowstride + (size_t)g * 64 * per),
                16, 0, 0);
        }
    }
}

}  // namespace gh
