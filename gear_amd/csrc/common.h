// Shared device helpers for the gfx950 kernels.  Every file is built with
// -ffp-contract=off: the only fused multiply-adds are the explicit fma()
// calls, so each arithmetic chain has exactly one meaning.
#pragma once
#include <hip/hip_runtime.h>
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

// xor tree over the 16 lanes of a group: p[l] += p[l^8], p[l^4], p[l^2],
// p[l^1].  Every lane of the group ends with the same value.
__device__ __forceinline__ double tree16(double v)
{
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, GH_LANES);
    return v;
}
__device__ __forceinline__ float tree16(float v)
{
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, GH_LANES);
    return v;
}

}  // namespace gh
