// Microbenchmark of gfx950's FP64 pipes (measurement only, not part of the
// bound path): in each 512-thread workgroup, waves 0-3 run FP64 MFMA chains
// and waves 4-7 FP64 FMA chains.  Waves w and w + 4 share a SIMD (waves of
// a workgroup are placed round-robin over the CU's four SIMDs), so timing
// mode 1 (MFMA waves only), mode 2 (FMA waves only) and mode 0 (both) shows
// whether the two kinds of work overlap on one SIMD.
#include "common.h"
#include "gearhip_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(512) void pipe_probe_kernel(int mode, int n_mfma, int n_fma,
                                                         double* __restrict__ out)
{
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    double r = 0.0;
    if (wv < 4) {
        if (mode == 2) return;
        d4 a0 = {0.0, 0.0, 0.0, 0.0}, a1 = a0, a2 = a0, a3 = a0;
        const double x = 1.0 + l * 1e-3, y = 0.5;
        for (int i = 0; i < n_mfma; ++i) {   // four independent accumulator chains
            a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, x, a1, 0, 0, 0);
            a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x, x, a2, 0, 0, 0);
            a3 = __builtin_amdgcn_mfma_f64_16x16x4f64(y, y, a3, 0, 0, 0);
        }
        r = a0[0] + a1[1] + a2[2] + a3[3];
    } else {
        if (mode == 1) return;
        double v0 = l, v1 = l + 1, v2 = l + 2, v3 = l + 3, v4 = 1, v5 = 2, v6 = 3, v7 = 4;
        const double m = 0.999999, c = 1e-9;
        for (int i = 0; i < n_fma; ++i) {    // eight independent FMA chains
            v0 = fma(v0, m, c); v1 = fma(v1, m, c); v2 = fma(v2, m, c); v3 = fma(v3, m, c);
            v4 = fma(v4, m, c); v5 = fma(v5, m, c); v6 = fma(v6, m, c); v7 = fma(v7, m, c);
        }
        r = v0 + v1 + v2 + v3 + v4 + v5 + v6 + v7;
    }
    out[(size_t)blockIdx.x * 512 + threadIdx.x] = r;
}

}  // namespace

extern "C" int gh_pipe_probe(int mode, int blocks, int n_mfma, int n_fma, double* out,
                             hipStream_t stream)
{
    if (mode < 0 || mode > 2 || blocks < 1 || n_mfma < 0 || n_fma < 0 || !out) return GH_EINVAL;
    hipLaunchKernelGGL(pipe_probe_kernel, dim3(blocks), dim3(512), 0, stream, mode, n_mfma, n_fma,
                       out);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_ELAUNCH;
}
