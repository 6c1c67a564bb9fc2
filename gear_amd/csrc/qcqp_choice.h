// The FP64 path's McCormick plane choice of one element, for the FP32-sound
// kernels' near-tie cases.
#pragma once
#include "common.h"

namespace gh {

// The FP64 kernel's plane choice for element j of row s of box b: the same
// FMA chains over i (qcqp_crown.hip phase 1) and the same selection
// arithmetic, so the result is the FP64 path's r bit for bit.
__device__ __forceinline__ bool fp64_choice(int n, const double* __restrict__ Qrow,
                                         const double* __restrict__ lo,
                                         const double* __restrict__ hi, int j, double kn)
{
    double c = 0.0, dv = 0.0, mg = 0.0;
    for (int i = 0; i < n; ++i) {
        double l = lo[i], h = hi[i];
        double x0 = (h + l) / 2.0, ep = (h - l) / 2.0;
        double mx = finite_abs(fmax(-l, h));
        double q = Qrow[i], qa = fabs(q);
        c = fma(q, x0, c);
        dv = fma(qa, ep, dv);
        mg = fma(qa, mx, mg);
    }
    double e = mg * kn;
    double yl = widen((c - dv) + 0.0, e, -1.0);
    double yu = widen((c + dv) + 0.0, e, 1.0);
    double xl = lo[j], xu = hi[j];
    double a1 = yl, b1 = xl, c1 = (-xl) * yl;
    double a2 = yu, b2 = xu, c2 = (-xu) * yu;
    double m1 = a1 * (a1 >= 0 ? xl : xu);
    m1 = m1 + b1 * (b1 >= 0 ? yl : yu);
    m1 = m1 + c1;
    double m2 = a2 * (a2 >= 0 ? xl : xu);
    m2 = m2 + b2 * (b2 >= 0 ? yl : yu);
    m2 = m2 + c2;
    return m1 >= m2;
}

}  // namespace gh
