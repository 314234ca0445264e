// Far-field pair algebra for the fast path.
//
// For a pair that is not touching and whose surface gap h = d - S is at least
// h_min (S = Ri + Rj, d = |r|), the reference's vdW force
//     F = Ha/6*64 Ri^3 Rj^3 (h+S)/(s^2 t^2) * r/d,   s = h^2+2(Ri+Rj)h, t = s+4RiRj
// (ColliderSingleCore.cpp:211-225) simplifies exactly, because h + S = d:
//     s = d^2 - (Ri+Rj)^2,   t = d^2 - (Ri-Rj)^2,   F = C_ij / (s^2 t^2) * r
// with C_ij = (Ha/6*64) Ri^3 Rj^3 = q_i q_j.  No square root is needed, and the
// pair is "near" (routed to the exact path) iff s < 2 S h_min + h_min^2.
// Gravity on all pairs is G m_i m_j r / d^3 = g_i g_j r * rsqrt(d^2)^3.
// The vdW energy of the same pair is U = -Ha/6 (2RiRj/s + 2RiRj/t + ln(s/t))
// (ColliderSingleCore.cpp:231-238 with z = S + h = d).
#pragma once

#include "sl_common.hpp"

namespace sl {

// 1/x: hardware reciprocal (24.4 good bits, measured over 4M operands spanning 2^+-60)
// and one Newton-Raphson step y <- y + y (1 - x y), 48.7 good bits (2.2e-15 relative):
// the size of the far-field algebra's other reassociation errors, far inside the
// fast path's 1e-13 force bar, for two FMAs less per pair than a second step.
#ifndef SL_RCP_NEWTON_STEPS
#define SL_RCP_NEWTON_STEPS 1
#endif
__device__ __forceinline__ double rcp_f64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    for (int k = 0; k < SL_RCP_NEWTON_STEPS; k++) {
        const double e = __builtin_fma(-x, y, 1.0);
        y = __builtin_fma(y, e, y);
    }
    return y;
}

// 1/sqrt(x): hardware estimate and one Newton step y <- y (3/2 - x/2 y^2), about
// 46 good bits; used only for the gravity term, which is many orders of magnitude
// below the vdW term inside a dust aggregate.
__device__ __forceinline__ double rsq_f64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    const double e = __builtin_fma(-hx * y, y, 0.5);
    return __builtin_fma(y, e, y);
}

}  // namespace sl
