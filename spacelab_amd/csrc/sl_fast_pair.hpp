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

// 1/x to binary64 precision: hardware reciprocal (about 23 good bits) and two
// Newton-Raphson steps y <- y + y (1 - x y).
__device__ __forceinline__ double rcp_f64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
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

#include "sl_pair_ref.hpp"

namespace sl {

// 1/sqrt(x) to binary64 precision: hardware estimate and two Newton steps.
__device__ __forceinline__ double rsq_f64_full(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    double e = __builtin_fma(-hx * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-hx * y, y, 0.5);
    return __builtin_fma(y, e, y);
}

// One near pair in the fast path: the same physics, orientation (A, B) and branch
// structure as pair_ref (ColliderSingleCore.cpp:104-250), with each division by a
// length or magnitude replaced by a product with a Newton-refined reciprocal or
// reciprocal square root (about 1 ulp each) -- roughly a quarter of the instructions
// of the correctly rounded quotients.  The write-step energy keeps the reference
// expressions.
template <class ContactData>
__device__ __forceinline__ void pair_fast(const Phys& P, bool write, const PairIn& in, const ContactData& cd,
                                          PairOut& o)
{
    const double S = in.RA + in.RB;
    const D3 r = in.pB - in.pA;
    const double d2 = dot(r, r);
    const double inv_d = rsq_f64_full(d2);
    const double dist = d2 * inv_d;
    const double overlap = S - dist;
    const D3 rhat = r * inv_d;
    o.tA = d3(0, 0, 0);
    o.tB = d3(0, 0, 0);
    o.pe = 0;
    const double grav = P.G * in.mA * in.mB * (inv_d * inv_d);
    if (overlap > 0) {
        const double old_dist = cd.old_dist();
        const double k = (dist >= old_dist) ? P.kout : P.kin;
        const double h = P.h_min;
        const double twoRah = 2 * in.RA * h, twoRbh = 2 * in.RB * h;
        const double s = h * h + twoRah + twoRbh, t = s + 4 * in.RA * in.RB;
        const double st = s * t;
        const double fv = P.Ha / 6 * 64 * (in.RA * in.RA * in.RA) * (in.RB * in.RB * in.RB) * (h + S) *
                          rcp_f64(st * st);
        const double fel = -k * overlap * .5;
        const double f_el_mag = fabs(fel);
        const double invS = rcp_f64(S);
        const D3 r_a = r * (in.RA * invS);
        const D3 r_b = r * (-in.RB * invS);
        const D3 wA = cd.wA(), wB = cd.wB();
        const D3 w_diff = wA - wB;
        const D3 d_vel = cd.vB() - cd.vA();
        const D3 v_rel = ((d_vel - r * (dot(d_vel, r) * (inv_d * inv_d))) - cross(wA, r_a)) - cross(wB, r_a);
        D3 slide = d3(0, 0, 0), roll = d3(0, 0, 0);
        const double vr2 = dot(v_rel, v_rel);
        if (vr2 > 1e-26) slide = v_rel * (P.u_s * f_el_mag * rsq_f64_full(vr2));
        if (dot(w_diff, w_diff) > 1e-26) {
            const D3 c = cross(w_diff, r_a);
            roll = c * (-P.u_r * f_el_mag * rsq_f64_full(dot(c, c)));
        }
        o.F = ((rhat * (grav + fel)) + slide) + rhat * fv;
        o.tA = cross(r_a, slide + roll);
        o.tB = cross(r_b, (-slide) + roll);
        if (write) {
            double pe = vdw_energy(P.Ha, in.RA, in.RB, S, h) + 0.5 * k * overlap * overlap;
            if (P.far_gravity) pe = pe + -P.G * in.mA * in.mB * inv_d;
            o.pe = pe;
        }
        o.contact = 1;
    } else {
        double h = fabs(overlap);
        if (h < P.h_min) h = P.h_min;
        const double twoRah = 2 * in.RA * h, twoRbh = 2 * in.RB * h;
        const double s = h * h + twoRah + twoRbh, t = s + 4 * in.RA * in.RB;
        const double st = s * t;
        double f = P.Ha / 6 * 64 * (in.RA * in.RA * in.RA) * (in.RB * in.RB * in.RB) * (h + S) * rcp_f64(st * st);
        if (P.far_gravity) f = f + grav;
        o.F = rhat * f;
        if (write) {
            double pe = vdw_energy(P.Ha, in.RA, in.RB, S, h);
            if (P.far_gravity) pe = pe + -P.G * in.mA * in.mB * inv_d;
            o.pe = pe;
        }
        o.contact = 0;
    }
}

}  // namespace sl
