// Device evaluation of one sphere pair in the reference's exact arithmetic.
//
// The reference computes every pair (A > B) in the body of its triangular loop,
// ColliderSingleCore/ColliderSingleCore.cpp:104-250 (and the same body in
// init_conditions, ball_group.hpp:784-936).  To be bitwise-equal to it on the
// GPU the device must perform the same IEEE binary64 operations in the same
// order: the expressions below keep the reference's evaluation order
// (left-to-right products, vec3 component order, norm = sqrt((x*x+y*y)+z*z),
// normalized = 3 divisions, vec3.hpp:44-128,180-184), the translation unit is
// compiled with -ffp-contract=off, and division and sqrt are the correctly
// rounded forms clang emits for binary64 on gfx950.
#pragma once

#include "sl_common.hpp"

#pragma clang fp contract(off)

namespace sl {

struct D3 {
    double x, y, z;
};

__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator-(D3 a) { return d3(-a.x, -a.y, -a.z); }
// vector times scalar: the reference multiplies scalar * component (vec3.hpp:66)
__device__ __forceinline__ D3 operator*(D3 a, double s) { return d3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ D3 operator/(D3 a, double s) { return d3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double norm(D3 a) { return __builtin_sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ D3 cross(D3 a, D3 b)
{
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Magnitude factor of the vdW force at surface separation h, in the reference's
// product order: Ha/6*64*Ra*Ra*Ra*Rb*Rb*Rb*((h+Ra+Rb)/(s*s*t*t)),
// s = h^2 + 2Ra h + 2Rb h, t = s + 4 Ra Rb.
__device__ __forceinline__ double vdw_factor(double Ha, double Ra, double Rb, double h, double h2)
{
    const double twoRah = 2 * Ra * h;
    const double twoRbh = 2 * Rb * h;
    const double s = h2 + twoRah + twoRbh;
    const double t = h2 + twoRah + twoRbh + 4 * Ra * Rb;
    return Ha / 6 * 64 * Ra * Ra * Ra * Rb * Rb * Rb * ((h + Ra + Rb) / (s * s * t * t));
}

// vdW potential between spheres at surface separation h (ColliderSingleCore.cpp:195-202).
__device__ __forceinline__ double vdw_energy(double Ha, double Ra, double Rb, double S, double h)
{
    const double dR = Ra - Rb;
    const double z = S + h;
    const double two_RaRb = 2 * Ra * Rb;
    const double ds = z * z - (S * S);
    const double dd = z * z - (dR * dR);
    return -Ha / 6 * (two_RaRb / ds + two_RaRb / dd + ::log(ds / dd));
}

// a / b, correctly rounded, from y = RN(1/b), the correctly rounded reciprocal (one
// binary64 division per divisor): q0 = a y, then two corrections q <- q + (a - b q) y whose
// residual a - b q an FMA forms exactly.  The first correction leaves q within one ulp of
// a/b, so the second residual is exact and, y being within half an ulp of 1/b, the second
// correction rounds to the correctly rounded quotient (Markstein's theorem; P. Markstein,
// IBM J. Res. Dev. 34(1), 1990).  Valid while a, a/b and the residuals are normal numbers.
// A zero quotient can come out +0 where a / b gives -0; every use here adds it to a sum
// that started at +0, whose bits that does not change.
__device__ __forceinline__ double div_rcp(double a, double b, double y)
{
    const double q0 = a * y;
    const double q1 = __builtin_fma(__builtin_fma(-q0, b, a), y, q0);
    return __builtin_fma(__builtin_fma(-q1, b, a), y, q1);
}

struct PairIn {
    D3 pA, pB;    // positions (after this step's drift)
    double RA, RB, mA, mB;
};

// Force on A of a pair that does not touch (overlap = S - dist <= 0) in the reference's
// arithmetic (ColliderSingleCore.cpp:211-227, and gravity on all pairs as pair_ref adds
// it): bitwise pair_ref's far arm, with the three divisions by dist taken from one
// correctly rounded reciprocal (div_rcp).  h receives the clamped gap.
__device__ __forceinline__ D3 far_force_ref(const Phys& P, double RA, double RB, double mA, double mB, const D3& r,
                                            double dist, double overlap, double& h)
{
    h = ::fabs(overlap);                                          // :211-215
    if (h < P.h_min) h = P.h_min;
    const double fv = vdw_factor(P.Ha, RA, RB, h, h * h);
    const double yd = 1.0 / dist;
    const D3 rh = d3(div_rcp(r.x, dist, yd), div_rcp(r.y, dist, yd), div_rcp(r.z, dist, yd));   // r / norm(r)
    D3 F = rh * fv;                                               // :221-225
    if (P.far_gravity) F = F + rh * (P.G * mA * mB / (dist * dist));
    return F;
}

struct PairOut {
    D3 F;         // force on A
    D3 tA, tB;    // torques on A and B (touching pairs)
    double pe;    // pair energy (write steps)
    int contact;
};

// Full pair evaluation.  Velocities and spins are the previous full-step ones
// (the loop reads O.vel and O.w); they and old_dist are only read for touching
// pairs, through the callables, so the far-field arm loads nothing extra.
template <class ContactData>
__device__ __forceinline__ void pair_ref(const Phys& P, bool write, const PairIn& in,
                                         const ContactData& cd, PairOut& o)
{
    const double S = in.RA + in.RB;
    const D3 r = in.pB - in.pA;                   // vector from A to B
    const double dist = norm(r);
    const double overlap = S - dist;
    o.tA = d3(0, 0, 0);
    o.tB = d3(0, 0, 0);
    o.pe = 0;
    if (overlap > 0) {
        const double old_dist = cd.old_dist();
        const double k = (dist >= old_dist) ? P.kout : P.kin;   // :122-126
        const double h = P.h_min;                                // :129
        const double fv = vdw_factor(P.Ha, in.RA, in.RB, h, h * h);
        const D3 f_vdw = (r / norm(r)) * fv;                     // :135-139
        const D3 f_el = (r / dist) * (-k * overlap * .5);        // :142
        const D3 f_g = (r / dist) * (P.G * in.mA * in.mB / (dist * dist));  // :145
        const double f_el_mag = norm(f_el);                      // :154
        const D3 r_a = (r * in.RA) / S;                          // :155
        const D3 r_b = ((-r) * in.RB) / S;                       // :156
        const D3 wA = cd.wA(), wB = cd.wB();
        const D3 w_diff = wA - wB;                               // :157
        const D3 d_vel = cd.vB() - cd.vA();                      // :160
        // :161-162, the second spin term uses r_a as the reference writes it
        const D3 v_rel = ((d_vel - (r / (dist * dist)) * dot(d_vel, r)) - cross(wA, r_a)) - cross(wB, r_a);
        D3 slide = d3(0, 0, 0), roll = d3(0, 0, 0);
        const double v_rel_mag = norm(v_rel);
        if (v_rel_mag > 1e-13) slide = (v_rel / v_rel_mag) * (P.u_s * f_el_mag);       // :166-170
        const double w_diff_mag = norm(w_diff);
        if (w_diff_mag > 1e-13) {                                                   // :174-178
            const D3 c = cross(w_diff, r_a);
            roll = (c * (-P.u_r * f_el_mag)) / norm(c);
        }
        o.F = ((f_g + f_el) + slide) + f_vdw;                    // :181
        o.tA = cross(r_a, slide + roll);                         // :184
        o.tB = cross(r_b, (-slide) + roll);                      // :185
        if (write) {
            double pe = vdw_energy(P.Ha, in.RA, in.RB, S, h) + 0.5 * k * overlap * overlap;  // :195-203
            if (P.far_gravity) pe = pe + -P.G * in.mA * in.mB / dist;
            o.pe = pe;
        }
        o.contact = 1;
    } else {
        double h = ::fabs(overlap);                              // :211-215
        if (h < P.h_min) h = P.h_min;
        const double fv = vdw_factor(P.Ha, in.RA, in.RB, h, h * h);
        o.F = (r / norm(r)) * fv;                                // :221-227
        if (P.far_gravity) o.F = o.F + (r / dist) * (P.G * in.mA * in.mB / (dist * dist));
        if (write) {
            double pe = vdw_energy(P.Ha, in.RA, in.RB, S, h);    // :231-239
            if (P.far_gravity) pe = pe + -P.G * in.mA * in.mB / dist;
            o.pe = pe;
        }
        o.contact = 0;
    }
}

}  // namespace sl
