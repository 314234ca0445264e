/*
 * CPU oracle for the SpaceLab DEM hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Restates, operation for operation and in the same evaluation order:
 *   - sim_one_step            ColliderSingleCore/ColliderSingleCore.cpp:78-288
 *   - init_conditions         ball_group.hpp:777-947
 *   - vec3 arithmetic order   vec3.hpp:44-128,180-184
 * Must be compiled with -ffp-contract=off (no FMA contraction) and without
 * -ffast-math: every +,-,*,/ and sqrt is one IEEE binary64 operation, as in the
 * reference built with /fp:precise.  See sl_oracle.h for what pins each entry.
 */
#include "sl_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } v3;

static inline v3 v3_make(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_neg(v3 a) { return v3_make(-a.x, -a.y, -a.z); }
/* vec3::operator*(double) computes scalar * component (vec3.hpp:66). */
static inline v3 v3_scale(v3 a, double s) { return v3_make(s * a.x, s * a.y, s * a.z); }
static inline v3 v3_div(v3 a, double s) { return v3_make(a.x / s, a.y / s, a.z / s); }
/* vec3::dot and normsquared: (x*x + y*y) + z*z (vec3.hpp:121,127). */
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double v3_norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* vec3::cross (vec3.hpp:122-125). */
static inline v3 v3_cross(v3 a, v3 b)
{
    return v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 v3_load(const double* p, int i) { return v3_make(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void v3_store(double* p, int i, v3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }

/* vdW force magnitude factor (ColliderSingleCore.cpp:130-139, 216-225):
 * Ha/6*64*Ra^3*Rb^3 * ((h+Ra+Rb)/(s*s*t*t)), left to right. */
static inline double vdw_scalar(double Ha, double Ra, double Rb, double h, double h2)
{
    const double twoRah = 2 * Ra * h;
    const double twoRbh = 2 * Rb * h;
    const double s = h2 + twoRah + twoRbh;
    const double t = h2 + twoRah + twoRbh + 4 * Ra * Rb;
    return Ha / 6 * 64 * Ra * Ra * Ra * Rb * Rb * Rb * ((h + Ra + Rb) / (s * s * t * t));
}

/* U_vdw (ColliderSingleCore.cpp:195-202). */
static inline double vdw_pot(double Ha, double Ra, double Rb, double sumRaRb, double h)
{
    const double diffRaRb = Ra - Rb;
    const double z = sumRaRb + h;
    const double two_RaRb = 2 * Ra * Rb;
    const double denom_sum = z * z - (sumRaRb * sumRaRb);
    const double denom_diff = z * z - (diffRaRb * diffRaRb);
    return -Ha / 6 * (two_RaRb / denom_sum + two_RaRb / denom_diff + log(denom_sum / denom_diff));
}

double orc_vdw_force_mag(double Ha, double Ra, double Rb, double h)
{
    return vdw_scalar(Ha, Ra, Rb, h, h * h);
}

double orc_vdw_potential(double Ha, double Ra, double Rb, double h)
{
    return vdw_pot(Ha, Ra, Rb, Ra + Rb, h);
}

typedef struct {
    v3 F;        /* total force on A */
    v3 tA, tB;   /* torques on A and B (contact only) */
    double pe;   /* pair potential energy (write steps) */
    double dist; /* |x_B - x_A| */
    int contact;
} pair_out;

/* One unique pair (A > B) exactly as the body of the reference's pair loop
 * (ColliderSingleCore.cpp:104-250).  pB/pA are the drifted positions, vA/vB and
 * wA/wB the full-step velocities of the previous step (the loop reads O.vel and
 * O.w, not the half-step ones).  old_dist is the pair distance of the previous
 * step (distances[e]). */
static inline void pair_eval(const orc_params* P, int write_step, v3 pA, v3 pB, v3 vA, v3 vB,
                             v3 wA, v3 wB, double RA, double RB, double mA, double mB,
                             double old_dist, pair_out* o)
{
    const double sumRaRb = RA + RB;
    const v3 rVecab = v3_sub(pB, pA);
    const v3 rVecba = v3_neg(rVecab);
    const double dist = v3_norm(rVecab);
    const double overlap = sumRaRb - dist;
    v3 totalForceOnA = v3_make(0, 0, 0);
    o->tA = v3_make(0, 0, 0);
    o->tB = v3_make(0, 0, 0);
    o->pe = 0;
    o->dist = dist;

    if (overlap > 0) {
        double k;
        if (dist >= old_dist) {
            k = P->kout;
        } else {
            k = P->kin;
        }
        const double h = P->h_min;
        const double h2 = h * h;
        const double fv = vdw_scalar(P->Ha, RA, RB, h, h2);
        /* rVecab.normalized() = rVecab / rVecab.norm() (vec3.hpp:128) */
        const v3 vdwForceOnA = v3_scale(v3_div(rVecab, v3_norm(rVecab)), fv);
        /* -k * overlap * .5 * (rVecab / dist) */
        const v3 elasticForceOnA = v3_scale(v3_div(rVecab, dist), -k * overlap * .5);
        /* (G * mA * mB / (dist * dist)) * (rVecab / dist) */
        const v3 gravForceOnA = v3_scale(v3_div(rVecab, dist), P->G * mA * mB / (dist * dist));

        v3 slideForceOnA = v3_make(0, 0, 0);
        v3 rollForceA = v3_make(0, 0, 0);

        const double elastic_force_A_mag = v3_norm(elasticForceOnA);
        const v3 r_a = v3_div(v3_scale(rVecab, RA), sumRaRb);
        const v3 r_b = v3_div(v3_scale(rVecba, RB), sumRaRb);
        const v3 w_diff = v3_sub(wA, wB);

        const v3 d_vel = v3_sub(vB, vA);
        /* d_vel - d_vel.dot(r) * (r / (dist*dist)) - wA.cross(r_a) - wB.cross(r_a)
         * (the second cross product uses r_a, as in ColliderSingleCore.cpp:162) */
        const v3 frame_A_vel_B = v3_sub(
            v3_sub(v3_sub(d_vel, v3_scale(v3_div(rVecab, dist * dist), v3_dot(d_vel, rVecab))),
                   v3_cross(wA, r_a)),
            v3_cross(wB, r_a));

        const double rel_vel_mag = v3_norm(frame_A_vel_B);
        if (rel_vel_mag > 1e-13) {
            slideForceOnA = v3_scale(v3_div(frame_A_vel_B, rel_vel_mag), P->u_s * elastic_force_A_mag);
        }
        const double w_diff_mag = v3_norm(w_diff);
        if (w_diff_mag > 1e-13) {
            /* -u_r * mag * (w_diff).cross(r_a) / (w_diff).cross(r_a).norm() */
            const v3 c = v3_cross(w_diff, r_a);
            rollForceA = v3_div(v3_scale(c, -P->u_r * elastic_force_A_mag), v3_norm(c));
        }
        totalForceOnA = v3_add(v3_add(v3_add(gravForceOnA, elasticForceOnA), slideForceOnA), vdwForceOnA);
        o->tA = v3_cross(r_a, v3_add(slideForceOnA, rollForceA));
        o->tB = v3_cross(r_b, v3_add(v3_neg(slideForceOnA), rollForceA));
        if (write_step) {
            const double U_vdw = vdw_pot(P->Ha, RA, RB, sumRaRb, h);
            double pe = U_vdw + 0.5 * k * overlap * overlap;
            if (P->far_gravity) pe = pe + -P->G * mA * mB / dist;
            o->pe = pe;
        }
        o->contact = 1;
    } else {
        double h = fabs(overlap);
        if (h < P->h_min) {
            h = P->h_min;
        }
        const double h2 = h * h;
        const double fv = vdw_scalar(P->Ha, RA, RB, h, h2);
        const v3 vdwForceOnA = v3_scale(v3_div(rVecab, v3_norm(rVecab)), fv);
        totalForceOnA = vdwForceOnA;
        if (P->far_gravity) {
            /* the term the reference leaves commented out (ColliderSingleCore.cpp:208,227) */
            const v3 gravForceOnA = v3_scale(v3_div(rVecab, dist), P->G * mA * mB / (dist * dist));
            totalForceOnA = v3_add(vdwForceOnA, gravForceOnA);
        }
        if (write_step) {
            double pe = vdw_pot(P->Ha, RA, RB, sumRaRb, h);
            if (P->far_gravity) pe = pe + -P->G * mA * mB / dist;
            o->pe = pe;
        }
        o->contact = 0;
    }
    o->F = totalForceOnA;
}

static inline size_t tri_index(int A, int B)
{
    /* static_cast<unsigned>(A * (A - 1) * .5) + B (ColliderSingleCore.cpp:115) */
    return (size_t)(unsigned)(A * (A - 1) * .5) + (size_t)B;
}

static void energy_zero(orc_energy* e)
{
    memset(e, 0, sizeof(*e));
}

/* Third-pass energy terms for one ball (ColliderSingleCore.cpp:282-285). */
static inline void energy_add_ball(orc_energy* e, v3 pos, v3 vel, v3 w, double m, double moi)
{
    e->KE += .5 * m * v3_dot(vel, vel) + .5 * moi * v3_dot(w, w);
    const v3 p = v3_scale(vel, m);
    e->p[0] += p.x; e->p[1] += p.y; e->p[2] += p.z;
    const v3 L = v3_add(v3_scale(v3_cross(pos, vel), m), v3_scale(w, moi));
    e->L[0] += L.x; e->L[1] += L.y; e->L[2] += L.z;
}

/* First pass (ColliderSingleCore.cpp:82-97). */
static void drift(int n, double dt, double* pos, const double* vel, double* velh, const double* w,
                  double* wh, double* acc, double* aacc)
{
    for (int b = 0; b < n; b++) {
        const v3 vh = v3_add(v3_load(vel, b), v3_scale(v3_scale(v3_load(acc, b), .5), dt));
        v3_store(velh, b, vh);
        v3_store(wh, b, v3_add(v3_load(w, b), v3_scale(v3_scale(v3_load(aacc, b), .5), dt)));
        v3_store(pos, b, v3_add(v3_load(pos, b), v3_scale(vh, dt)));
        v3_store(acc, b, v3_make(0, 0, 0));
        v3_store(aacc, b, v3_make(0, 0, 0));
    }
}

/* Third pass (ColliderSingleCore.cpp:263-287). */
static void kick(int n, double dt, int write_step, const double* pos, double* vel,
                 const double* velh, double* w, const double* wh, const double* acc,
                 const double* aacc, const double* m, const double* moi, orc_energy* e)
{
    for (int b = 0; b < n; b++) {
        const v3 v = v3_add(v3_load(velh, b), v3_scale(v3_scale(v3_load(acc, b), .5), dt));
        const v3 ww = v3_add(v3_load(wh, b), v3_scale(v3_scale(v3_load(aacc, b), .5), dt));
        v3_store(vel, b, v);
        v3_store(w, b, ww);
        if (write_step) energy_add_ball(e, v3_load(pos, b), v, ww, m[b], moi[b]);
    }
}

/* Triangular pair loop of the reference (ColliderSingleCore.cpp:100-256). */
static void pairs_tri(int n, const orc_params* P, int write_step, const double* pos,
                      const double* vel, const double* w, double* acc, double* aacc,
                      const double* R, const double* m, const double* moi, double* dist,
                      double* PE)
{
    pair_out o;
    for (int A = 1; A < n; A++) {
        for (int B = 0; B < A; B++) {
            const v3 pA = v3_load(pos, A), pB = v3_load(pos, B);
            double old_dist;
            if (dist) {
                old_dist = dist[tri_index(A, B)];
            } else {
                old_dist = v3_norm(v3_sub(pB, pA));
            }
            pair_eval(P, write_step, pA, pB, v3_load(vel, A), v3_load(vel, B), v3_load(w, A),
                      v3_load(w, B), R[A], R[B], m[A], m[B], old_dist, &o);
            if (o.contact) {
                v3_store(aacc, A, v3_add(v3_load(aacc, A), v3_div(o.tA, moi[A])));
                v3_store(aacc, B, v3_add(v3_load(aacc, B), v3_div(o.tB, moi[B])));
            }
            if (write_step) *PE += o.pe;
            v3_store(acc, A, v3_add(v3_load(acc, A), v3_div(o.F, m[A])));
            v3_store(acc, B, v3_sub(v3_load(acc, B), v3_div(o.F, m[B])));
            if (dist) dist[tri_index(A, B)] = o.dist;
        }
    }
}

void orc_init_conditions(int n, const orc_params* P, const double* pos, const double* vel,
                         const double* w, double* acc, double* aacc, const double* R,
                         const double* m, const double* moi, double* dist, orc_energy* e)
{
    energy_zero(e);
    pairs_tri(n, P, 1, pos, vel, w, acc, aacc, R, m, moi, dist, &e->PE);
    for (int b = 0; b < n; b++) energy_add_ball(e, v3_load(pos, b), v3_load(vel, b), v3_load(w, b), m[b], moi[b]);
}

void orc_step_tri(int n, const orc_params* P, int write_step, double* pos, double* vel,
                  double* velh, double* w, double* wh, double* acc, double* aacc,
                  const double* R, const double* m, const double* moi, double* dist,
                  orc_energy* e)
{
    energy_zero(e);
    drift(n, P->dt, pos, vel, velh, w, wh, acc, aacc);
    pairs_tri(n, P, write_step, pos, vel, w, acc, aacc, R, m, moi, dist, &e->PE);
    kick(n, P->dt, write_step, pos, vel, velh, w, wh, acc, aacc, m, moi, e);
}

/* Ordered gather: ball i collects the contribution of every pair it belongs to
 * in ascending partner index j, which is the order in which the reference's
 * triangular loop touches acc[i] and aacc[i] (j < i: i is A and adds; j > i:
 * i is B and subtracts).  The pair itself is always evaluated in the
 * reference's orientation (A = max(i,j), B = min(i,j)).  Row PE partials over
 * j < i are summed afterwards in ascending i. */
static void pairs_gather(int n, const orc_params* P, int write_step, const double* pos,
                         const double* vel, const double* w, double* acc, double* aacc,
                         const double* R, const double* m, const double* moi,
                         const double* pos_prev, double* pe_row, int nthreads)
{
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int i = 0; i < n; i++) {
        pair_out o;
        v3 a = v3_load(acc, i), aa = v3_load(aacc, i);
        double pe = 0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const int A = i > j ? i : j;
            const int B = i > j ? j : i;
            const v3 pA = v3_load(pos, A), pB = v3_load(pos, B);
            /* The previous distance only matters in contact; recompute it from the
             * previous positions exactly as the reference computed it then. */
            const double old_dist = v3_norm(v3_sub(v3_load(pos_prev, B), v3_load(pos_prev, A)));
            pair_eval(P, write_step && (i == A), pA, pB, v3_load(vel, A), v3_load(vel, B),
                      v3_load(w, A), v3_load(w, B), R[A], R[B], m[A], m[B], old_dist, &o);
            if (i == A) {
                if (o.contact) aa = v3_add(aa, v3_div(o.tA, moi[A]));
                a = v3_add(a, v3_div(o.F, m[A]));
                if (write_step) pe += o.pe;
            } else {
                if (o.contact) aa = v3_add(aa, v3_div(o.tB, moi[B]));
                a = v3_sub(a, v3_div(o.F, m[B]));
            }
        }
        v3_store(acc, i, a);
        v3_store(aacc, i, aa);
        if (pe_row) pe_row[i] = pe;
    }
}

#include <stdlib.h>

static double sum_rows(int n, const double* pe_row)
{
    double s = 0;
    for (int i = 0; i < n; i++) s += pe_row[i];
    return s;
}

void orc_init_gather(int n, const orc_params* P, const double* pos, const double* vel,
                     const double* w, double* acc, double* aacc, const double* R,
                     const double* m, const double* moi, orc_energy* e, int nthreads)
{
    double* pe_row = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    energy_zero(e);
    pairs_gather(n, P, 1, pos, vel, w, acc, aacc, R, m, moi, pos, pe_row, nthreads);
    e->PE = sum_rows(n, pe_row);
    for (int b = 0; b < n; b++) energy_add_ball(e, v3_load(pos, b), v3_load(vel, b), v3_load(w, b), m[b], moi[b]);
    free(pe_row);
}

void orc_step_gather(int n, const orc_params* P, int write_step, double* pos, double* vel,
                     double* velh, double* w, double* wh, double* acc, double* aacc,
                     const double* R, const double* m, const double* moi, double* pos_prev,
                     orc_energy* e, int nthreads)
{
    double* pe_row = write_step ? (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)) : NULL;
    energy_zero(e);
    memcpy(pos_prev, pos, sizeof(double) * 3 * (size_t)n);
    drift(n, P->dt, pos, vel, velh, w, wh, acc, aacc);
    pairs_gather(n, P, write_step, pos, vel, w, acc, aacc, R, m, moi, pos_prev, pe_row, nthreads);
    if (write_step) e->PE = sum_rows(n, pe_row);
    kick(n, P->dt, write_step, pos, vel, velh, w, wh, acc, aacc, m, moi, e);
    free(pe_row);
}

int orc_run(int n, const orc_params* P, int n_steps, int skip, int do_init, double* pos,
            double* vel, double* velh, double* w, double* wh, double* acc, double* aacc,
            const double* R, const double* m, const double* moi, double* pos_prev,
            orc_energy* e_init, orc_energy* e_out, int e_cap, int mode, int nthreads)
{
    double* dist = NULL;
    size_t np = (size_t)n * (size_t)(n > 0 ? n - 1 : 0) / 2;
    orc_energy e;
    int nw = 0;
    if (mode == 0) {
        dist = (double*)malloc(sizeof(double) * (np > 0 ? np : 1));
        /* distances of the current positions: the pair distances of the previous
         * step for the first call (pos_prev, when given, else pos). */
        const double* pp = pos_prev ? pos_prev : pos;
        for (int A = 1; A < n; A++)
            for (int B = 0; B < A; B++)
                dist[tri_index(A, B)] = v3_norm(v3_sub(v3_load(pp, B), v3_load(pp, A)));
    }
    if (do_init) {
        if (mode == 0) {
            orc_init_conditions(n, P, pos, vel, w, acc, aacc, R, m, moi, dist, &e);
        } else {
            orc_init_gather(n, P, pos, vel, w, acc, aacc, R, m, moi, &e, nthreads);
        }
        if (e_init) *e_init = e;
    }
    for (int step = 1; step < n_steps; step++) {
        const int write_step = skip > 0 && (step % skip == 0);
        if (mode == 0) {
            orc_step_tri(n, P, write_step, pos, vel, velh, w, wh, acc, aacc, R, m, moi, dist, &e);
        } else {
            orc_step_gather(n, P, write_step, pos, vel, velh, w, wh, acc, aacc, R, m, moi, pos_prev, &e, nthreads);
        }
        if (write_step) {
            if (nw < e_cap && e_out) e_out[nw] = e;
            nw++;
        }
    }
    if (mode == 0 && pos_prev) {
        /* leave pos_prev consistent with the gather form: positions before the
         * last drift are not kept by the triangular form, so report none. */
    }
    free(dist);
    return nw;
}
