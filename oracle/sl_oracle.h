/*
 * CPU oracle for the SpaceLab DEM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's per-step physics
 * (ColliderSingleCore/ColliderSingleCore.cpp:78-288, ball_group.hpp:777-947),
 * operation for operation, in IEEE binary64 with no FMA contraction.  It is
 * the checker for the HIP path: only the tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load it.  The product library never
 * calls it.
 *
 * Pinning (each function lists what checks it):
 *   orc_step_tri / orc_init_conditions : bitwise against golden vectors written
 *       by the compiled reference (tests/golden), plus closed-form checks
 *       (free drift, two-sphere restitution, vdW force = dU/dz, Newton-3).
 *   orc_step_gather : bitwise against orc_step_tri in every state array.
 *   far_gravity = 1 (no reference counterpart; the reference comments the
 *       far-field gravity term out at ColliderSingleCore.cpp:208,227): pinned
 *       by the Kepler two-body closed form (period, energy, angular momentum).
 *   Energy sums (PE) on the gather path: summed per row, then over rows; within
 *       1e-13 relative of the reference's running sum (golden check).
 *
 * Arrays of 3-vectors are AoS xyz (n*3 doubles), the layout of the
 * reference's vec3 arrays (ball_group.hpp:44-50).
 */
#ifndef SL_ORACLE_H
#define SL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double G;      /* gravitational constant (dust_const.hpp:21) */
    double Ha;     /* Hamaker constant (dust_const.hpp:37) */
    double h_min;  /* minimum vdW separation (dust_const.hpp:38) */
    double u_s;    /* sliding friction coefficient (dust_const.hpp:23) */
    double u_r;    /* rolling friction coefficient (dust_const.hpp:24) */
    double kin;    /* loading spring constant (dust_const.hpp:52) */
    double kout;   /* unloading spring constant (dust_const.hpp:53) */
    double dt;     /* time step (dust_const.hpp:51) */
    int far_gravity; /* 0: gravity on contacts only (reference); 1: all pairs */
} orc_params;

typedef struct {
    double PE, KE;
    double p[3];   /* total linear momentum (vector) */
    double L[3];   /* total angular momentum about the origin (vector) */
} orc_energy;

/* Step-0 force and energy pass (ball_group.hpp:777-947).  acc/aacc accumulate
 * from their incoming values (the reference starts from zeroed arrays).
 * dist: triangular N(N-1)/2 array indexed A(A-1)/2+B, or NULL.  When NULL the
 * previous distance is taken equal to the current one (so k = kout for an
 * overlapping pair at step 0); the reference reads uninitialised memory there. */
void orc_init_conditions(int n, const orc_params* P, const double* pos, const double* vel,
                         const double* w, double* acc, double* aacc, const double* R,
                         const double* m, const double* moi, double* dist, orc_energy* e);

/* One sim_one_step (ColliderSingleCore.cpp:78-288), triangular pair loop in the
 * reference order, hysteresis through the triangular dist array (updated). */
void orc_step_tri(int n, const orc_params* P, int write_step, double* pos, double* vel,
                  double* velh, double* w, double* wh, double* acc, double* aacc,
                  const double* R, const double* m, const double* moi, double* dist,
                  orc_energy* e);

/* Same step as an ordered gather: for each ball i, pair contributions are added
 * in ascending j.  The previous distance of a pair is recomputed from pos_prev
 * (positions before this step's drift, written by this call).  nthreads > 1
 * uses OpenMP over i (each i is independent, so results do not depend on it). */
void orc_step_gather(int n, const orc_params* P, int write_step, double* pos, double* vel,
                     double* velh, double* w, double* wh, double* acc, double* aacc,
                     const double* R, const double* m, const double* moi, double* pos_prev,
                     orc_energy* e, int nthreads);

/* Step-0 pass as a gather (pos_prev = pos, i.e. k = kout for overlaps). */
void orc_init_gather(int n, const orc_params* P, const double* pos, const double* vel,
                     const double* w, double* acc, double* aacc, const double* R,
                     const double* m, const double* moi, orc_energy* e, int nthreads);

/* Run: init (if do_init) then steps 1..n_steps-1 as sim_looper does
 * (ColliderSingleCore.cpp:298-331), write step when Step % skip == 0.  Energies of
 * write steps go to e_out[k] (k-th write), capacity e_cap.  mode 0 = triangular,
 * 1 = gather.  Returns the number of write records. */
int orc_run(int n, const orc_params* P, int n_steps, int skip, int do_init, double* pos,
            double* vel, double* velh, double* w, double* wh, double* acc, double* aacc,
            const double* R, const double* m, const double* moi, double* pos_prev,
            orc_energy* e_init, orc_energy* e_out, int e_cap, int mode, int nthreads);

/* Pair kernel pieces exposed for closed-form tests. */
double orc_vdw_force_mag(double Ha, double Ra, double Rb, double h);
double orc_vdw_potential(double Ha, double Ra, double Rb, double h);

#ifdef __cplusplus
}
#endif
#endif
