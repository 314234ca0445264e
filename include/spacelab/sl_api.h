/*
 * spacelab_amd C ABI -- the MI355X (gfx950) implementation of SpaceLab's per-step
 * DEM hot path: the Ball_group all-pairs force loop (van der Waals cohesion,
 * gravity, elastic contact with loading/unloading hysteresis, sliding and rolling
 * friction) and the velocity-Verlet update with its energy and momentum sums.
 *
 * What it replaces in the reference:
 *   sl_init_forces   <- Ball_group::init_conditions        (ball_group.hpp:777-947)
 *   sl_step          <- sim_one_step(bool write_step), called once per Step by
 *                       sim_looper                          (ColliderSingleCore.cpp:78-288, 298-331)
 *   sl_upload        <- the state a Ball_group holds in pos/vel/w/acc/aacc/R/m/moi
 *                       (ball_group.hpp:21-53); AoS xyz, so &O.pos[0].x can be passed
 *   sl_download      <- reading O.pos/O.vel/O.w/O.acc/O.aacc back after a step
 *   sl_params        <- the constants and globals sim_one_step reads
 *                       (dust_const.hpp:21-53: G, Ha, h_min, u_s, u_r, kin, kout, dt)
 *   sl_energy        <- O.PE, O.KE, O.mom, O.ang_mom accumulated on write steps
 *                       (ColliderSingleCore.cpp:190-204, 228-240, 282-285)
 *   sl_batch_*       <- no reference counterpart: many independent Ball_groups
 *                       (collision ensembles) stepped in one launch.
 *
 * Conventions: plain C types only; all host pointers are borrowed for the call.
 * Every call returns 0 on success or a negative SL_E* code, with a message in
 * sl_last_error(ctx); the library never exits or throws.  A context owns its
 * device memory and one HIP stream and is used from one thread at a time;
 * separate contexts may run concurrently.
 */
#ifndef SPACELAB_SL_API_H
#define SPACELAB_SL_API_H

#ifdef __cplusplus
extern "C" {
#endif

#define SL_API_VERSION 1

/* error codes */
#define SL_OK 0
#define SL_E_ARG -1       /* invalid argument (NULL, n out of range, bad parameter) */
#define SL_E_HIP -2       /* HIP runtime error */
#define SL_E_STATE -3     /* call out of order (e.g. step before upload) */
#define SL_E_CAPACITY -4  /* a capacity was exceeded (sl_gen_bpca found no free placement) */
#define SL_E_NOGPU -5     /* no usable gfx950 device */

/* precision */
#define SL_PREC_FP64 0    /* every pair term in binary64 */
#define SL_PREC_MIXED 1   /* far-field vdW/gravity in binary32 per tile, compensated
                             binary64 across tiles; contacts and integration binary64 */
/* order */
#define SL_ORDER_ORDERED 0 /* bitwise the reference's arithmetic and summation order */
#define SL_ORDER_FAST 1    /* each pair once (Newton 3), reassociated far-field algebra */

typedef struct sl_ctx sl_ctx;

typedef struct {
    double G;        /* gravitational constant, cm^3 g^-1 s^-2 */
    double Ha;       /* Hamaker constant, erg */
    double h_min;    /* minimum vdW surface separation, cm */
    double u_s;      /* sliding friction coefficient */
    double u_r;      /* rolling friction coefficient */
    double kin;      /* spring constant while loading (distance shrinking) */
    double kout;     /* spring constant while unloading */
    double dt;       /* time step, s */
    int far_gravity; /* 0: gravity on touching pairs only, as the reference codes it
                        (ColliderSingleCore.cpp:208,227); 1: gravity on all pairs */
    int precision;   /* SL_PREC_* */
    int order;       /* SL_ORDER_* */
} sl_params;

typedef struct {
    double PE;         /* vdW + elastic (+ gravitational when far_gravity) potential */
    double KE;         /* translational + rotational kinetic energy */
    double p[3];       /* total momentum */
    double L[3];       /* total angular momentum about the origin */
    int nonfinite;     /* 1 if any position/velocity/spin became Inf/NaN */
    long long contacts;/* overlapping pairs in this step */
} sl_energy;

/* Reference presets.  preset 0 = "reference": dust_const.hpp verbatim (cor 0.4,
 * u_s 0.1, far_gravity 0).  preset 1 = "config": dust_const.hpp with cor 0.64,
 * u_s 0.3 and far_gravity 1.  kin, kout and dt are left at -1 (set them with
 * sl_calibrate or by hand). */
int sl_params_preset(sl_params* out, int preset);

/* kin, kout and dt for a set of radii and masses.
 *   dt_rule 0: the reference's updateDTK (ball_group.hpp:1313-1329):
 *              kin = kConsts*r_max*v^2, kout = cor*kin, dt = .01*sqrt(...)
 *   dt_rule 1: same kin/kout, dt = t_c/50 with t_c = pi*sqrt(m_min/kin), the
 *              half-period of the contact spring (force (k/2)*overlap). */
int sl_calibrate(sl_params* p, int n, const double* R, const double* m, double velocity,
                 double cor, int dt_rule);

int sl_create(sl_ctx** out, int n_max, const sl_params* p);
int sl_set_params(sl_ctx* ctx, const sl_params* p);

/* Upload a state.  pos3/vel3/w3 are required; acc3/aacc3 may be NULL (zero);
 * pos_prev3 are the positions the previous pair pass used (NULL: pos3, which
 * is what they are after any completed step or init pass).  Resets the step
 * counter to 0. */
int sl_upload(sl_ctx* ctx, int n, const double* pos3, const double* vel3, const double* w3,
              const double* acc3, const double* aacc3, const double* R, const double* m,
              const double* moi, const double* pos_prev3);

/* Step-0 force and energy pass (init_conditions): acc/aacc from the current
 * positions, energies of the current state. */
int sl_init_forces(sl_ctx* ctx, sl_energy* e);

/* n_steps calls of sim_one_step.  The context counts steps from the last upload;
 * a step is a write step when count % write_every == 0 (write_every <= 0: never).
 * Energies of the write steps are stored in order into e_at_writes (capacity
 * e_cap, may be NULL).  Returns the number of write steps taken (>= 0) or an
 * error code. */
int sl_step(sl_ctx* ctx, int n_steps, int write_every, sl_energy* e_at_writes, int e_cap);

int sl_download(sl_ctx* ctx, double* pos3, double* vel3, double* w3, double* acc3,
                double* aacc3);
/* Half-step velocities of the last step (velh, wh in the reference). */
int sl_download_half(sl_ctx* ctx, double* velh3, double* wh3);

int sl_synchronize(sl_ctx* ctx);
long long sl_step_count(const sl_ctx* ctx);
const char* sl_last_error(const sl_ctx* ctx);
void sl_destroy(sl_ctx* ctx);

/* Options (no reference counterpart).
 *   SL_OPT_GRAPHS      1 (default): runs of steps between write steps are captured
 *                      once as HIP graphs and replayed, in both orders (not while
 *                      sl_profile_enable times the launches, nor under a profiling
 *                      tool); 0: every kernel launched directly.  Results are identical
 *                      either way.
 * The others are options of the SL_ORDER_FAST engine:
 *   SL_OPT_TILE_SPLIT  wavefronts per 64x64 pair tile: 0 (default) = chosen by the
 *                      built-in autotuner on first use of a configuration and stored in
 *                      a JSON cache ($SPACELAB_TUNE_CACHE, else ~/.cache/spacelab_amd/
 *                      tune.json; a shipped data/tune.json beside the library's lib/ is
 *                      read too), so later runs of it use the same split; or 1, 2, 4.
 *                      The split changes the summation order inside a tile, so runs
 *                      compare bitwise only at equal split.  sl_get_option reports the
 *                      split in use (0 before the first fast step).
 *   SL_OPT_BALL_WAVES  wavefronts per workgroup of the per-ball kernel: 0 (default) = 4
 *                      when the state has more 64-ball blocks than the device has
 *                      compute units and no system has more than 80 blocks (many small
 *                      systems), else 8; or 4, 8.  Results are identical either way.
 *                      sl_get_option reports the shape the next step uses. */
#define SL_OPT_GRAPHS 1
#define SL_OPT_TILE_SPLIT 2
#define SL_OPT_BALL_WAVES 3
/*   SL_OPT_ORDERED_SHAPE launch shape of the SL_ORDER_ORDERED pair pass (balls per
 *                      workgroup, evaluating wavefronts, partners per chunk): 0 (default)
 *                      = chosen by timing each shape on the first step and stored with
 *                      the tile splits, or 1..n to force one.  Every shape gives the same
 *                      results bit for bit.  sl_get_option reports the shape in use (0
 *                      before the first step). */
#define SL_OPT_ORDERED_SHAPE 4
/*   SL_OPT_TILE_QUEUE  how the fast engine's tile kernel is launched: 0 (default) = chosen by
 *                      the autotuner with the split, 1 = one workgroup per tile group, 2 =
 *                      as many workgroups as the device holds at once pulling tiles from a
 *                      work queue.  Results are identical either way.  sl_get_option reports
 *                      1 or 2 for the launch in use (0 before the first fast step). */
#define SL_OPT_TILE_QUEUE 5
int sl_set_option(sl_ctx* ctx, int option, int value);
int sl_get_option(sl_ctx* ctx, int option, int* value);

/* ---- batched ensembles (no reference counterpart) --------------------------
 * Many independent Ball_groups (for instance one collision each) advanced together:
 * every step is one tile launch and one per-ball launch over all systems, and runs
 * of steps between write steps are replayed as captured HIP graphs.  Each system
 * has its own kin, kout and dt; G, Ha, h_min, u_s, u_r, far_gravity and precision are
 * common.  Only SL_ORDER_FAST is offered.  Arrays are the systems' arrays
 * concatenated in system order (sum of n_per_sys balls). */
typedef struct sl_batch sl_batch;

/* max_balls: capacity in balls over all systems (each system is padded to a
 * multiple of 64 slots internally). */
int sl_batch_create(sl_batch** out, int max_systems, int max_balls, const sl_params* p);
/* kin, kout, dt: per system (n_sys values each), or NULL for the value in p. */
int sl_batch_upload(sl_batch* b, int n_sys, const int* n_per_sys, const double* pos3,
                    const double* vel3, const double* w3, const double* R, const double* m,
                    const double* moi, const double* kin, const double* kout, const double* dt);
/* New per-system kin/kout/dt (NULL keeps the current values); state is kept. */
int sl_batch_set_system_params(sl_batch* b, const double* kin, const double* kout, const double* dt);
/* Step-0 pass; e_sys[s] receives system s's energies (may be NULL). */
int sl_batch_init_forces(sl_batch* b, sl_energy* e_sys);
/* As sl_step for every system; energies of write step w for system s go to
 * e[w * n_sys + s] for the first e_cap write steps. */
int sl_batch_step(sl_batch* b, int n_steps, int write_every, sl_energy* e, int e_cap);
int sl_batch_download(sl_batch* b, double* pos3, double* vel3, double* w3, double* acc3,
                      double* aacc3);
int sl_batch_set_option(sl_batch* b, int option, int value);
int sl_batch_get_option(sl_batch* b, int option, int* value);
long long sl_batch_step_count(const sl_batch* b);
const char* sl_batch_last_error(const sl_batch* b);
void sl_batch_destroy(sl_batch* b);

/* ---- measurement ---------------------------------------------------------- */
typedef struct {
    double pair_ms;        /* summed duration of the pair kernel launches (HIP events); in
                              fast mode the far-field tile sweep */
    long long pair_launches;
    double ball_ms;        /* summed duration of the per-ball kernels; in fast mode they
                              also evaluate the near pairs */
    long long ball_launches;
    long long contacts;    /* overlapping pairs summed over profiled steps */
    long long near_pairs;  /* pairs routed to the near/contact path (fast mode) */
} sl_profile;

/* Record HIP events around each kernel launch of sl_step (on the context's
 * stream) and accumulate; enable=0 stops, reset clears. */
int sl_profile_enable(sl_ctx* ctx, int enable, int reset);
int sl_profile_read(sl_ctx* ctx, sl_profile* out);

/* ---- run control: calibrate_dt's velocity rule ---------------------------- */
/* calc_helpfuls (ball_group.hpp:188-195): out5 = {r_min, r_max, m_total, radius (the largest
 * distance of a centre from the centre of mass; R[0] for one ball), soc = 4 r_max + radius}. */
int sl_group_stats(int n, const double* pos3, const double* R, const double* m, double* out5);
/* calc_v_collapse (ball_group.hpp:302-313): fall speed onto a group of mass m_total and
 * radius initial_radius, integrated in steps of 0.1. */
int sl_v_collapse(double G, double m_total, double initial_radius, double* v_collapse);
/* getVelMax (ball_group.hpp:317-350): the largest speed among the balls moving toward the
 * centre of mass (angle between velocity and outward radius above `cone`) within `soc` of
 * it; with soc <= 0 the largest speed of all.  `ignored` (may be NULL): balls left out. */
int sl_vel_max(int n, const double* pos3, const double* vel3, const double* m, double soc, double cone,
               double* v_max, int* ignored);
/* calibrate_dt's run length (ball_group.hpp:225-237): steps = sim_time / dt at step 0 (or
 * when dt_old < 0), else (int)(dt_old / dt) * (steps_old - step) + step; skip = the steps per
 * time_resolution (1 s when dt exceeds it). */
int sl_run_length(double dt, double dt_old, int step, int steps_old, double sim_time, double time_resolution,
                  int* steps, int* skip);
/* The run-control state calibrate_dt keeps between calls. */
typedef struct {
    double kin, kout, dt;            /* springs and time step (-1 until calibrated) */
    int steps, skip;                 /* run length and write interval */
    double v_collapse;               /* calc_v_collapse of the group at initialisation */
    double v_max, v_max_prev;        /* speed of the last velocity-rule call; the smallest so far */
    double soc;                      /* sphere of consideration, from the last updateDTK */
} sl_run_state;
/* The state after a group is built: calc_helpfuls and calc_v_collapse of the current
 * positions, nothing calibrated yet (the continue constructor, ball_group.hpp:92-98). */
int sl_run_init(sl_run_state* rs, int n, const double* pos3, const double* R, const double* m, double G);
/* calibrate_dt(step, custom_speed) (ball_group.hpp:197-238): updateDTK at custom_speed when
 * it is positive; otherwise at max(getVelMax, v_collapse) when that is below every speed
 * used before; then steps and skip.  `base` supplies Ha and h_min; kout = cor kin.
 * SL_E_STATE if no calibration has happened yet. */
int sl_calibrate_dt(sl_run_state* rs, const sl_params* base, int n, const double* pos3, const double* vel3,
                    const double* R, const double* m, int step, double custom_speed, double cor);
/* to_origin (ball_group.hpp:542-547): subtracts the centre of mass (getCOM, :519-530) from
 * every position, in place.  SL_E_ARG when the total mass is not positive. */
int sl_to_origin(int n, double* pos3, const double* m);

/* ---- initial conditions --------------------------------------------------- */
/* Seeded ballistic particle-cluster aggregate: each new sphere is placed touching
 * a uniformly chosen earlier one in a uniformly random direction, retried until it
 * overlaps no other sphere.  Radii lognormal (median r_med, log-sd sigma), density
 * rho, moi = 0.4 m R^2.  Positions are centred on the centre of mass; velocities
 * are isotropic Gaussian (sd v_sigma per component) with the mass-weighted mean
 * removed; spins zero.  Deterministic for a seed on every platform. */
int sl_gen_bpca(int n, unsigned long long seed, double r_med, double sigma, double rho,
                double v_sigma, double* pos3, double* vel3, double* w3, double* R, double* m,
                double* moi);

/* ---- collisions ------------------------------------------------------------- */
/* The reference's two-cluster set-up of two loaded groups (sim_init_two_cluster,
 * ball_group.hpp:1368-1442; the collision constructor, :100-112): the projectile (n_p
 * balls) is moved by (its radius + the target's radius + 2 r_max of the target) along
 * (cos a, sin a, 0) for the impact angle a (dust_const.hpp:36 impactParameter), given
 * velocity -v_impact along x (dust_const.hpp:34 v_custom) and the target the balancing
 * +(M_p/M_t) v_impact; the output arrays hold the target, then the projectile (n_t + n_p
 * entries).  Radii are the largest centre distance from each group's centre of mass.
 * v2 (may be NULL) receives {projectile speed, target speed} along x.  Centring and
 * calibration follow separately, as in the constructor: sl_run_init, sl_calibrate_dt,
 * then sl_to_origin. */
int sl_two_cluster(int n_t, const double* t_pos3, const double* t_vel3, const double* t_w3, const double* t_R,
                   const double* t_m, const double* t_moi, int n_p, const double* p_pos3, const double* p_vel3,
                   const double* p_w3, const double* p_R, const double* p_m, const double* p_moi, double v_impact,
                   double impact_angle, double* pos3, double* vel3, double* w3, double* R, double* m, double* moi,
                   double* v2);
/* Two-aggregate collision (the two-cluster set-up of ball_group.hpp:1368-1442, with
 * the impact parameter as a distance, no reference counterpart for the rest): a
 * target of n_t spheres and a projectile of n_p spheres, each a seeded aggregate as
 * sl_gen_bpca builds it (projectile seed derived from seed) with no internal motion,
 * each in a uniformly random orientation.  The projectile's centre of mass sits at
 * (-(R_t + R_p + gap * r_max), b, 0) with b = b_frac * (R_t + R_p), R_t and R_p the
 * bodies' outer radii; it moves along +x at `speed` relative to the target and the
 * total momentum is zero.  Arrays hold n_t + n_p entries, target first.  geom (may be
 * NULL) receives {R_t, R_p, b, M_t, M_p, r_max}. */
int sl_gen_collision(int n_t, int n_p, unsigned long long seed, double speed, double b_frac, double gap,
                     double r_med, double sigma, double rho, double* pos3, double* vel3, double* w3,
                     double* R, double* m, double* moi, double* geom);

/* Outcome of a finished collision from its contact graph: spheres closer than
 * R_i + R_j + link are bound.  With C1, C2 the two heaviest components:
 *   stick    C1 holds at least half of each body's mass and 90% of the total;
 *   bounce   otherwise, if C1 and C2 hold at least half of one body each and all
 *            other components together less than 10% of the mass;
 *   fragment otherwise. */
#define SL_OUTCOME_STICK 0
#define SL_OUTCOME_BOUNCE 1
#define SL_OUTCOME_FRAGMENT 2
typedef struct {
    int label;                /* SL_OUTCOME_* */
    int n_components;
    double largest_mass_frac; /* mass of C1 / total mass */
    double second_mass_frac;  /* mass of C2 / total mass */
    double mass_transfer;     /* projectile mass in the component holding most of the
                                 target's mass, / projectile mass */
    double v_largest[3];      /* centre-of-mass velocity of C1 */
    double v_second[3];       /* centre-of-mass velocity of C2 (0 if none) */
} sl_outcome;
int sl_collision_outcome(int n_t, int n_p, const double* pos3, const double* vel3, const double* R,
                         const double* m, double link, sl_outcome* out);

/* Moves the projectile (balls n_t..) along +x so that its first contact with the
 * target comes after `gap` (cm) of travel.  Returns 1 and the shift applied (cm) when
 * the bodies meet on that line, 0 with no shift when they miss, < 0 on bad input. */
int sl_collision_approach(int n_t, int n_p, double* pos3, const double* R, double gap, double* shift);

/* Descriptors of a collision set-up (inputs of the surrogate model), SL_N_DESCRIPTORS
 * values:
 *   0-3   target {ln N, radius of gyration / mean radius, mass-radius fractal dimension
 *         ln N / ln(Rg/r), mean coordination (neighbours closer than R_i + R_j + link)}
 *   4-7   the same for the projectile
 *   8-11  log10 speed, b_frac, projectile mass fraction, log10 kinetic energy in the
 *         centre-of-mass frame
 *   12    1 if the bodies meet moving along +x, else 0
 *   13-14 mass share of the projectile in the target's shadow along x, and of the
 *         target in the projectile's
 *   15    ln(1 + target balls crossed per projectile mass unit along x)
 *   16-20 at the first contact (0 on a miss): cosine of the contact normal with the
 *         approach line, distance of the contact point from the target's centre of
 *         mass / R_t and from the projectile's / R_p, and the neighbour counts of the
 *         two touching balls within their own bodies */
#define SL_N_DESCRIPTORS 21
int sl_collision_descriptors(int n_t, int n_p, const double* pos3, const double* R, const double* m,
                             double speed, double b_frac, double link, double* desc);
/* The same for `count` collisions at once on up to `nthreads` host threads (0: one per
 * core, at most 16).  Collision k's balls are first[k] .. first[k] + n_t[k] + n_p[k] - 1 of the
 * concatenated pos3 (xyz per ball), R and m; its descriptors go to
 * desc[k * SL_N_DESCRIPTORS ...].  Returns the first error of any collision, else SL_OK. */
int sl_collision_descriptors_batch(int count, const int* n_t, const int* n_p, const long long* first,
                                   const double* pos3, const double* R, const double* m, const double* speed,
                                   const double* b_frac, double link, double* desc, int nthreads);

/* ---- surrogate model (no reference counterpart) ---------------------------
 * A 4-layer perceptron n_in -> width -> width -> width -> n_out with ReLU between
 * layers, trained and evaluated on the GPU; its matrix products use bfloat16
 * operands on the matrix cores with binary32 accumulation, binary32 master weights
 * and Adam.  Outputs 0..2 are the logits of {stick, bounce, fragment}; outputs 3..
 * are regression targets.  Loss per sample: softmax cross-entropy + mean squared
 * error of the regression outputs; a batch's loss is the mean over its samples.
 * Parameters are exchanged flat: for each layer W[in][out] (row-major) then b[out].
 * Every reduction has a fixed order, so training is reproducible bit for bit. */
typedef struct sl_mlp sl_mlp;
int sl_mlp_create(sl_mlp** out, int n_in, int width, int n_out, int max_batch, unsigned long long seed);
int sl_mlp_n_params(const sl_mlp* m);
int sl_mlp_get_params(sl_mlp* m, float* params);
int sl_mlp_set_params(sl_mlp* m, const float* params);   /* also resets the Adam state */
/* y: batch x n_out outputs; prob: batch x 3 class probabilities; ent: predictive
 * entropy (nats) of each sample's class distribution.  Any of them may be NULL. */
int sl_mlp_forward(sl_mlp* m, int batch, const float* x, float* y, float* prob, float* ent);
/* Mean loss of a labelled batch and (if grads is not NULL) its gradient with respect
 * to every parameter, without updating them. */
int sl_mlp_gradients(sl_mlp* m, int batch, const float* x, const int* cls, const float* reg, float* loss,
                     float* grads);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on a labelled batch; loss before the step. */
int sl_mlp_train_step(sl_mlp* m, int batch, const float* x, const int* cls, const float* reg, float lr,
                      float* loss);
const char* sl_mlp_last_error(const sl_mlp* m);
void sl_mlp_destroy(sl_mlp* m);

/* ---- allocation-free stepping ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory
 * while they step: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The same ordered step as sl_step (the same state and energies, bit
 * for bit) on a state the caller keeps in device memory.  The step's scratch (energy
 * partials, touching-pair counters, the non-finite flag) lives in a fixed buffer of the
 * library, one per device, held until the process ends; no call allocates or frees.
 * Every call on a device shares that device's buffer: calls that overlap on one device
 * race.
 *
 * dev_state  SL_NFIELD * stride doubles on the current device, field f of ball i at
 *            dev_state[f * stride + i] (stride >= n), fields in the order below.  The
 *            caller sets position, velocity, spin, acc and aacc (as sl_init_forces or a
 *            step leaves them), R, m and moi; the half-step and previous-position fields
 *            are overwritten by the step.  Out: the state after the steps.
 * first_step the step count before the call: step first_step + k (k = 1..n_steps) is a
 *            write step when it is a multiple of write_every (> 0; <= 0: none).
 * dev_energy device array of SL_STATIC_E_STRIDE doubles per write step, or NULL: PE, KE,
 *            p.x, p.y, p.z, L.x, L.y, L.z, touching pairs of that step, and 1.0 if a
 *            non-finite state appeared since the start of the call (else 0.0).
 * stream     a hipStream_t, or NULL for the null stream.  The call only enqueues work.
 * p          precision must be SL_PREC_FP64; p->order is not read.
 * Returns the number of write steps, SL_E_CAPACITY when n > SL_STATIC_MAX_BALLS (the
 * size the buffer holds), SL_E_ARG for other invalid arguments, SL_E_HIP on a runtime
 * error. */
#define SL_STATIC_MAX_BALLS 65536
#define SL_STATIC_E_STRIDE 10
#define SL_NFIELD 27
#define SL_F_POS 0       /* x, y, z in fields 0, 1, 2; likewise for the vectors below */
#define SL_F_VEL 3
#define SL_F_W 6
#define SL_F_ACC 9
#define SL_F_AACC 12
#define SL_F_VELH 15     /* half-step velocity (written by the step) */
#define SL_F_WH 18       /* half-step spin (written by the step) */
#define SL_F_POS_PREV 21 /* positions of the previous pair pass (written by the step) */
#define SL_F_R 24
#define SL_F_M 25
#define SL_F_MOI 26
int sl_step_ordered_static(const sl_params* p, int n, int stride, double* dev_state, long long first_step,
                           int n_steps, int write_every, double* dev_energy, void* stream);

/* ---- arithmetic probe ------------------------------------------------------ */
/* q = a/b, s = sqrt(a), lg = log(a) element-wise on the device, compiled exactly as
 * the ordered kernels are.  Lets tests check that binary64 division and square
 * root are correctly rounded on the device, which bitwise parity relies on. */
int sl_probe_divsqrt(int n, const double* a, const double* b, double* q, double* s, double* lg);

/* q = a/b element-wise on the device as the ordered kernels divide by a distance or a
 * mass: from the correctly rounded reciprocal y = 1/b, q = a y corrected twice by the
 * FMA-exact residual a - b q.  Must equal the correctly rounded quotient bit for bit. */
int sl_probe_div_rcp(int n, const double* a, const double* b, double* q);

/* Lane layout of the wavefront register rotation the fast pair kernel relies on:
 * out[l] (binary32) and out[64+l] (binary64) must be (l+1)%64, out[128+l] (l+31)%64. */
int sl_probe_rotation(int* out192);

/* A plain device-to-device copy of `bytes` (16 bytes per thread), launched `reps`
 * times: the reference kernel for scaling HBM traffic counters to bytes. */
int sl_probe_copy(unsigned long long bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
