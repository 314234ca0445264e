// spacelab_collider: a ColliderSingleCore-shaped driver on the GPU.
//
// Builds the Ball_group (continued from a run's CSV files, as
// Ball_group(path, filename, customVel) does, ball_group.hpp:92-98; two loaded groups set
// up for a collision, as Ball_group(path, projectile, target, customVel) does, :100-112
// and :1368-1442; or generated),
// calibrates kin, kout, dt, steps and skip as calibrate_dt does (ball_group.hpp:197-238:
// at --v-custom when it is positive, else by the velocity rule; a continued run defaults
// to the velocity rule like the reference's continue constructor, a generated group and a
// collision to dust_const.hpp's v_custom, ColliderSingleCore.cpp:32-34), centres it on its centre of mass, runs init_conditions,
// sim_init_write and sim_looper (ColliderSingleCore.cpp:291-384) and writes
// <out>constants.csv, <out>simData.csv and <out>energy.csv.  Explicit --dt/--kin/--kout
// and --steps/--skip override the calibration; --dynamic-time recalibrates at write
// steps (the reference's dynamicTime).
//
//   spacelab_collider (--continue PATH PREFIX | --two-cluster PATH PROJECTILE TARGET |
//       --generate N SEED) [--out PREFIX] [--v-impact V] [--impact-angle A] [--v-custom V] [--cor C] [--dt DT --kin KIN --kout KOUT] [--steps N] [--skip K]
//       [--order ordered|fast] [--precision fp64|mixed] [--preset reference|config]
//       [--no-center] [--dynamic-time] [--timing] [--quiet]
//
// --v-impact (default dust_const.hpp:34 v_custom) and --impact-angle (radians, default
// dust_const.hpp:36 impactParameter = 0) set the two-cluster collision.
// --timing prints the wall time of sim_looper and the final flush (write steps, state
// downloads and CSV output included) on stdout.
#include "collider.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

namespace {

[[noreturn]] void usage(const char* msg)
{
    std::fprintf(stderr, "spacelab_collider: %s\n"
                         "usage: spacelab_collider (--continue PATH PREFIX | --two-cluster PATH PROJECTILE TARGET |\n"
                         "  --generate N SEED) [--out PREFIX] [--v-impact V] [--impact-angle A] [--v-custom V] [--cor C] [--dt DT --kin KIN --kout KOUT] [--steps N] [--skip K]\n"
                         "  [--order ordered|fast] [--precision fp64|mixed] [--preset reference|config]\n"
                         "  [--no-center] [--dynamic-time] [--timing] [--quiet]\n",
                 msg);
    std::exit(2);
}

}  // namespace

int main(int argc, char** argv)
{
    using namespace spacelab;
    std::string path, prefix, projectile, out = "Unnamed_";   // output_prefix default, dust_const.hpp:67
    int gen_n = 0;
    unsigned long long gen_seed = 0;
    constexpr double kVCustom = 0.36301555459799423;   // dust_const.hpp:34
    double v_custom = -1;                         // unset: chosen by the mode below
    double v_impact = kVCustom, impact_angle = 0;   // dust_const.hpp:34,36
    double cor = 0.4;                             // dust_const.hpp:27
    double dt = 0, kin = 0, kout = 0;
    int steps = -1, skip = -1, order = SL_ORDER_ORDERED, precision = SL_PREC_FP64, preset = 0;
    bool center = true, quiet = false, cont = false, two = false, dynamic = false, timing = false;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) usage(("missing value after " + a).c_str());
            return argv[++i];
        };
        if (a == "--continue") {
            path = next();
            prefix = next();
            cont = true;
        } else if (a == "--two-cluster") {
            path = next();
            projectile = next();
            prefix = next();
            two = true;
        } else if (a == "--generate") {
            gen_n = std::atoi(next());
            gen_seed = std::strtoull(next(), nullptr, 0);
        } else if (a == "--out") out = next();
        else if (a == "--v-custom") v_custom = std::atof(next());
        else if (a == "--v-impact") v_impact = std::atof(next());
        else if (a == "--impact-angle") impact_angle = std::atof(next());
        else if (a == "--cor") cor = std::atof(next());
        else if (a == "--dt") dt = std::atof(next());
        else if (a == "--kin") kin = std::atof(next());
        else if (a == "--kout") kout = std::atof(next());
        else if (a == "--steps") steps = std::atoi(next());
        else if (a == "--skip") skip = std::atoi(next());
        else if (a == "--order") {
            const std::string v = next();
            order = v == "fast" ? SL_ORDER_FAST : v == "ordered" ? SL_ORDER_ORDERED : -1;
            if (order < 0) usage("--order is ordered or fast");
        } else if (a == "--precision") {
            const std::string v = next();
            precision = v == "mixed" ? SL_PREC_MIXED : v == "fp64" ? SL_PREC_FP64 : -1;
            if (precision < 0) usage("--precision is fp64 or mixed");
        } else if (a == "--preset") {
            const std::string v = next();
            preset = v == "config" ? 1 : v == "reference" ? 0 : -1;
            if (preset < 0) usage("--preset is reference or config");
        } else if (a == "--no-center") center = false;
        else if (a == "--dynamic-time") dynamic = true;
        else if (a == "--timing") timing = true;
        else if (a == "--quiet") quiet = true;
        else usage(("unknown argument " + a).c_str());
    }
    if ((int)cont + (int)two + (int)(gen_n > 0) != 1) usage("give exactly one of --continue, --two-cluster and --generate");
    try {
        Ball_group O = cont  ? Ball_group::from_csv(path, prefix)
                       : two ? Ball_group::two_cluster(Ball_group::from_csv(path, prefix),
                                                       Ball_group::from_csv(path, projectile), v_impact, impact_angle)
                             : Ball_group::generate(gen_n, gen_seed);
        sl_params p;
        sl_params_preset(&p, preset);
        // the continue constructor calibrates with customVel 0 (the velocity rule), the
        // generator and the collision at v_custom; all before centring (ball_group.hpp:67-112)
        if (v_custom < 0) v_custom = cont ? 0.0 : kVCustom;
        RunControl rc(O, p, cor);
        if (!(dt > 0 && kin > 0 && kout > 0 && steps >= 0 && skip > 0)) rc.calibrate(O, 0, v_custom);
        if (center) O.to_origin();
        if (dt > 0) rc.rs.dt = dt;
        if (kin > 0) rc.rs.kin = kin;
        if (kout > 0) rc.rs.kout = kout;
        if (steps >= 0) rc.rs.steps = steps;
        if (skip > 0) rc.rs.skip = skip;
        p.kin = rc.rs.kin;
        p.kout = rc.rs.kout;
        p.dt = rc.rs.dt;
        p.order = order;
        p.precision = precision;
        if (!quiet)
            std::cerr << "==================\ndt: " << p.dt << "\nk: " << p.kin << "\nSkip: " << rc.rs.skip
                      << "\nSteps: " << rc.rs.steps << "\n==================\n";
        Collider c(O, p);
        c.progress = !quiet;
        c.init_conditions();
        c.sim_init_write(out);
        const time_t t0 = std::time(nullptr);
        const auto c0 = std::chrono::steady_clock::now();
        c.sim_looper(rc.rs.steps, rc.rs.skip, dynamic ? &rc : nullptr);
        c.flush();
        const double loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
        if (timing) std::printf("loop_seconds %.6f steps %d balls %d\n", loop_s, rc.rs.steps, O.num_particles);
        if (!quiet)
            std::cerr << "Simulation complete!\n"
                      << O.num_particles << " Particles and " << rc.rs.steps << " Steps.\n"
                      << "Simulated time: " << rc.rs.steps * p.dt << " seconds\n"
                      << "Computation time: " << std::time(nullptr) - t0 << " seconds\n"
                      << "Output: " << c.output_prefix << "{constants,simData,energy}.csv\n";
    } catch (const Error& e) {
        std::fprintf(stderr, "spacelab_collider: %s\n", e.what());
        return 1;
    }
    return 0;
}
