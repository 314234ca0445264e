// Harness that drives the reference's own sim_one_step -- TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile, in this container only, straight from the
// reference's source (REF_DRIVER = ColliderSingleCore/ColliderSingleCore.cpp,
// which includes ball_group.hpp, dust_const.hpp, vec3.hpp, Utils.hpp and
// linalg.hpp).  Nothing of the reference is copied: the file is included where
// it lies.  Output goes to oracle/_ref/ only.
//
// Usage: ref_harness <input.bin> <output.bin>          state and energies
//        ref_harness --csv <input.bin> <output-dir>    the reference's own CSV files
//        ref_harness --calib <input.bin> <output.bin>  the reference's calibrate_dt
//        ref_harness --two-cluster <dir/> <projectile> <target> <custom> <output.bin>
//                                                      the reference's two-cluster constructor
//
// Protocol (little-endian binary files):
//   input : int32 {magic 0x534C4731, n, n_steps, skip}
//           f64   {dt, kin, kout}
//           f64   pos[3n], vel[3n], w[3n], R[n], m[n], moi[n]
//   output: f64   pos, vel, velh, w, wh, acc, aacc   (each 3n)
//           f64   e_init[8]  (PE, KE, p.xyz, L.xyz after the step-0 pass)
//           int32 n_writes, then f64 e[n_writes][8]
//
// Flow, as the reference runs it (ball_group.hpp:92-98 -> init_conditions,
// then sim_looper, ColliderSingleCore.cpp:298-331):
//   1. distances[e] = |x_B - x_A| of the input positions (the reference leaves
//      them uninitialised before init_conditions, ball_group.hpp:741,796).
//   2. step-0 pass: sim_one_step(true) with dt = 0.  With dt = 0 the first and
//      third passes leave positions and velocities unchanged, and the pair pass is
//      the same computation as init_conditions (ball_group.hpp:780-946); this
//      avoids reaching into the class's private section.
//   3. for Step = 1 .. n_steps-1: sim_one_step(Step % skip == 0); energies of
//      write steps are recorded and reset as sim_looper does (:333-344).
//
// Calibration mode: calc_helpfuls and calc_v_collapse of the input group, then three
// calls of the reference's calibrate_dt: (step 0, velocity rule), (step 100, velocity rule
// after the velocities are halved), (step 200, custom speed = the input's dt field).
// After each: f64 {kin, kout, dt, steps, skip, v_collapse, v_max, v_max_prev, soc}.
//
// Two-cluster mode: Ball_group(dir, projectile, target, custom) -- the reference's own
// constructor (ball_group.hpp:100-112): sim_init_two_cluster loads both groups from
// <dir><name>simData.csv and constants.csv, offsets, kicks and merges them (:1368-1442),
// then calc_v_collapse, calibrate_dt(0, custom) and simInit_cond_and_center (to_origin,
// init_conditions).  Output: int32 n; f64 pos[3n], vel[3n], w[3n], R[n], m[n], moi[n];
// f64 {kin, kout, dt, steps, skip, v_collapse, v_max, v_max_prev, soc, m_total,
// initial_radius, r_max}.  init_conditions reads the uninitialised distances array
// (ball_group.hpp:741,796), so its forces and energies are not part of the output.
//
// CSV mode: after the step-0 pass, O.sim_init_write(<output-dir>/) writes
// constants.csv and the headers and first rows of simData.csv and energy.csv
// (ball_group.hpp:421-516), then the reference's own sim_looper
// (ColliderSingleCore.cpp:291-384) runs Step = 1 .. n_steps-1 with its write,
// energy-row and flush logic (precision 12 on its energy stream, :42).  Rows still
// buffered when sim_looper returns are not written, as in the reference.
//
// The reference's translation unit constructs a global Ball_group from two CSV
// files at a fixed relative path (ColliderSingleCore.cpp:33, dust_const.hpp:65-66)
// during static initialisation.  ref_stage below runs first (same TU, defined
// earlier), moves into a fresh temporary directory and writes a two-sphere
// input there so that constructor succeeds; main() then replaces the group.

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

// The reference calls isnan unqualified (ball_group.hpp:1419), which MSVC's
// <cmath> allows; libstdc++ needs the using-declaration.
using std::isnan;

namespace {
std::string g_tmpdir;

void mkdirs(const std::string& p)
{
    std::string cur;
    for (size_t i = 0; i < p.size(); ++i) {
        cur.push_back(p[i]);
        if (p[i] == '/' || i + 1 == p.size()) ::mkdir(cur.c_str(), 0755);
    }
}

struct RefStage {
    RefStage()
    {
        char tmpl[] = "/tmp/sl_ref_XXXXXX";
        char* d = ::mkdtemp(tmpl);
        if (!d || ::chdir(d) != 0) {
            std::fprintf(stderr, "ref_harness: cannot create temp dir\n");
            std::exit(3);
        }
        g_tmpdir = d;
        const std::string dir = "C:/Users/milin/Desktop/VSOUT/";
        const std::string tag = "2_R2e-05_v1e+00_cor0.16_mu0.1_rho2.25_k3e+01_Ha5e-12_dt2e-10_";
        mkdirs(dir);
        FILE* f = std::fopen((dir + tag + "simData.csv").c_str(), "w");
        std::fprintf(f, "x0,y0,z0,wx0,wy0,wz0,wmag0,vx0,vy0,vz0,bound0,x1,y1,z1,wx1,wy1,wz1,wmag1,vx1,vy1,vz1,bound1\n");
        std::fprintf(f, "0,1.101e-05,0,0,0,0,0,0,0,0,0,0,-1.101e-05,0,0,0,0,0,0,0,0,0");
        std::fclose(f);
        f = std::fopen((dir + tag + "constants.csv").c_str(), "w");
        std::fprintf(f, "1e-05,9.42477e-15,3.76991e-25\n1e-05,9.42477e-15,3.76991e-25\n");
        std::fclose(f);
    }
} ref_stage;
}  // namespace

#define main spacelab_reference_main
#include REF_DRIVER
#undef main

namespace {
bool read_all(FILE* f, void* p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

void cleanup()
{
    if (!g_tmpdir.empty()) {
        std::string cmd = "rm -rf '" + g_tmpdir + "'";
        if (std::system(cmd.c_str()) != 0) std::fprintf(stderr, "ref_harness: cleanup failed\n");
    }
}

void put_energy(std::vector<double>& out)
{
    out.push_back(O.PE);
    out.push_back(O.KE);
    out.push_back(O.mom.x);
    out.push_back(O.mom.y);
    out.push_back(O.mom.z);
    out.push_back(O.ang_mom.x);
    out.push_back(O.ang_mom.y);
    out.push_back(O.ang_mom.z);
    O.KE = 0;
    O.PE = 0;
    O.mom = {0, 0, 0};
    O.ang_mom = {0, 0, 0};
}
}  // namespace

int two_cluster(char** argv)
{
    const Ball_group g(argv[2], argv[3], argv[4], std::atof(argv[5]));
    FILE* fo = std::fopen(argv[6], "wb");
    if (!fo) return 2;
    const int32_t n = g.num_particles;
    std::fwrite(&n, 4, 1, fo);
    for (const vec3* a : {g.pos, g.vel, g.w})
        for (int i = 0; i < n; i++) {
            const double v[3] = {a[i].x, a[i].y, a[i].z};
            std::fwrite(v, 8, 3, fo);
        }
    for (const double* a : {g.R, g.m, g.moi}) std::fwrite(a, 8, n, fo);
    const double tail[12] = {kin, kout, dt, (double)steps, (double)skip, g.v_collapse, g.v_max,
                             g.v_max_prev, g.soc, g.m_total, g.initial_radius, g.r_max};
    std::fwrite(tail, 8, 12, fo);
    std::fclose(fo);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 7 && std::strcmp(argv[1], "--two-cluster") == 0) {
        const int rc = two_cluster(argv);
        cleanup();
        return rc;
    }
    const bool csv = argc == 4 && std::strcmp(argv[1], "--csv") == 0;
    const bool calib = argc == 4 && std::strcmp(argv[1], "--calib") == 0;
    if (argc != 3 && !csv && !calib) {
        std::fprintf(stderr, "usage: ref_harness <input.bin> <output.bin> | --csv <input.bin> <output-dir>\n");
        cleanup();
        return 2;
    }
    if (csv || calib) argv++;
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) { cleanup(); return 2; }
    int32_t hdr[4];
    double prm[3];
    if (!read_all(fi, hdr, sizeof hdr) || hdr[0] != 0x534C4731 || !read_all(fi, prm, sizeof prm)) {
        std::fprintf(stderr, "ref_harness: bad header\n");
        cleanup();
        return 2;
    }
    const int n = hdr[1], n_steps = hdr[2], skip_in = hdr[3];
    std::vector<double> pos(3 * n), vel(3 * n), ww(3 * n), R(n), m(n), moi(n);
    bool ok = read_all(fi, pos.data(), 8 * pos.size()) && read_all(fi, vel.data(), 8 * vel.size()) &&
              read_all(fi, ww.data(), 8 * ww.size()) && read_all(fi, R.data(), 8 * R.size()) &&
              read_all(fi, m.data(), 8 * m.size()) && read_all(fi, moi.data(), 8 * moi.size());
    std::fclose(fi);
    if (!ok) { std::fprintf(stderr, "ref_harness: short input\n"); cleanup(); return 2; }

    Ball_group g(n);
    for (int i = 0; i < n; i++) {
        g.pos[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        g.vel[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        g.w[i] = {ww[3 * i], ww[3 * i + 1], ww[3 * i + 2]};
        g.R[i] = R[i];
        g.m[i] = m[i];
        g.moi[i] = moi[i];
    }
    for (int A = 1; A < n; A++)
        for (int B = 0; B < A; B++)
            g.distances[static_cast<unsigned>(A * (A - 1) * .5) + B] = (g.pos[B] - g.pos[A]).norm();
    O = g;
    O.KE = 0;
    O.PE = 0;
    O.mom = {0, 0, 0};
    O.ang_mom = {0, 0, 0};

    if (calib) {
        std::vector<double> rec;
        auto put = [&]() {
            const double v[9] = {kin, kout, dt, (double)steps, (double)skip, O.v_collapse, O.v_max, O.v_max_prev, O.soc};
            rec.insert(rec.end(), v, v + 9);
        };
        dt = -1;
        kin = -1;
        kout = -1;
        steps = 0;
        skip = 0;
        O.calc_helpfuls();
        O.calc_v_collapse();
        O.calibrate_dt(0, 0.0);
        put();
        for (int i = 0; i < n; i++) O.vel[i] = O.vel[i] * 0.5;
        O.calibrate_dt(100, 0.0);
        put();
        O.calibrate_dt(200, prm[0]);
        put();
        FILE* fo = std::fopen(argv[2], "wb");
        if (!fo) { cleanup(); return 2; }
        std::fwrite(rec.data(), 8, rec.size(), fo);
        std::fclose(fo);
        cleanup();
        return 0;
    }
    kin = prm[1];
    kout = prm[2];
    skip = skip_in;
    std::vector<double> e_init, e_w;
    dt = 0;
    sim_one_step(true);
    if (csv) {
        // the step-0 pass stands for init_conditions: its row is not the reference's
        ballBuffer.str("");
        std::string dir = argv[2];
        if (dir.empty() || dir.back() != '/') dir += '/';
        O.sim_init_write(dir);
        energyBuffer.precision(12);   // as the reference's main does (ColliderSingleCore.cpp:42)
        steps = n_steps;
        skip = skip_in;
        dt = prm[0];
        sim_looper();
        cleanup();
        return 0;
    }
    put_energy(e_init);
    ballBuffer.str("");
    dt = prm[0];
    int n_writes = 0;
    for (int Step = 1; Step < n_steps; Step++) {
        const bool w = skip > 0 && (Step % skip == 0);
        sim_one_step(w);
        if (w) {
            put_energy(e_w);
            n_writes++;
            ballBuffer.str("");
        }
    }

    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) { cleanup(); return 2; }
    auto put3 = [&](const vec3* a) {
        for (int i = 0; i < n; i++) {
            const double v[3] = {a[i].x, a[i].y, a[i].z};
            std::fwrite(v, 8, 3, fo);
        }
    };
    put3(O.pos);
    put3(O.vel);
    put3(O.velh);
    put3(O.w);
    put3(O.wh);
    put3(O.acc);
    put3(O.aacc);
    std::fwrite(e_init.data(), 8, 8, fo);
    const int32_t nw = n_writes;
    std::fwrite(&nw, 4, 1, fo);
    if (n_writes) std::fwrite(e_w.data(), 8, e_w.size(), fo);
    std::fclose(fo);
    cleanup();
    return 0;
}
