// C++ host driver over the C ABI: a Ball_group with the reference's arrays and
// loaders, and a Collider that runs sim_one_step / sim_looper on the GPU and writes
// the run's constants, simData and energy files in the reference's layout.
//
// Reference counterparts: Ball_group (ball_group.hpp:21-53 arrays, :1236-1247 and
// :1352-1364 loadSim/sim_continue via getLastLine/parseSimData/loadConsts, :1368-1442
// sim_init_two_cluster, :421-516 sim_init_write, :519-547 getCOM/to_origin, :1117-1124
// getMass), calibrate_dt's first call
// (:197-238: updateDTK at a custom speed, then steps and skip from the simulated time
// and the time resolution) and the driver's sim_one_step / sim_looper
// (ColliderSingleCore.cpp:78-288, :291-384).
//
// The state lives on the GPU between steps; it is downloaded on write steps only
// (the simData row needs it).  Non-write steps between two write steps go to the
// library in one call, which gives the same results as one call per step.
#pragma once

#include "spacelab/sl_api.h"
#include "spacelab/sl_csv.h"

#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <ctime>
#include <deque>
#include <exception>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace spacelab {

struct vec3 {
    double x = 0, y = 0, z = 0;
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    vec3& operator+=(const vec3& o)
    {
        x += o.x;
        y += o.y;
        z += o.z;
        return *this;
    }
};
static_assert(sizeof(vec3) == 3 * sizeof(double), "vec3 must be three packed doubles (the ABI's AoS layout)");

// Constants of the reference's run control (dust_const.hpp:28-29).
constexpr double kSimTimeSeconds = 0.5e-3;
constexpr double kTimeResolution = 1e-5;

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what + " (code " + std::to_string(code) + ")"), code(code) {}
    int code;
};

inline std::string read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(SL_E_ARG, "cannot open " + path);
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

inline bool file_exists(const std::string& path) { return std::ifstream(path).good(); }

class Ball_group {
public:
    int num_particles = 0;
    std::vector<vec3> pos, vel, w, acc, aacc;
    std::vector<double> R, m, moi;
    double PE = 0, KE = 0;
    vec3 mom, ang_mom;

    Ball_group() = default;
    explicit Ball_group(int n) { allocate(n); }

    void allocate(int n)
    {
        num_particles = n;
        pos.assign(n, vec3{});
        vel.assign(n, vec3{});
        w.assign(n, vec3{});
        acc.assign(n, vec3{});
        aacc.assign(n, vec3{});
        R.assign(n, 0.0);
        m.assign(n, 0.0);
        moi.assign(n, 0.0);
    }

    // Continue a run: the last simData row and the constants file of <path><prefix>.
    static Ball_group from_csv(const std::string& path, const std::string& prefix)
    {
        const std::string sim = path + prefix + "simData.csv";
        const long long len = sl_csv_last_line(sim.c_str(), nullptr, 0);
        if (len < 0) throw Error((int)len, "cannot read " + sim);
        std::string line((size_t)len + 1, '\0');
        sl_csv_last_line(sim.c_str(), &line[0], len + 1);
        line.resize((size_t)len);
        Ball_group g(sl_csv_count_balls(line.c_str()));
        int rc = sl_csv_parse_sim_row(line.c_str(), g.num_particles, &g.pos[0].x, &g.w[0].x, &g.vel[0].x);
        if (rc) throw Error(rc, "bad simData row in " + sim);
        // <prefix>constants.csv, or the Generator's <prefix>Constants.csv where only that exists
        std::string cname = path + prefix + "constants.csv";
        if (!file_exists(cname) && file_exists(path + prefix + "Constants.csv")) cname = path + prefix + "Constants.csv";
        const std::string consts = read_file(cname);
        rc = sl_csv_parse_constants(consts.c_str(), g.num_particles, g.R.data(), g.m.data(), g.moi.data());
        if (rc) throw Error(rc, "bad constants file " + cname);
        return g;
    }

    // Seeded ballistic particle-cluster aggregate (sl_gen_bpca).
    static Ball_group generate(int n, unsigned long long seed, double r_med = 1e-5, double sigma = 0.2,
                               double rho = 2.25, double v_sigma = 1e-3)
    {
        Ball_group g(n);
        const int rc = sl_gen_bpca(n, seed, r_med, sigma, rho, v_sigma, &g.pos[0].x, &g.vel[0].x, &g.w[0].x,
                                   g.R.data(), g.m.data(), g.moi.data());
        if (rc) throw Error(rc, "sl_gen_bpca");
        return g;
    }

    double getMass() const
    {
        double t = 0;
        for (int i = 0; i < num_particles; i++) t += m[i];
        return t;
    }

    vec3 getCOM() const
    {
        const double M = getMass();
        vec3 c;
        for (int i = 0; i < num_particles; i++) {
            c.x += pos[i].x * m[i];
            c.y += pos[i].y * m[i];
            c.z += pos[i].z * m[i];
        }
        c.x /= M;
        c.y /= M;
        c.z /= M;
        return c;
    }

    void to_origin()
    {
        const int rc = sl_to_origin(num_particles, &pos[0].x, m.data());
        if (rc) throw Error(rc, "to_origin: the group has no mass");
    }

    // The two-cluster set-up (sim_init_two_cluster, ball_group.hpp:1368-1442, over
    // sl_two_cluster): the projectile offset beside the target at the impact angle,
    // moving at -v_impact along x, the target at the balancing speed; target first.
    static Ball_group two_cluster(const Ball_group& target, const Ball_group& projectile, double v_impact,
                                  double impact_angle, double* v2 = nullptr)
    {
        Ball_group g(target.num_particles + projectile.num_particles);
        const int rc = sl_two_cluster(target.num_particles, &target.pos[0].x, &target.vel[0].x, &target.w[0].x,
                                      target.R.data(), target.m.data(), target.moi.data(), projectile.num_particles,
                                      &projectile.pos[0].x, &projectile.vel[0].x, &projectile.w[0].x,
                                      projectile.R.data(), projectile.m.data(), projectile.moi.data(), v_impact,
                                      impact_angle, &g.pos[0].x, &g.vel[0].x, &g.w[0].x, g.R.data(), g.m.data(),
                                      g.moi.data(), v2);
        if (rc) throw Error(rc, "sl_two_cluster");
        return g;
    }

    void zeroVel()
    {
        for (auto& v : vel) v = vec3{};
    }
    void zeroAngVel()
    {
        for (auto& v : w) v = vec3{};
    }
};

// calibrate_dt's run control (sl_run_init / sl_calibrate_dt): built from a group as the
// reference's constructors do (calc_helpfuls, calc_v_collapse), then calibrate(step,
// custom_speed): updateDTK at custom_speed when it is positive, else at the largest
// speed of the balls falling toward the centre of mass (at least the collapse speed)
// when that is below every speed used before; steps and skip follow.
struct RunControl {
    sl_run_state rs{};
    sl_params base{};
    double cor = 0.4;   // dust_const.hpp:27

    RunControl(const Ball_group& g, const sl_params& base_params, double cor_ = 0.4) : base(base_params), cor(cor_)
    {
        const int rc = sl_run_init(&rs, g.num_particles, &g.pos[0].x, g.R.data(), g.m.data(), base.G);
        if (rc) throw Error(rc, "sl_run_init");
    }

    void calibrate(const Ball_group& g, int step, double custom_speed)
    {
        const int rc = sl_calibrate_dt(&rs, &base, g.num_particles, &g.pos[0].x, &g.vel[0].x, g.R.data(), g.m.data(),
                                       step, custom_speed, cor);
        if (rc) throw Error(rc, "sl_calibrate_dt");
    }
};

// simData rows formatted on a host thread of their own, in the order they were queued,
// while the GPU runs the next steps: a write step hands over a copy of the state and
// returns.  take() waits for the queue to drain and returns the text.
class RowFormatter {
public:
    RowFormatter() : th_([this] { run(); }) {}
    ~RowFormatter()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    RowFormatter(const RowFormatter&) = delete;
    RowFormatter& operator=(const RowFormatter&) = delete;

    void push(int n, const double* pos3, const double* w3, const double* vel3)
    {
        Job j{n, std::vector<double>(pos3, pos3 + 3 * (size_t)n), std::vector<double>(w3, w3 + 3 * (size_t)n),
              std::vector<double>(vel3, vel3 + 3 * (size_t)n)};
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(std::move(j));
        }
        cv_.notify_all();
    }

    // the rows of every job queued so far, each preceded by '\n'
    std::string take()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return q_.empty() && !busy_; });
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
        std::string out;
        out.swap(out_);
        return out;
    }

private:
    struct Job {
        int n;
        std::vector<double> pos, w, vel;
    };
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    std::string out_;
    bool busy_ = false, stop_ = false;
    std::exception_ptr err_;
    std::thread th_;

    void run()
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;   // stop_ with nothing left
            Job j = std::move(q_.front());
            q_.pop_front();
            busy_ = true;
            lk.unlock();
            std::string row;
            std::exception_ptr e;
            try {
                // a number takes at most 13 characters at 6 digits ("-1.23457e-308")
                row.assign((size_t)j.n * (11 * 14) + 16, '\0');
                long long len = sl_csv_sim_row(j.n, j.pos.data(), j.w.data(), j.vel.data(), &row[0], (long long)row.size());
                if (len >= (long long)row.size()) {
                    row.assign((size_t)len + 1, '\0');
                    len = sl_csv_sim_row(j.n, j.pos.data(), j.w.data(), j.vel.data(), &row[0], (long long)row.size());
                }
                if (len < 0) throw Error((int)len, "CSV formatting");
                row.resize((size_t)len);
            } catch (...) {
                e = std::current_exception();
            }
            lk.lock();
            if (e) {
                if (!err_) err_ = e;
            } else {
                out_ += '\n';
                out_ += row;
            }
            busy_ = false;
            cv_.notify_all();
        }
    }
};

class Collider {
public:
    std::string output_prefix;
    double simTimeElapsed = 0;
    bool progress = true;

    Collider(Ball_group& g, const sl_params& p) : O(g), params(p)
    {
        int rc = sl_create(&ctx, g.num_particles, &p);
        if (rc) throw Error(rc, rc == SL_E_NOGPU ? "no usable gfx950 GPU" : "sl_create");
        rc = sl_upload(ctx, g.num_particles, &g.pos[0].x, &g.vel[0].x, &g.w[0].x, &g.acc[0].x, &g.aacc[0].x,
                       g.R.data(), g.m.data(), g.moi.data(), nullptr);
        check(rc, "sl_upload");
        dt = p.dt;
    }

    // New springs and time step for the following steps (the state stays on the GPU).
    void set_run(double kin, double kout, double dt_new)
    {
        params.kin = kin;
        params.kout = kout;
        params.dt = dt_new;
        check(sl_set_params(ctx, &params), "sl_set_params");
        dt = dt_new;
    }
    ~Collider()
    {
        if (ctx) sl_destroy(ctx);
    }
    Collider(const Collider&) = delete;
    Collider& operator=(const Collider&) = delete;

    // init_conditions (ball_group.hpp:777-947): forces of the current state and the
    // energies PE, KE, momentum and angular momentum, into O.
    void init_conditions()
    {
        sl_energy e;
        check(sl_init_forces(ctx, &e), "sl_init_forces");
        add_energy(e);
        download();
    }

    // sim_init_write (ball_group.hpp:421-516): picks a free file name (a counter
    // prefix while <prefix>simData.csv exists), writes the constants, both headers,
    // the current energy row at 6 digits and the current state row, then resets the
    // energy sums.
    void sim_init_write(std::string filename, int counter = 0)
    {
        bool taken = file_exists(filename + "simData.csv");
        while (taken) {
            counter++;
            taken = file_exists(std::to_string(counter) + '_' + filename + "simData.csv");
        }
        if (counter > 0) filename.insert(0, std::to_string(counter) + '_');
        output_prefix = filename;
        const int n = O.num_particles;
        std::string e = SL_CSV_ENERGY_HEADER;
        e += '\n';
        e += energy_row(SL_CSV_STATE_DIGITS);
        append(filename + "energy.csv", e);
        std::string s = sim_header(n);
        s += '\n';
        s += sim_row();
        append(filename + "simData.csv", s);
        append(filename + "constants.csv", constants());
        reset_energy();
    }

    // sim_one_step (ColliderSingleCore.cpp:78-288): one velocity-Verlet step; on a
    // write step the state row goes to the simData buffer and the energies are added
    // to O's sums.
    void sim_one_step(bool write_step) { steps_then_write(write_step ? 0 : 1, write_step); }

    // sim_looper (ColliderSingleCore.cpp:291-384): Step = 1 .. steps-1, write steps
    // every `skip`, an energy row at 12 digits after each, the buffers flushed to the
    // files at the first write and at every tenth (or after 30 minutes).  With `dynamic`
    // (the reference's dynamicTime, off in dust_const.hpp:10) each write step ends with
    // calibrate_dt(Step) by the velocity rule; new springs, dt, run length and write
    // interval take effect from the next step.
    void sim_looper(int steps, int skip, RunControl* dynamic = nullptr)
    {
        if (skip < 1) throw Error(SL_E_ARG, "skip must be at least 1");
        const time_t t0 = std::time(nullptr);
        int pending = 0;   // non-write steps not yet run
        if (dynamic) dynamic->rs.steps = steps;
        for (int Step = 1; Step < (dynamic ? dynamic->rs.steps : steps); Step++) {
            if (Step % skip != 0) {
                pending++;
                continue;
            }
            simTimeElapsed += dt * skip;
            if (progress)
                std::fprintf(stderr, "%d\t%2.0f%%\tReal: %.1f s\tSim: %.4g s\n", Step,
                             100.0 * Step / steps, difftime(std::time(nullptr), t0), simTimeElapsed);
            steps_then_write(pending, true);
            pending = 0;
            energyBuffer += '\n';
            energyBuffer += energy_row(SL_CSV_ENERGY_DIGITS);
            reset_energy();
            if (std::time(nullptr) - lastWrite > 1800 || Step / skip % 10 == 0) flush();
            if (dynamic) {
                const double dt_before = dynamic->rs.dt;
                dynamic->calibrate(O, Step, 0.0);
                if (dynamic->rs.dt != dt_before) set_run(dynamic->rs.kin, dynamic->rs.kout, dynamic->rs.dt);
                skip = dynamic->rs.skip;
            }
        }
        if (pending) steps_then_write(pending, false);
    }

    // Appends the buffered rows to the files.  The reference flushes only inside the
    // loop, so rows buffered after its last flush are not written by it.
    void flush()
    {
        ballBuffer += rows.take();
        append(output_prefix + "simData.csv", ballBuffer);
        append(output_prefix + "energy.csv", energyBuffer);
        ballBuffer.clear();
        energyBuffer.clear();
        lastWrite = std::time(nullptr);
    }

    void download()
    {
        check(sl_download(ctx, &O.pos[0].x, &O.vel[0].x, &O.w[0].x, &O.acc[0].x, &O.aacc[0].x), "sl_download");
    }

    sl_ctx* context() { return ctx; }

private:
    Ball_group& O;
    sl_params params;
    sl_ctx* ctx = nullptr;
    double dt = 0;
    std::string ballBuffer, energyBuffer;
    RowFormatter rows;
    time_t lastWrite = 0;

    void check(int rc, const char* what)
    {
        if (rc < 0) throw Error(rc, std::string(what) + ": " + sl_last_error(ctx));
    }

    // `n_plain` non-write steps, then one step that is a write step if `write`
    void steps_then_write(int n_plain, bool write)
    {
        if (n_plain > 0) check(sl_step(ctx, n_plain, 0, nullptr, 0), "sl_step");
        if (!write) return;
        sl_energy e;
        check(sl_step(ctx, 1, 1, &e, 1), "sl_step");
        add_energy(e);
        download();
        // the row is formatted on the formatter's thread while the next steps run
        rows.push(O.num_particles, &O.pos[0].x, &O.w[0].x, &O.vel[0].x);
    }

    void add_energy(const sl_energy& e)
    {
        O.PE += e.PE;
        O.KE += e.KE;
        O.mom += vec3{e.p[0], e.p[1], e.p[2]};
        O.ang_mom += vec3{e.L[0], e.L[1], e.L[2]};
    }

    void reset_energy()
    {
        O.PE = 0;
        O.KE = 0;
        O.mom = vec3{};
        O.ang_mom = vec3{};
    }

    // One formatting pass into a buffer of `guess` bytes (an upper bound for the text;
    // a second pass only if the text turns out longer).
    template <class F>
    static std::string text(F f, size_t guess)
    {
        std::string s(guess + 1, '\0');
        long long len = f(&s[0], (long long)s.size());
        if (len < 0) throw Error((int)len, "CSV formatting");
        if ((size_t)len >= s.size()) {
            s.assign((size_t)len + 1, '\0');
            f(&s[0], len + 1);
        }
        s.resize((size_t)len);
        return s;
    }

    // a number takes at most 13 characters at 6 digits ("-1.23457e-308") and 19 at 12
    std::string sim_header(int n)
    {
        return text([&](char* b, long long c) { return sl_csv_sim_header(n, b, c); }, (size_t)n * 80 + 16);
    }
    std::string sim_row()
    {
        return text([&](char* b, long long c) {
            return sl_csv_sim_row(O.num_particles, &O.pos[0].x, &O.w[0].x, &O.vel[0].x, b, c);
        }, (size_t)O.num_particles * (11 * 14) + 16);
    }
    std::string energy_row(int digits)
    {
        return text([&](char* b, long long c) {
            return sl_csv_energy_row(simTimeElapsed, O.PE, O.KE, &O.mom.x, &O.ang_mom.x, digits, b, c);
        }, 6 * 20 + 16);
    }
    std::string constants()
    {
        return text([&](char* b, long long c) {
            return sl_csv_constants(O.num_particles, O.R.data(), O.m.data(), O.moi.data(), b, c);
        }, (size_t)O.num_particles * (3 * 14) + 16);
    }

    static void append(const std::string& path, const std::string& s)
    {
        std::ofstream f(path, std::ios::binary | std::ios::app);
        if (!f) throw Error(SL_E_ARG, "cannot write " + path);
        f << s;
    }
};

}  // namespace spacelab
