// Host-side collision set-up and outcome analysis: two seeded aggregates aimed at
// each other (the two-cluster set-up of ball_group.hpp:1368-1442, with the impact
// parameter as a distance), the contact-graph outcome of a finished collision
// (stick / bounce / fragment, mass transfer, fragment velocities) and the
// descriptors a surrogate model reads.  Plain C++, no device code.
#include "spacelab/sl_api.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

// SplitMix64 (Steele, Lea, Flood 2014), as the aggregate generator uses.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};

// Uniform random rotation matrix from three uniforms (Shoemake, Graphics Gems III,
// "Uniform random rotations", 1992).
void random_rotation(Rng& rng, double Q[9])
{
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double w = a * std::sin(2 * kPi * u2), x = a * std::cos(2 * kPi * u2);
    const double y = b * std::sin(2 * kPi * u3), z = b * std::cos(2 * kPi * u3);
    Q[0] = 1 - 2 * (y * y + z * z); Q[1] = 2 * (x * y - z * w);     Q[2] = 2 * (x * z + y * w);
    Q[3] = 2 * (x * y + z * w);     Q[4] = 1 - 2 * (x * x + z * z); Q[5] = 2 * (y * z - x * w);
    Q[6] = 2 * (x * z - y * w);     Q[7] = 2 * (y * z + x * w);     Q[8] = 1 - 2 * (x * x + y * y);
}

struct Body {
    int n;
    double M = 0, com[3] = {0, 0, 0}, radius = 0;
};

// Mass, centre of mass and outer radius max(|x - com| + R) of balls [b0, b0 + n).
Body body_of(const double* pos3, const double* R, const double* m, int b0, int n)
{
    Body b;
    b.n = n;
    for (int i = b0; i < b0 + n; i++) {
        b.M += m[i];
        for (int k = 0; k < 3; k++) b.com[k] += m[i] * pos3[3 * i + k];
    }
    for (int k = 0; k < 3; k++) b.com[k] /= b.M;
    for (int i = b0; i < b0 + n; i++) {
        double d2 = 0;
        for (int k = 0; k < 3; k++) d2 += (pos3[3 * i + k] - b.com[k]) * (pos3[3 * i + k] - b.com[k]);
        b.radius = std::max(b.radius, std::sqrt(d2) + R[i]);
    }
    return b;
}

// Calls f(i, j), i < j, for every pair closer than R_i + R_j + link, found through a
// uniform grid whose cell (2 max R + link) holds every such partner in the 27 cells
// around a ball.
template <class F>
void near_pairs(int n, const double* pos3, const double* R, double link, F&& f)
{
    double rmax = 0;
    for (int i = 0; i < n; i++) rmax = std::max(rmax, R[i]);
    const double cell = 2 * rmax + link;
    auto ckey = [&](long long x, long long y, long long z) {
        return ((x & 0x1FFFFF) << 42) | ((y & 0x1FFFFF) << 21) | (z & 0x1FFFFF);
    };
    std::unordered_map<long long, std::vector<int>> grid;
    grid.reserve((size_t)n * 2);
    std::vector<long long> cx(n), cy(n), cz(n);
    for (int i = 0; i < n; i++) {
        cx[i] = (long long)std::floor(pos3[3 * i] / cell);
        cy[i] = (long long)std::floor(pos3[3 * i + 1] / cell);
        cz[i] = (long long)std::floor(pos3[3 * i + 2] / cell);
        grid[ckey(cx[i], cy[i], cz[i])].push_back(i);
    }
    for (int i = 0; i < n; i++)
        for (long long ax = cx[i] - 1; ax <= cx[i] + 1; ax++)
            for (long long ay = cy[i] - 1; ay <= cy[i] + 1; ay++)
                for (long long az = cz[i] - 1; az <= cz[i] + 1; az++) {
                    auto it = grid.find(ckey(ax, ay, az));
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        if (j <= i) continue;
                        double d2 = 0;
                        for (int k = 0; k < 3; k++) d2 += (pos3[3 * i + k] - pos3[3 * j + k]) * (pos3[3 * i + k] - pos3[3 * j + k]);
                        const double lim = R[i] + R[j] + link;
                        if (d2 < lim * lim) f(i, j);
                    }
                }
}

// Union-find over the pairs closer than R_i + R_j + link.
std::vector<int> components(int n, const double* pos3, const double* R, double link)
{
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    near_pairs(n, pos3, R, link, [&](int i, int j) { parent[find(i)] = find(j); });
    std::vector<int> comp(n);
    for (int i = 0; i < n; i++) comp[i] = find(i);
    return comp;
}

// The first contact of the projectile (balls n_t..) moving along +x: for each pair,
// |d + s x|^2 = (R_i + R_j)^2 gives the travel s at which it touches (the first root;
// a pair with s < 0 is moving apart).  Returns false when no pair ever touches.
struct Contact {
    int i = -1, j = -1;
    double s = INFINITY;
};

bool first_contact(int n_t, int n_p, const double* pos3, const double* R, Contact& c)
{
    c = Contact{};
    for (int j = n_t; j < n_t + n_p; j++)
        for (int i = 0; i < n_t; i++) {
            const double dx = pos3[3 * j] - pos3[3 * i], dy = pos3[3 * j + 1] - pos3[3 * i + 1],
                         dz = pos3[3 * j + 2] - pos3[3 * i + 2];
            const double S = R[i] + R[j];
            const double disc = dx * dx - (dx * dx + dy * dy + dz * dz - S * S);
            if (disc < 0) continue;
            const double s = -dx - std::sqrt(disc);
            if (s >= 0 && s < c.s) c = Contact{i, j, s};
        }
    return c.i >= 0;
}

}  // namespace

extern "C" {

int sl_gen_collision(int n_t, int n_p, unsigned long long seed, double speed, double b_frac, double gap,
                     double r_med, double sigma, double rho, double* pos3, double* vel3, double* w3, double* R,
                     double* m, double* moi, double* geom)
{
    if (n_t < 1 || n_p < 1 || !(speed >= 0) || !(b_frac >= 0 && b_frac <= 1) || !(gap >= 0) || !pos3 || !vel3 ||
        !w3 || !R || !m || !moi)
        return SL_E_ARG;
    // two independent aggregates: target from seed, projectile from a derived seed
    const unsigned long long seed_p = seed ^ 0xD1B54A32D192ED03ull;
    int rc = sl_gen_bpca(n_t, seed, r_med, sigma, rho, 0.0, pos3, vel3, w3, R, m, moi);
    if (rc) return rc;
    rc = sl_gen_bpca(n_p, seed_p, r_med, sigma, rho, 0.0, pos3 + 3 * (size_t)n_t, vel3 + 3 * (size_t)n_t,
                     w3 + 3 * (size_t)n_t, R + n_t, m + n_t, moi + n_t);
    if (rc) return rc;
    // random orientations (each body is centred on its centre of mass)
    Rng rng(seed ^ 0x94D049BB133111EBull);
    for (int body = 0; body < 2; body++) {
        double Q[9];
        random_rotation(rng, Q);
        const int b0 = body ? n_t : 0, nb = body ? n_p : n_t;
        for (int i = b0; i < b0 + nb; i++) {
            const double x = pos3[3 * i], y = pos3[3 * i + 1], z = pos3[3 * i + 2];
            pos3[3 * i] = Q[0] * x + Q[1] * y + Q[2] * z;
            pos3[3 * i + 1] = Q[3] * x + Q[4] * y + Q[5] * z;
            pos3[3 * i + 2] = Q[6] * x + Q[7] * y + Q[8] * z;
        }
    }
    const Body T = body_of(pos3, R, m, 0, n_t), P = body_of(pos3, R, m, n_t, n_p);
    double r_max = 0;
    for (int i = 0; i < n_t + n_p; i++) r_max = std::max(r_max, R[i]);
    // projectile on the -x side with surface clearance gap * r_max, offset by the
    // impact parameter b along y, moving along +x; total momentum zero
    const double b = b_frac * (T.radius + P.radius);
    const double shift[3] = {-(T.radius + P.radius + gap * r_max), b, 0.0};
    const double M = T.M + P.M;
    const double vt = -speed * P.M / M, vp = speed * T.M / M;
    for (int i = 0; i < n_t + n_p; i++) {
        const bool proj = i >= n_t;
        if (proj)
            for (int k = 0; k < 3; k++) pos3[3 * i + k] += shift[k];
        vel3[3 * i] = proj ? vp : vt;
        vel3[3 * i + 1] = 0;
        vel3[3 * i + 2] = 0;
    }
    if (geom) {
        geom[0] = T.radius;
        geom[1] = P.radius;
        geom[2] = b;
        geom[3] = T.M;
        geom[4] = P.M;
        geom[5] = r_max;
    }
    return SL_OK;
}

int sl_collision_outcome(int n_t, int n_p, const double* pos3, const double* vel3, const double* R,
                         const double* m, double link, sl_outcome* out)
{
    if (n_t < 1 || n_p < 1 || !pos3 || !vel3 || !R || !m || !out || !(link >= 0)) return SL_E_ARG;
    const int n = n_t + n_p;
    const std::vector<int> comp = components(n, pos3, R, link);
    std::unordered_map<int, int> idx;
    std::vector<double> cm, ct, cp;         // mass, target mass, projectile mass per component
    std::vector<std::array<double, 3>> mv;  // momentum per component
    for (int i = 0; i < n; i++) {
        auto it = idx.find(comp[i]);
        int c;
        if (it == idx.end()) {
            c = (int)cm.size();
            idx[comp[i]] = c;
            cm.push_back(0);
            ct.push_back(0);
            cp.push_back(0);
            mv.push_back({0, 0, 0});
        } else {
            c = it->second;
        }
        cm[c] += m[i];
        (i < n_t ? ct[c] : cp[c]) += m[i];
        for (int k = 0; k < 3; k++) mv[c][k] += m[i] * vel3[3 * i + k];
    }
    const int nc = (int)cm.size();
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cm[a] > cm[b]; });
    double Mt = 0, Mp = 0;
    for (int c = 0; c < nc; c++) {
        Mt += ct[c];
        Mp += cp[c];
    }
    const double Mtot = Mt + Mp;
    const int c1 = order[0], c2 = nc > 1 ? order[1] : -1;
    const double f1 = cm[c1] / Mtot, f2 = c2 >= 0 ? cm[c2] / Mtot : 0.0;
    // the component holding most of the target's mass
    int ctgt = 0;
    for (int c = 1; c < nc; c++)
        if (ct[c] > ct[ctgt]) ctgt = c;
    sl_outcome o{};
    o.n_components = nc;
    o.largest_mass_frac = f1;
    o.second_mass_frac = f2;
    o.mass_transfer = cp[ctgt] / Mp;
    for (int k = 0; k < 3; k++) {
        o.v_largest[k] = mv[c1][k] / cm[c1];
        o.v_second[k] = c2 >= 0 ? mv[c2][k] / cm[c2] : 0.0;
    }
    const bool merged = ct[c1] >= 0.5 * Mt && cp[c1] >= 0.5 * Mp && f1 >= 0.9;
    const bool two_bodies = c2 >= 0 && 1.0 - f1 - f2 < 0.1 &&
                            ((ct[c1] >= 0.5 * Mt && cp[c2] >= 0.5 * Mp) || (cp[c1] >= 0.5 * Mp && ct[c2] >= 0.5 * Mt));
    o.label = merged ? SL_OUTCOME_STICK : (two_bodies ? SL_OUTCOME_BOUNCE : SL_OUTCOME_FRAGMENT);
    *out = o;
    return SL_OK;
}

int sl_collision_descriptors(int n_t, int n_p, const double* pos3, const double* R, const double* m, double speed,
                             double b_frac, double link, double* desc)
{
    if (n_t < 1 || n_p < 1 || !pos3 || !R || !m || !desc || !(link >= 0)) return SL_E_ARG;
    double r_mean = 0;
    for (int i = 0; i < n_t + n_p; i++) r_mean += R[i];
    r_mean /= (n_t + n_p);
    double out[SL_N_DESCRIPTORS];
    for (int body = 0; body < 2; body++) {
        const int b0 = body ? n_t : 0, nb = body ? n_p : n_t;
        const Body B = body_of(pos3, R, m, b0, nb);
        double rg2 = 0;
        for (int i = b0; i < b0 + nb; i++)
            for (int k = 0; k < 3; k++) rg2 += m[i] * (pos3[3 * i + k] - B.com[k]) * (pos3[3 * i + k] - B.com[k]);
        const double rg = std::sqrt(rg2 / B.M) / r_mean;
        // mass-radius fractal dimension N ~ (Rg / r)^D, D = ln N / ln(Rg / r) (3 for a
        // single sphere or a compact cluster with Rg below r)
        const double df = (nb > 1 && rg > 1.0) ? std::log((double)nb) / std::log(rg) : 3.0;
        // coordination: touching or linked neighbours per ball
        long long links = 0;
        near_pairs(nb, pos3 + 3 * (size_t)b0, R + b0, link, [&](int, int) { links++; });
        const double coord = 2.0 * links / nb;
        const int o0 = body ? 4 : 0;
        out[o0 + 0] = std::log((double)nb);
        out[o0 + 1] = rg;
        out[o0 + 2] = df;
        out[o0 + 3] = coord;
    }
    out[8] = std::log10(std::max(speed, 1e-30));
    out[9] = b_frac;
    double Mt = 0, Mp = 0;
    for (int i = 0; i < n_t; i++) Mt += m[i];
    for (int i = n_t; i < n_t + n_p; i++) Mp += m[i];
    out[10] = Mp / (Mt + Mp);
    out[11] = std::log10(std::max(0.5 * (Mt * Mp / (Mt + Mp)) * speed * speed, 1e-300));
    // impact geometry along the approach line (+x): shadows, column depth, first contact
    std::vector<int> hit_t(n_t, 0), col_p(n_p, 0);
    for (int j = n_t; j < n_t + n_p; j++)
        for (int i = 0; i < n_t; i++) {
            const double dy = pos3[3 * j + 1] - pos3[3 * i + 1], dz = pos3[3 * j + 2] - pos3[3 * i + 2];
            const double S = R[i] + R[j];
            if (dy * dy + dz * dz < S * S) {
                hit_t[i] = 1;
                col_p[j - n_t]++;
            }
        }
    double sh_p = 0, sh_t = 0, depth = 0;
    for (int i = 0; i < n_t; i++) sh_t += hit_t[i] ? m[i] : 0.0;
    for (int j = 0; j < n_p; j++) {
        sh_p += col_p[j] ? m[n_t + j] : 0.0;
        depth += m[n_t + j] * col_p[j];
    }
    Contact c;
    const bool met = first_contact(n_t, n_p, pos3, R, c);
    out[12] = met ? 1.0 : 0.0;
    out[13] = sh_p / Mp;
    out[14] = sh_t / Mt;
    out[15] = std::log1p(depth / Mp);
    for (int k = 16; k < 21; k++) out[k] = 0.0;
    if (met) {
        const Body T = body_of(pos3, R, m, 0, n_t), P = body_of(pos3, R, m, n_t, n_p);
        const int i = c.i, j = c.j;
        // centre-line of the touching pair at contact, target ball -> projectile ball
        double u[3] = {pos3[3 * j] + c.s - pos3[3 * i], pos3[3 * j + 1] - pos3[3 * i + 1],
                       pos3[3 * j + 2] - pos3[3 * i + 2]};
        const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double& x : u) x /= un;
        double dt2 = 0, dp2 = 0;
        for (int k = 0; k < 3; k++) {
            const double cp = pos3[3 * i + k] + u[k] * R[i];   // contact point
            const double pc = P.com[k] + (k == 0 ? c.s : 0.0);  // projectile COM at contact
            dt2 += (cp - T.com[k]) * (cp - T.com[k]);
            dp2 += (cp - pc) * (cp - pc);
        }
        auto neighbours = [&](int b0, int nb, int q) {
            int cnt = 0;
            for (int k = b0; k < b0 + nb; k++) {
                if (k == q) continue;
                double d2 = 0;
                for (int e = 0; e < 3; e++) d2 += (pos3[3 * k + e] - pos3[3 * q + e]) * (pos3[3 * k + e] - pos3[3 * q + e]);
                const double lim = R[k] + R[q] + link;
                cnt += d2 < lim * lim;
            }
            return cnt;
        };
        out[16] = -u[0];
        out[17] = std::sqrt(dt2) / T.radius;
        out[18] = std::sqrt(dp2) / P.radius;
        out[19] = neighbours(0, n_t, i);
        out[20] = neighbours(n_t, n_p, j);
    }
    for (int k = 0; k < SL_N_DESCRIPTORS; k++) desc[k] = out[k];
    return SL_OK;
}

int sl_collision_descriptors_batch(int count, const int* n_t, const int* n_p, const long long* first,
                                   const double* pos3, const double* R, const double* m, const double* speed,
                                   const double* b_frac, double link, double* desc, int nthreads)
{
    if (count < 0) return SL_E_ARG;
    if (count == 0) return SL_OK;
    if (!n_t || !n_p || !first || !pos3 || !R || !m || !speed || !b_frac || !desc) return SL_E_ARG;
    // collisions are handed out one at a time: their costs differ by orders of magnitude
    std::atomic<int> next{0};
    std::atomic<int> err{SL_OK};
    auto work = [&] {
        for (int k = next.fetch_add(1); k < count; k = next.fetch_add(1)) {
            const long long o = first[k];
            const int rc = o < 0 ? SL_E_ARG
                                 : sl_collision_descriptors(n_t[k], n_p[k], pos3 + 3 * o, R + o, m + o, speed[k],
                                                            b_frac[k], link, desc + (size_t)k * SL_N_DESCRIPTORS);
            if (rc != SL_OK) {
                int expected = SL_OK;
                err.compare_exchange_strong(expected, rc);
            }
        }
    };
    int nt = nthreads > 0 ? nthreads : std::min(16, (int)std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, std::min(count, 64)));
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return err.load();
}

// Moves the projectile (balls n_t..) along x so that its first contact with the
// target comes after a surface gap `gap` of travel at the current relative motion
// (+x).  Returns 1 and the shift applied when the bodies meet on that line, 0 (no
// shift) when they miss.
int sl_collision_approach(int n_t, int n_p, double* pos3, const double* R, double gap, double* shift)
{
    if (n_t < 1 || n_p < 1 || !pos3 || !R || !(gap >= 0)) return SL_E_ARG;
    Contact c;
    if (shift) *shift = 0;
    if (!first_contact(n_t, n_p, pos3, R, c)) return 0;
    const double d = c.s - gap;
    for (int j = n_t; j < n_t + n_p; j++) pos3[3 * j] += d;
    if (shift) *shift = d;
    return 1;
}

}  // extern "C"
