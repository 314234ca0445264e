// Internal interfaces between the C ABI (sl_api.hip) and the kernel translation units.
#pragma once

#include "sl_common.hpp"

#include <string>

namespace sl {

// ---- ordered (bitwise) path: sl_kernels_ordered.hip
hipError_t launch_drift(const State& s, int n, double dt, hipStream_t st);
hipError_t launch_copy_prev(const State& s, int n, hipStream_t st);
int ordered_blocks(int n, int* bi);
hipError_t launch_pairs_ordered(const State& s, const Phys& P, int n, int write, double* pe_part,
                                unsigned long long* contacts, hipStream_t st);
int kick_blocks(int n);
hipError_t launch_kick(const State& s, int n, double dt, int write, int kick, double* e_part,
                       int* nonfinite, hipStream_t st);

// ---- small utilities: sl_util.hip
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
hipError_t launch_zero_i32(int* p, int n, hipStream_t st);

// ---- fast path (Newton-3 tiles, fp64 or mixed): sl_kernels_fast.hip
constexpr int FAST_TILE = 64;       // tile edge (balls) of the Newton-3 pair kernel
constexpr int FAST_TILES_PER_BLOCK = 4;
constexpr int FAST_MAX_NEAR = 32;   // near partners (touching or within h_min) per ball

// Per-ball static data derived from R and m for the tile kernel.
struct FastAux {
    double* q = nullptr;     // sqrt(Ha/6*64) R^3       (binary64 path: C_ij = q_i q_j)
    double* g = nullptr;     // sqrt(G) m               (G m_i m_j = g_i g_j)
    float* r32 = nullptr;    // R / L0                  (mixed path, scaled units)
    float* q32 = nullptr;    // (R / L0)^3
    float* g32 = nullptr;    // m / M0
};

// One evaluated near pair, filed with each of its two balls: the partner, the
// force and torque on this ball, and (for the reference orientation's A ball only)
// the pair energy.  contact: 0 none, 1 touching and this ball is A, 2 touching and
// this ball is B.
struct NearEntry {
    int j;
    int contact;
    double fx, fy, fz;
    double tx, ty, tz;
    double pe;
};

struct FastBuffers {
    int n_max = 0, n_pad = 0;
    int n = 0;                    // system size the tile table was built for
    double* alt_base = nullptr;   // second state set (ping-pong)
    State alt{};
    bool cur_is_alt = false;      // which set holds the current state
    bool drifted = false;         // the next step's drift was already applied
    FastAux aux{};
    double L0 = 1, M0 = 1;        // length and mass units of the mixed path
    int2* tiles = nullptr;        // (I, J) block pairs, I >= J
    int n_tiles = 0;
    double* part = nullptr;       // far-field force partials [K][xyz][n_pad]
    int* nb_count = nullptr;      // near entries filed this step, per ball
    NearEntry* nb_ent = nullptr;  // [n_pad][FAST_MAX_NEAR]
    int* flags = nullptr;         // [0] near-list overflow
    unsigned long long* near_total = nullptr;
    int split = 0;                // wavefronts per 64x64 tile (0: not tuned yet)
    int split_key = -1;           // (precision, gravity) the choice was tuned for
    int tile_blocks = 0;          // blocks of the last tile launch (its PE partials)
};

int fast_alloc(FastBuffers& fb, int n_max, int n_pad);
void fast_free(FastBuffers& fb);
int fast_max_blocks(int n_max);
int fast_sync_from(FastBuffers& fb, State& canon, const Phys& P, int n, hipStream_t st, std::string* err);
int fast_sync_to(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err);
int fast_begin_init(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err);
int fast_drift(FastBuffers& fb, State& canon, int n, double dt, hipStream_t st, std::string* err);
int fast_pairs(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write,
               double* pe_part, unsigned long long* contacts, hipStream_t st, std::string* err);
int fast_ball(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write, int kick,
              int drift_next, double* pe_part, double* e_part, unsigned long long* contacts,
              int* nonfinite, hipStream_t st, int* n_pe_blocks, std::string* err);
int fast_check(FastBuffers& fb, hipStream_t st, std::string* err);
int fast_energy_blocks(int n);   // energy partial blocks written by the fast per-ball kernel
hipError_t fast_reset_counters(FastBuffers& fb, hipStream_t st);
long long fast_read_near_total(FastBuffers& fb, hipStream_t st);

}  // namespace sl
