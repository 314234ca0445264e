// Internal interfaces between the C ABI (sl_api.hip) and the kernel translation units.
#pragma once

#include "spacelab/sl_api.h"
#include "sl_common.hpp"

#include <map>
#include <string>
#include <vector>

namespace sl {

// ---- ordered (bitwise) path: sl_kernels_ordered.hip
hipError_t launch_drift(const State& s, int n, double dt, hipStream_t st);
hipError_t launch_copy_prev(const State& s, int n, hipStream_t st);
// Launch shapes of the pair pass (all give the same bits); shape = index below ordered_shapes().
int ordered_shapes();
int ordered_default_shape(int n);
int ordered_blocks(int n, int shape);
// pe_ball: on write steps, per ball i the energy of its pairs (i, j < i) in ascending j.
hipError_t launch_pairs_ordered(const State& s, const Phys& P, int n, int write, double* pe_ball,
                                unsigned long long* contacts, hipStream_t st, int shape);
int kick_blocks(int n);
hipError_t launch_kick(const State& s, int n, double dt, int write, int kick, double* e_part,
                       int* nonfinite, hipStream_t st);
// A step's closing kick (no energies) and the next step's drift in one launch: the same
// operations as launch_kick then launch_drift, so the same bits.
hipError_t launch_kick_drift(const State& s, int n, double dt, int* nonfinite, hipStream_t st);

// ---- small utilities: sl_util.hip
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
hipError_t launch_zero_i32(int* p, int n, hipStream_t st);

// ---- fast path (Newton-3 tiles, fp64 or mixed): sl_kernels_fast.hip
constexpr int FAST_TILE = 64;       // tile edge (balls) of the Newton-3 pair kernel

// Per-ball static data derived from R and m for the tile kernel.
struct FastAux {
    double* q = nullptr;     // sqrt(Ha/6*64) R^3       (binary64 path: C_ij = q_i q_j)
    double* g = nullptr;     // sqrt(G) m               (G m_i m_j = g_i g_j)
    float* q32 = nullptr;    // sqrt(Kv/Kf) (R/L0)^3    (mixed path, scaled units)
    float* g32 = nullptr;    // sqrt(Kg/Kf) m/M0
};

// Layout of one or more independent systems in one set of state arrays: system s
// owns balls [off[s], off[s] + n[s]); off[s] is a multiple of FAST_TILE and the
// slots up to the next system are padding.  A single simulation is one system.
struct FastLayout {
    int n_sys = 0;
    std::vector<int> n, off, nb;   // balls, first ball, 64-ball blocks per system
    int n_used = 0;                // ball slots in use (padded)
    int nb_max = 0;                // largest block count of one system
};

// Lay out systems of the given sizes (FAST_TILE-aligned offsets).
FastLayout fast_layout(int n_sys, const int* n_per_sys);

// Optional hooks called around each kernel launch (profiling with HIP events).
struct LaunchHook {
    void (*before)(void* user, int kind) = nullptr;   // kind 0: pair/tile sweep, 1: per-ball kernel
    void (*after)(void* user, int kind) = nullptr;
    void* user = nullptr;
};

struct FastBuffers {
    int n_pad = 0;                // capacity of the state arrays (balls)
    int max_sys = 0;
    FastLayout lay;
    std::vector<Phys> P;          // per system
    std::vector<int> blk_sys_h;   // per 64-ball block: system (host copy)
    std::vector<int> tile_sys_h;  // per tile: system (host copy)
    // The engine keeps its own two state sets (ping-pong), with each system's balls
    // in Morton order of their positions at configuration, so that the 64-ball blocks
    // are compact in space: most tiles then have no near pair.  perm maps engine slot
    // -> canonical slot (the caller's order), within each system.
    double* eng_base = nullptr;   // 2 * NFIELD * n_pad doubles
    State eng[2] = {};
    int cur = 0;                  // engine set holding the current state
    int* perm = nullptr;          // device, n_pad
    bool canon_valid = true;      // the canonical set holds the current state
    bool eng_valid = false;       // the engine sets hold the current state
    bool drifted = false;         // the next step's drift was already applied
    FastAux aux{};
    double4* bnd = nullptr;       // per 64-ball block: bounding box of the positions (lo, hi)
    float4* xl = nullptr;         // per ball: position relative to its block's box centre, R (units of L0)
    float* gq_blk = nullptr;      // per 64-ball block: largest g/q of the mixed path's scaled weights
    double L0 = 1, M0 = 1;        // length and mass units of the mixed path
    double Kf = 1;                // mixed path: force scale of the binary32 factor
    int* blk_sys = nullptr;       // device copies of the layout
    int* sys_off = nullptr;
    int* sys_n = nullptr;
    int* sys_nb = nullptr;
    Phys* sysP = nullptr;
    int4* tiles = nullptr;        // {I, J, system, first block of system}, I >= J
    int n_tiles = 0, cap_tiles = 0;
    double* part = nullptr;       // far-field force partials [K][xyz][n_pad]
    unsigned long long* nbm = nullptr;   // near-partner masks [K][n_pad]
    int part_slots = 0;           // K capacity of part and nbm
    int* flags = nullptr;         // [1] non-finite state
    unsigned long long* blk_ctr = nullptr;    // per 64-ball block: touching and near pairs (running sums)
    int cap_blocks = 0;
    double* e_blk = nullptr;      // per 64-ball block: near-pair PE, KE, p, L, contacts, nonfinite
    double* pe_tile = nullptr;    // per tile: pair energy (write steps)
    double* h_e_blk = nullptr;    // pinned host copies
    double* h_pe_tile = nullptr;
    int precision = 0;
    int split = 0;                // wavefronts per 64x64 tile (0: not tuned yet)
    bool queue_launch = false;    // tiles from the work queue (k_tiles_q) rather than one per workgroup
    int forced_queue = 0;         // 1 forces the per-workgroup launch, 2 the work queue
    int* queue = nullptr;         // k_tiles_q's per-segment next groups and finished workgroups (0 between launches)
    int split_key = -1;           // (precision, gravity) the choice was tuned for
    int forced_split = 0;         // 1, 2 or 4 forces the choice
    int n_cu = 256;               // compute units of the device (k_ball's shape)
    int forced_ball_waves = 0;    // 4 or 8 forces k_ball's wavefronts per workgroup
    bool use_graphs = true;
    LaunchHook hook{};
    struct GraphKey {
        int len, write_last, parity;
        bool operator<(const GraphKey& o) const
        {
            return len != o.len ? len < o.len : write_last != o.write_last ? write_last < o.write_last : parity < o.parity;
        }
    };
    std::map<GraphKey, hipGraphExec_t> graphs;
};

constexpr int FAST_E_STRIDE = 10;   // doubles per energy block

int fast_alloc(FastBuffers& fb, int n_pad, int max_sys);
void fast_free(FastBuffers& fb);
// Build the tile tables and per-ball constants for a layout already placed in canon.
int fast_configure(FastBuffers& fb, State& canon, const FastLayout& lay, const Phys* P, int precision,
                   hipStream_t st, std::string* err);
// New physics parameters for the same layout (re-derives per-ball constants).
int fast_set_phys(FastBuffers& fb, State& canon, const Phys* P, int precision, hipStream_t st, std::string* err);
// Bring the canonical state set up to date from the engine's sets.
int fast_sync_to(FastBuffers& fb, State& canon, hipStream_t st, std::string* err);
// The canonical set was changed by other code (the ordered path): the engine reloads it
// before its next step.
void fast_canon_changed(FastBuffers& fb);
// Step-0 pass: forces of the current positions and the energies (one per system).
int fast_init(FastBuffers& fb, State& canon, sl_energy* e_sys, hipStream_t st, std::string* err);
// n_steps steps; the energies of write steps go to e_writes[w * n_sys + s] for the first
// cap writes.  Returns the number of write steps or a negative error.
int fast_steps(FastBuffers& fb, State& canon, int n_steps, long long* step_count, int write_every,
               sl_energy* e_writes, int cap, hipStream_t st, std::string* err);
int fast_max_tile_blocks(int n_pad);
void fast_clear_graphs(FastBuffers& fb);
// wavefronts per k_ball workgroup for the current layout (forced, or by block count)
int fast_ball_waves(const FastBuffers& fb);
hipError_t fast_reset_counters(FastBuffers& fb, hipStream_t st);
int fast_read_counters(FastBuffers& fb, hipStream_t st, long long* contacts, long long* near_pairs);

// A profiling tool (rocprofv3) is attached to this process: launch kernels directly, no
// graph replay.
bool tool_attached();

// ---- the autotuner's persistent cache: sl_tune.cpp
std::string tune_key(int precision, int gravity, int n_sys, int n_used, int nb_max, int n_tiles);
int tune_cache_get(const std::string& key);   // the stored split, 0 if none
void tune_cache_put(const std::string& key, int split, const float* ms, int n);

}  // namespace sl
