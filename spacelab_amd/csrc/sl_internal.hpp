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
struct FastBuffers {
    int n_max = 0, n_pad = 0;
    double* alt_base = nullptr;   // second state set (ping-pong)
    State alt{};
    bool cur_is_alt = false;      // which set holds the current state
    double* part = nullptr;       // per-tile far-field partial accelerations
    double* comp = nullptr;       // unused slot kept for layout stability
    int* nb_count = nullptr;      // near-pair count per ball
    int* nb_list = nullptr;       // near-pair partners per ball
    int* flags = nullptr;         // [0] near-list overflow
    unsigned long long* near_total = nullptr;
    float* f32 = nullptr;         // binary32 copies for the mixed path
    int max_nb = 0;
};

int fast_alloc(FastBuffers& fb, int n_max, int n_pad);
void fast_free(FastBuffers& fb);
int fast_max_blocks(int n_max);
int fast_sync_from(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err);
int fast_sync_to(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err);
int fast_begin_init(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err);
int fast_drift(FastBuffers& fb, State& canon, int n, double dt, hipStream_t st, std::string* err);
int fast_pairs(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write,
               unsigned long long* contacts, hipStream_t st, std::string* err);
int fast_ball(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write, int kick,
              double* pe_part, double* e_part, unsigned long long* contacts, int* nonfinite,
              hipStream_t st, int* n_pe_blocks, std::string* err);
hipError_t fast_reset_counters(FastBuffers& fb, hipStream_t st);
long long fast_read_near_total(FastBuffers& fb, hipStream_t st);

}  // namespace sl
