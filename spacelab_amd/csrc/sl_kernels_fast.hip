// Fast path: every unique pair evaluated once (Newton's third law), far-field
// algebra without square roots, near pairs in the reference's arithmetic.
// Host-side management and launches; the kernels are in sl_fast_kernels.hpp.
//
// Step structure (two kernels per step once running): k_tiles (pair triangle
// in 64x64 tiles, one wavefront each) then k_ball (per-ball sums, near pairs,
// kick, energies and the next step's drift), ping-ponging between two state sets.
#include "spacelab/sl_api.h"
#include "sl_fast_kernels.hpp"
#include "sl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace sl {

using namespace fastk;

namespace {

// Padding balls [n, n_pad): far apart, zero cohesion and gravity weights, so the
// tile kernel can read whole 64-ball blocks without bounds tests.
__global__ void k_pad(State s, FastAux aux, int n, int n_pad)
{
    const int b = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_pad) return;
    const double x = 1e12 * (double)(b + 1);
    s.f[PX][b] = x;
    s.f[PY][b] = 0;
    s.f[PZ][b] = 0;
    s.f[RAD][b] = 1e-30;
    aux.q[b] = 0;
    aux.g[b] = 0;
    aux.r32[b] = 0.f;
    aux.q32[b] = 0.f;
    aux.g32[b] = 0.f;
}

__global__ void k_aux(State s, FastAux aux, int n, double sqrtC, double sqrtG, double L0inv, double M0inv)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const double R = s.f[RAD][b], m = s.f[MASS][b];
    aux.q[b] = sqrtC * R * R * R;
    aux.g[b] = sqrtG * m;
    const double r = R * L0inv;
    aux.r32[b] = (float)r;
    aux.q32[b] = (float)(r * r * r);
    aux.g32[b] = (float)(m * M0inv);
}

__global__ void k_copy_fields(State dst, State src, int n, int f0, int f1)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    for (int f = f0; f < f1; f++) dst.f[f][b] = src.f[f][b];
}

// Lane-layout probe for the DPP rotation the tile kernels rely on.
__global__ void k_probe_rot(int* out)
{
    const int l = threadIdx.x;
    out[l] = __float_as_int(rot1(__int_as_float(l)));
    out[64 + l] = (int)rot1((double)l);
    out[128 + l] = (int)rotk((double)l, 31);
}

State& cur_state(FastBuffers& fb, State& canon) { return fb.cur_is_alt ? fb.alt : canon; }
State& nxt_state(FastBuffers& fb, State& canon) { return fb.cur_is_alt ? canon : fb.alt; }

int nblocks64(int n) { return (n + TB - 1) / TB; }

}  // namespace

int fast_max_blocks(int n_max)
{
    const int nb = nblocks64(n_max);
    const long long tiles = (long long)nb * (nb + 1) / 2;
    // the finest tile split puts one tile per block
    return (int)tiles + (n_max + BPB - 1) / BPB;
}

int fast_energy_blocks(int n) { return (n + BPB - 1) / BPB; }

int fast_alloc(FastBuffers& fb, int n_max, int n_pad)
{
    fb.n_max = n_max;
    fb.n_pad = n_pad;
    const int nb = nblocks64(n_max);
    const long long tiles = (long long)nb * (nb + 1) / 2;
    if (hipMalloc(&fb.alt_base, sizeof(double) * (size_t)NFIELD * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.q, sizeof(double) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.g, sizeof(double) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.r32, sizeof(float) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.q32, sizeof(float) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.g32, sizeof(float) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.tiles, sizeof(int2) * (size_t)tiles) != hipSuccess) return -1;
    if (hipMalloc(&fb.part, sizeof(double) * (size_t)nb * 3 * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.nb_count, sizeof(int) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.nb_ent, sizeof(NearEntry) * (size_t)n_pad * MAXNB) != hipSuccess) return -1;
    if (hipMalloc(&fb.flags, sizeof(int) * 4) != hipSuccess) return -1;
    if (hipMalloc(&fb.near_total, sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemset(fb.alt_base, 0, sizeof(double) * (size_t)NFIELD * n_pad) != hipSuccess) return -1;
    if (hipMemset(fb.nb_count, 0, sizeof(int) * n_pad) != hipSuccess) return -1;
    if (hipMemset(fb.flags, 0, sizeof(int) * 4) != hipSuccess) return -1;
    if (hipMemset(fb.near_total, 0, sizeof(unsigned long long)) != hipSuccess) return -1;
    for (int f = 0; f < NFIELD; f++) fb.alt.f[f] = fb.alt_base + (size_t)f * n_pad;
    return 0;
}

void fast_free(FastBuffers& fb)
{
    void* ptrs[] = {fb.alt_base, fb.aux.q, fb.aux.g, fb.aux.r32, fb.aux.q32, fb.aux.g32, fb.tiles,
                    fb.part, fb.nb_count, fb.nb_ent, fb.flags, fb.near_total};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    fb = FastBuffers{};
}

int fast_sync_from(FastBuffers& fb, State& canon, const Phys& P, int n, hipStream_t st, std::string* err)
{
    auto fail = [&](const char* what) {
        if (err) *err = std::string("fast path set-up: ") + what;
        return SL_E_HIP;
    };
    if (!fb.alt_base) return 0;
    // the radius/mass/moi arrays are shared by both state sets
    fb.alt.f[RAD] = canon.f[RAD];
    fb.alt.f[MASS] = canon.f[MASS];
    fb.alt.f[MOI] = canon.f[MOI];
    fb.cur_is_alt = false;
    fb.drifted = false;
    // units of the mixed path: median radius and mass
    std::vector<double> R(n), m(n);
    if (hipMemcpyAsync(R.data(), canon.f[RAD], sizeof(double) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(m.data(), canon.f[MASS], sizeof(double) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail("read radii");
    std::nth_element(R.begin(), R.begin() + n / 2, R.end());
    std::nth_element(m.begin(), m.begin() + n / 2, m.end());
    fb.L0 = R[n / 2];
    fb.M0 = m[n / 2];
    const double C = P.Ha / 6 * 64;
    k_aux<<<(n + 255) / 256, 256, 0, st>>>(canon, fb.aux, n, std::sqrt(C), std::sqrt(P.G), 1.0 / fb.L0, 1.0 / fb.M0);
    const int np = fb.n_pad;
    if (np > n) {
        k_pad<<<(np - n + 255) / 256, 256, 0, st>>>(canon, fb.aux, n, np);
        k_pad<<<(np - n + 255) / 256, 256, 0, st>>>(fb.alt, fb.aux, n, np);
    }
    if (hipGetLastError() != hipSuccess) return fail("aux kernels");
    if (fb.n != n) {
        const int nb = nblocks64(n);
        std::vector<int2> t;
        t.reserve((size_t)nb * (nb + 1) / 2);
        for (int I = 0; I < nb; I++)
            for (int J = 0; J <= I; J++) t.push_back(make_int2(I, J));
        if (hipMemcpyAsync(fb.tiles, t.data(), sizeof(int2) * t.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail("tile table");
        fb.n_tiles = (int)t.size();
        fb.n = n;
        fb.split = 0;   // re-tune for the new size
    }
    if (launch_zero_i32(fb.nb_count, fb.n_pad, st) != hipSuccess || launch_zero_i32(fb.flags, 4, st) != hipSuccess)
        return fail("clear near lists");
    return 0;
}

int fast_sync_to(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err)
{
    if (!fb.cur_is_alt) return 0;
    k_copy_fields<<<(n + 255) / 256, 256, 0, st>>>(canon, fb.alt, n, PX, RAD);
    if (hipGetLastError() != hipSuccess) {
        if (err) *err = "fast path: copy state";
        return SL_E_HIP;
    }
    fb.cur_is_alt = false;
    return 0;
}

int fast_begin_init(FastBuffers& fb, State& canon, int n, hipStream_t st, std::string* err)
{
    if (launch_copy_prev(cur_state(fb, canon), n, st) != hipSuccess) {
        if (err) *err = "fast path: copy positions";
        return SL_E_HIP;
    }
    fb.drifted = false;
    return 0;
}

int fast_drift(FastBuffers& fb, State& canon, int n, double dt, hipStream_t st, std::string* err)
{
    if (fb.drifted) {
        fb.drifted = false;
        return 0;
    }
    if (launch_drift(cur_state(fb, canon), n, dt, st) != hipSuccess) {
        if (err) *err = "fast path: drift";
        return SL_E_HIP;
    }
    return 0;
}

namespace {

template <bool WRITE, bool GRAV, int SPLIT>
void launch_tiles(const TileArgs& a, int precision, hipStream_t st)
{
    constexpr int TPB = NW / SPLIT;
    const int blocks = (a.n_tiles + TPB - 1) / TPB;
    if (precision == SL_PREC_MIXED)
        k_tiles_mixed<WRITE, GRAV, SPLIT><<<blocks, NT, 0, st>>>(a);
    else
        k_tiles_f64<WRITE, GRAV, SPLIT><<<blocks, NT, 0, st>>>(a);
}

template <int SPLIT>
void launch_tiles_split(const TileArgs& a, int precision, bool write, bool grav, hipStream_t st)
{
    if (write)
        grav ? launch_tiles<true, true, SPLIT>(a, precision, st) : launch_tiles<true, false, SPLIT>(a, precision, st);
    else
        grav ? launch_tiles<false, true, SPLIT>(a, precision, st) : launch_tiles<false, false, SPLIT>(a, precision, st);
}

void launch_tiles_any(const TileArgs& a, int split, int precision, bool write, bool grav, hipStream_t st)
{
    if (split == 4)
        launch_tiles_split<4>(a, precision, write, grav, st);
    else if (split == 2)
        launch_tiles_split<2>(a, precision, write, grav, st);
    else
        launch_tiles_split<1>(a, precision, write, grav, st);
}

int tile_blocks_for(int n_tiles, int split)
{
    const int tpb = NW / split;
    return (n_tiles + tpb - 1) / tpb;
}

TileArgs tile_args(FastBuffers& fb, State& c, const Phys& P, double* pe_part)
{
    TileArgs a;
    a.px = c.f[PX];
    a.py = c.f[PY];
    a.pz = c.f[PZ];
    a.R = c.f[RAD];
    a.aux = fb.aux;
    a.cur = c;
    a.P = P;
    a.n_pad = fb.n_pad;
    a.n_tiles = fb.n_tiles;
    a.tiles = fb.tiles;
    a.L0inv2 = 1.0 / (fb.L0 * fb.L0);
    a.Kv = (P.Ha / 6 * 64) / (fb.L0 * fb.L0);
    a.Kg = P.G * fb.M0 * fb.M0;
    a.part = fb.part;
    a.nb_count = fb.nb_count;
    a.nb_ent = fb.nb_ent;
    a.flags = fb.flags;
    a.pe_part = pe_part;
    return a;
}

// Autotuner: the number of wavefronts sharing one 64x64 tile trades per-wave
// set-up against occupancy and load balance, which depend on the system size and
// the machine.  Time each choice on the live state (HIP events on the context's
// stream, best of three launches) and keep the fastest; the near lists these
// trial launches fill are cleared again.  SL_TILE_SPLIT=1|2|4 forces a choice.
int tune_split(FastBuffers& fb, const TileArgs& a, int precision, bool grav, hipStream_t st)
{
    const char* env = std::getenv("SL_TILE_SPLIT");
    if (env) {
        const int v = std::atoi(env);
        if (v == 1 || v == 2 || v == 4) return v;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return 1;
    }
    int best = 1;
    float best_ms = 1e30f;
    const int choices[3] = {1, 2, 4};
    for (int c : choices) {
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            hipEventRecord(e0, st);
            launch_tiles_any(a, c, precision, false, grav, st);
            hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            ms_min = ms < ms_min ? ms : ms_min;
            launch_zero_i32(fb.nb_count, fb.n_pad, st);
        }
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = c;
        }
    }
    launch_zero_i32(fb.nb_count, fb.n_pad, st);
    launch_zero_i32(fb.flags, 4, st);
    hipStreamSynchronize(st);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return best;
}

}  // namespace

int fast_pairs(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write,
               double* pe_part, unsigned long long* contacts, hipStream_t st, std::string* err)
{
    (void)contacts;
    (void)n;
    State& c = cur_state(fb, canon);
    const TileArgs a = tile_args(fb, c, P, pe_part);
    const bool g = P.far_gravity != 0;
    const int key = precision * 2 + (g ? 1 : 0);
    if (fb.split == 0 || fb.split_key != key) {
        fb.split = tune_split(fb, a, precision, g, st);
        fb.split_key = key;
    }
    launch_tiles_any(a, fb.split, precision, write != 0, g, st);
    fb.tile_blocks = tile_blocks_for(fb.n_tiles, fb.split);
    if (hipGetLastError() != hipSuccess) {
        if (err) *err = "fast path: tile kernel launch";
        return SL_E_HIP;
    }
    return 0;
}

int fast_ball(FastBuffers& fb, State& canon, const Phys& P, int n, int precision, int write, int kick,
              int drift_next, double* pe_part, double* e_part, unsigned long long* contacts,
              int* nonfinite, hipStream_t st, int* n_pe_blocks, std::string* err)
{
    const int tile_blocks = fb.tile_blocks;
    BallArgs a;
    a.cur = cur_state(fb, canon);
    a.nxt = nxt_state(fb, canon);
    a.P = P;
    a.n = n;
    a.n_pad = fb.n_pad;
    a.nblk = nblocks64(n);
    a.part = fb.part;
    a.nb_count = fb.nb_count;
    a.nb_ent = fb.nb_ent;
    a.write = write;
    a.kick = kick;
    a.drift_next = drift_next;
    a.compensated = precision == SL_PREC_MIXED;
    a.pe_part = pe_part + tile_blocks;
    a.e_part = e_part;
    a.contacts = contacts;
    a.near_total = fb.near_total;
    a.nonfinite = nonfinite;
    const int blocks = (n + BPB - 1) / BPB;
    k_ball<<<blocks, 256, 0, st>>>(a);
    if (hipGetLastError() != hipSuccess) {
        if (err) *err = "fast path: ball kernel launch";
        return SL_E_HIP;
    }
    fb.cur_is_alt = !fb.cur_is_alt;
    fb.drifted = drift_next != 0;
    if (n_pe_blocks) *n_pe_blocks = write ? tile_blocks + blocks : 0;
    return 0;
}

int fast_check(FastBuffers& fb, hipStream_t st, std::string* err)
{
    int h = 0;
    if (hipMemcpyAsync(&h, fb.flags, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        if (err) *err = "fast path: read flags";
        return SL_E_HIP;
    }
    if (h) {
        if (err)
            *err = "near-pair list overflow: a sphere has more than 32 partners within h_min of "
                   "contact; results of this call are invalid";
        return SL_E_CAPACITY;
    }
    return 0;
}

hipError_t fast_reset_counters(FastBuffers& fb, hipStream_t st) { return launch_zero_u64(fb.near_total, 1, st); }

long long fast_read_near_total(FastBuffers& fb, hipStream_t st)
{
    unsigned long long h = 0;
    if (hipMemcpyAsync(&h, fb.near_total, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    return (long long)(h / 2);   // each near pair is listed by both of its balls
}

}  // namespace sl

// Lane layout of the register rotation used by the fast tile kernels: out[l] =
// value of lane (l+1)%64 (binary32 and binary64) and of lane (l+31)%64.
extern "C" int sl_probe_rotation(int* out192)
{
    if (!out192) return SL_E_ARG;
    int* d = nullptr;
    if (hipMalloc(&d, sizeof(int) * 192) != hipSuccess) return SL_E_HIP;
    int rc = SL_OK;
    sl::k_probe_rot<<<1, 64>>>(d);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out192, d, sizeof(int) * 192, hipMemcpyDeviceToHost) != hipSuccess)
        rc = SL_E_HIP;
    hipFree(d);
    return rc;
}
