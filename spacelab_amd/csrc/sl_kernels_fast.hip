// Fast path engine: every unique pair evaluated once (Newton's third law), far-field
// algebra without square roots, near pairs in the reference's arithmetic.  Runs one
// system or a batch of independent systems laid out in one set of state arrays; the
// kernels are in sl_fast_kernels.hpp.
//
// The engine works on its own two state sets (ping-pong) in which each system's balls
// are in Morton order of their positions at configuration; the caller's (canonical)
// set is refreshed on demand.  Per step: k_tiles (the far-field sweep of every tile,
// with the per-pair near test only where the blocks' bounding boxes come within h_min;
// near pairs are recorded as partner masks) and k_ball (per-ball sums of the far-field
// partials and of the near pairs, evaluated there in the reference arithmetic, then
// kick, energies, the next step's drift and block bounds).  Runs of steps between
// write steps are captured once as HIP graphs and replayed.
#include "spacelab/sl_api.h"
#include "sl_fast_kernels.hpp"
#include "sl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <numeric>
#include <vector>

namespace sl {

using namespace fastk;

namespace {

// Padding slots (outside every system's balls): far apart, zero radius and zero
// cohesion and gravity weights, so the tile kernels can read whole 64-ball blocks
// without tests.  With R = 0 a padding pair's force and energy terms are exactly
// zero (s = t, so log(s/t) = 0), which keeps a system's sums independent of where
// its padding lies.
__global__ void k_pad_sys(State s, FastAux aux, const int* blk_sys, const int* sys_off, const int* sys_n,
                          int n_used, int n_pad)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_pad) return;
    if (b < n_used) {
        const int sy = blk_sys[b / TB];
        if (b - sys_off[sy] < sys_n[sy]) return;
    }
    s.f[PX][b] = 1e12 * (double)(b + 1);
    s.f[PY][b] = 0;
    s.f[PZ][b] = 0;
    s.f[RAD][b] = 0;
    s.f[MASS][b] = 1.0;
    s.f[MOI][b] = 1.0;
    aux.q[b] = 0;
    aux.g[b] = 0;
    aux.q32[b] = 0.f;
    aux.g32[b] = 0.f;
}

__global__ void k_aux(State s, FastAux aux, int n, double sqrtC, double sqrtG, double L0inv, double M0inv,
                      double wq, double wg)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const double R = s.f[RAD][b], m = s.f[MASS][b];
    aux.q[b] = sqrtC * R * R * R;
    aux.g[b] = sqrtG * m;
    const double r = R * L0inv;
    aux.q32[b] = (float)(wq * (r * r * r));
    aux.g32[b] = (float)(wg * (m * M0inv));
}

// Per 64-ball block (one wavefront each): the largest gravity-to-cohesion weight ratio
// g/q of its balls, the mixed path's bound on where gravity can matter in a far tile.
__global__ void k_gq_blk(FastAux aux, const int* blk_sys, const int* sys_off, const int* sys_n, float* gq)
{
    const int l = threadIdx.x;
    const int b = blockIdx.x * TB + l;
    const int sy = blk_sys[blockIdx.x];
    const bool valid = b - sys_off[sy] < sys_n[sy];
    const float q = aux.q32[b], g = aux.g32[b];
    // a ball without cohesion (q = 0) but with gravity has no bound: +inf keeps gravity on
    float r = !valid ? 0.f : (q > 0.f ? g / q : (g > 0.f ? __builtin_huge_valf() : 0.f));
    for (int m = 1; m < TB; m <<= 1) r = fmaxf(r, __shfl_xor(r, m, TB));
    if (l == 0) gq[blockIdx.x] = r;
}

// dst[k] = src[perm[k]] (to engine order) for fields [f0, f1)
__global__ void k_gather(State dst, State src, const int* perm, int n, int f0, int f1)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const int c = perm[b];
    for (int f = f0; f < f1; f++) dst.f[f][b] = src.f[f][c];
}

// dst[perm[k]] = src[k] (to canonical order) for fields [f0, f1)
__global__ void k_scatter(State dst, State src, const int* perm, int n, int f0, int f1)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const int c = perm[b];
    for (int f = f0; f < f1; f++) dst.f[f][c] = src.f[f][b];
}

// Lane-layout probe for the DPP rotation the tile kernels rely on.
__global__ void k_probe_rot(int* out)
{
    const int l = threadIdx.x;
    out[l] = __float_as_int(rot1(__int_as_float(l)));
    out[64 + l] = (int)rot1((double)l);
    out[128 + l] = (int)rotk((double)l, 31);
}

constexpr int GRAPH_MAX_STEPS = 128;   // steps per captured graph

int err_hip(std::string* err, const char* what)
{
    if (err) *err = std::string("fast path: ") + what;
    return SL_E_HIP;
}

// One workgroup per TPB tiles, or (queue) as many workgroups as the device holds at once
// pulling tiles from the work queue.  Both give the same bits.
template <int PREC, bool WRITE, bool GRAV, int SPLIT>
void launch_far(const TileArgs& a, bool queue, int n_cu, hipStream_t st)
{
    constexpr int TPB = NW / SPLIT;
    const int blocks = (a.n_tiles + TPB - 1) / TPB;
    if (!queue) {
        k_tiles<PREC, WRITE, GRAV, SPLIT><<<blocks, NT, 0, st>>>(a);
        return;
    }
    // workgroups one CU holds at once, from the kernel's registers and LDS (512 VGPRs per
    // SIMD lane in granules of 8, at most 8 wavefronts per SIMD, 4 SIMDs and 160 KB of
    // LDS per CU)
    static const int per_cu = [] {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_tiles_q<PREC, WRITE, GRAV, SPLIT>)) !=
            hipSuccess)
            return 1;
        const int vg = std::max(8, (fa.numRegs + 7) / 8 * 8);
        const int waves_per_simd = std::min(8, 512 / vg);
        const int by_regs = waves_per_simd * 4 / (NT / 64);
        const int by_lds = fa.sharedSizeBytes > 0 ? (int)(163840 / fa.sharedSizeBytes) : 64;
        return std::max(1, std::min(by_regs, by_lds));
    }();
    const int grid = std::max(1, std::min(blocks, per_cu * n_cu));
    k_tiles_q<PREC, WRITE, GRAV, SPLIT><<<grid, NT, 0, st>>>(a);
}

template <int SPLIT>
void launch_far_split(const TileArgs& a, int precision, bool write, bool grav, bool q, int n_cu, hipStream_t st)
{
    if (precision == SL_PREC_MIXED) {
        if (write)
            grav ? launch_far<SL_PREC_MIXED, true, true, SPLIT>(a, q, n_cu, st)
                 : launch_far<SL_PREC_MIXED, true, false, SPLIT>(a, q, n_cu, st);
        else
            grav ? launch_far<SL_PREC_MIXED, false, true, SPLIT>(a, q, n_cu, st)
                 : launch_far<SL_PREC_MIXED, false, false, SPLIT>(a, q, n_cu, st);
    } else {
        if (write)
            grav ? launch_far<SL_PREC_FP64, true, true, SPLIT>(a, q, n_cu, st)
                 : launch_far<SL_PREC_FP64, true, false, SPLIT>(a, q, n_cu, st);
        else
            grav ? launch_far<SL_PREC_FP64, false, true, SPLIT>(a, q, n_cu, st)
                 : launch_far<SL_PREC_FP64, false, false, SPLIT>(a, q, n_cu, st);
    }
}

// choice: the tile split (1, 2, 4), plus 8 for the work-queue launch
void launch_far_any(const TileArgs& a, int choice, int precision, bool write, bool grav, int n_cu, hipStream_t st)
{
    const int split = choice & 7;
    const bool q = (choice & 8) != 0;
    if (split == 4)
        launch_far_split<4>(a, precision, write, grav, q, n_cu, st);
    else if (split == 2)
        launch_far_split<2>(a, precision, write, grav, q, n_cu, st);
    else
        launch_far_split<1>(a, precision, write, grav, q, n_cu, st);
}

TileArgs tile_args(FastBuffers& fb)
{
    const State& c = fb.eng[fb.cur];
    TileArgs a;
    a.px = c.f[PX];
    a.py = c.f[PY];
    a.pz = c.f[PZ];
    a.R = c.f[RAD];
    a.aux = fb.aux;
    a.xl = fb.xl;
    a.gq_blk = fb.gq_blk;
    a.sysP = fb.sysP;
    a.bnd = fb.bnd;
    a.n_pad = fb.n_pad;
    a.n_tiles = fb.n_tiles;
    a.tiles = fb.tiles;
    a.L0inv = (float)(1.0 / fb.L0);
    a.L0inv2 = (float)(1.0 / (fb.L0 * fb.L0));
    a.Kf = fb.Kf;
    a.part = fb.part;
    a.nbm = fb.nbm;
    a.pe_tile = fb.pe_tile;
    a.queue = fb.queue;
    return a;
}

bool gravity_on(const FastBuffers& fb) { return fb.P[0].far_gravity != 0; }

int split_key(const FastBuffers& fb) { return fb.precision * 2 + (gravity_on(fb) ? 1 : 0); }

// The tile split has been chosen for the current precision and gravity setting.
bool split_ready(const FastBuffers& fb) { return fb.split != 0 && fb.split_key == split_key(fb); }

// Autotuner: the number of wavefronts sharing one far tile trades per-wave set-up
// against occupancy and load balance, which depend on the systems' sizes and the
// machine, and the work-queue launch trades a per-tile atomic for the launch's ramp and
// tail.  A choice stored for this configuration (sl_tune.cpp) is used as it is;
// otherwise each choice is warmed up once and then timed over three rounds of five
// launches on the live state (HIP events on the engine's stream), and the one kept is
// stored.  The trial launches only write partial sums that the real launch overwrites.
// Returns the split (1, 2, 4) plus 8 for the queue; a forced split takes the grid launch.
int tune_split(FastBuffers& fb, const TileArgs& a, hipStream_t st)
{
    if (fb.forced_split == 1 || fb.forced_split == 2 || fb.forced_split == 4) return fb.forced_split;
    const std::string key = tune_key(fb.precision, gravity_on(fb) ? 1 : 0, fb.lay.n_sys, fb.lay.n_used,
                                     fb.lay.nb_max, fb.n_tiles);
    if (const int cached = tune_cache_get(key)) {
        const int sp = cached & 7;
        if (sp == 1 || sp == 2 || sp == 4) return cached & 15;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 1;
    }
    const bool grav = gravity_on(fb);
    constexpr int NC = 6;
    const int choices[NC] = {1, 2, 4, 9, 10, 12};
    float t_min[NC];                            // best of three rounds per choice
    for (int k = 0; k < NC; k++) t_min[k] = 1e30f;
    for (int round = 0; round < 3; round++)
        for (int k = 0; k < NC; k++) {
            const int c = choices[k];
            launch_far_any(a, c, fb.precision, false, grav, fb.n_cu, st);
            (void)hipEventRecord(e0, st);
            for (int rep = 0; rep < 5; rep++) launch_far_any(a, c, fb.precision, false, grav, fb.n_cu, st);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < t_min[k]) t_min[k] = ms;
        }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // The split: the smallest within 2% of the fastest (splits that time alike then give
    // the same choice, and the same sums, from run to run more often).  The launch: the
    // queue only where it beats the grid launch at that split by more than 2% (both give
    // the same bits).
    float best_of_split[3];
    for (int k = 0; k < 3; k++) best_of_split[k] = std::min(t_min[k], t_min[k + 3]);
    const float fastest = std::min(best_of_split[0], std::min(best_of_split[1], best_of_split[2]));
    int ks = 0;
    for (int k = 0; k < 3; k++)
        if (best_of_split[k] <= 1.02f * fastest) {
            ks = k;
            break;
        }
    const int pick = t_min[ks + 3] < t_min[ks] / 1.02f ? choices[ks + 3] : choices[ks];
    tune_cache_put(key, pick, t_min, NC);
    return pick;
}

int launch_tiles_step(FastBuffers& fb, bool write, hipStream_t st, std::string* err)
{
    const TileArgs a = tile_args(fb);
    if (!split_ready(fb)) {
        const int choice = tune_split(fb, a, st);
        fb.split = choice & 7;
        fb.queue_launch = fb.forced_queue ? fb.forced_queue == 2 : (choice & 8) != 0;
        fb.split_key = split_key(fb);
        fast_clear_graphs(fb);
    }
    if (fb.hook.before) fb.hook.before(fb.hook.user, 0);
    launch_far_any(a, fb.split | (fb.queue_launch ? 8 : 0), fb.precision, write, gravity_on(fb), fb.n_cu, st);
    if (fb.hook.after) fb.hook.after(fb.hook.user, 0);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "tile kernel launch");
    return 0;
}

int launch_ball_step(FastBuffers& fb, bool write, bool kick, bool drift_next, hipStream_t st, std::string* err)
{
    BallArgs a;
    a.cur = fb.eng[fb.cur];
    a.nxt = fb.eng[1 - fb.cur];
    a.sysP = fb.sysP;
    a.blk_sys = fb.blk_sys;
    a.sys_off = fb.sys_off;
    a.sys_n = fb.sys_n;
    a.sys_nb = fb.sys_nb;
    a.n_used = fb.lay.n_used;
    a.n_pad = fb.n_pad;
    a.part = fb.part;
    a.Kf = fb.Kf;
    a.nbm = fb.nbm;
    a.perm = fb.perm;
    a.kick = kick;
    a.drift_next = drift_next;
    a.bnd = fb.bnd;
    a.xl = fb.xl;
    a.L0inv = 1.0 / fb.L0;
    a.e_blk = fb.e_blk;
    a.blk_ctr = fb.blk_ctr;
    a.flags = fb.flags;
    const int blocks = fb.lay.n_used / TB;
    if (fb.hook.before) fb.hook.before(fb.hook.user, 1);
    const bool mixed = fb.precision == SL_PREC_MIXED;
    if (fast_ball_waves(fb) == 8) {
        constexpr int T = kb_threads(8);
        if (mixed) write ? k_ball<true, true, 8><<<blocks, T, 0, st>>>(a) : k_ball<true, false, 8><<<blocks, T, 0, st>>>(a);
        else write ? k_ball<false, true, 8><<<blocks, T, 0, st>>>(a) : k_ball<false, false, 8><<<blocks, T, 0, st>>>(a);
    } else {
        constexpr int T = kb_threads(4);
        if (mixed) write ? k_ball<true, true, 4><<<blocks, T, 0, st>>>(a) : k_ball<true, false, 4><<<blocks, T, 0, st>>>(a);
        else write ? k_ball<false, true, 4><<<blocks, T, 0, st>>>(a) : k_ball<false, false, 4><<<blocks, T, 0, st>>>(a);
    }
    if (fb.hook.after) fb.hook.after(fb.hook.user, 1);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "ball kernel launch");
    fb.cur = 1 - fb.cur;
    fb.drifted = drift_next;
    return 0;
}

int launch_bounds(FastBuffers& fb, hipStream_t st, std::string* err)
{
    k_bounds<<<fb.lay.n_used / TB, TB, 0, st>>>(fb.eng[fb.cur], fb.blk_sys, fb.sys_off, fb.sys_n, fb.bnd, fb.xl,
                                                1.0 / fb.L0);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "block bounds");
    return 0;
}

int launch_drift_sys(FastBuffers& fb, hipStream_t st, std::string* err)
{
    k_drift_sys<<<(fb.lay.n_used + 255) / 256, 256, 0, st>>>(fb.eng[fb.cur], fb.blk_sys, fb.sys_off, fb.sys_n,
                                                             fb.sysP, fb.lay.n_used);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "drift");
    return launch_bounds(fb, st, err);
}

// Read back the partials of a write step and sum them per system, in a fixed order:
// the far-field energy of the tiles in table order, then the near-pair energy,
// kinetic energy, momenta and contacts of the 64-ball blocks in ascending order.
int read_energies(FastBuffers& fb, sl_energy* out, hipStream_t st, std::string* err)
{
    const int nblk = fb.lay.n_used / BPB;
    if (hipMemcpyAsync(fb.h_pe_tile, fb.pe_tile, sizeof(double) * fb.n_tiles, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(fb.h_e_blk, fb.e_blk, sizeof(double) * FAST_E_STRIDE * nblk, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return err_hip(err, "energy read-back");
    const int ns = fb.lay.n_sys;
    std::vector<sl_energy> e(ns);
    for (auto& x : e) x = sl_energy{};
    for (int t = 0; t < fb.n_tiles; t++) e[fb.tile_sys_h[t]].PE += fb.h_pe_tile[t];
    for (int b = 0; b < nblk; b++) {
        const int s = fb.blk_sys_h[(b * BPB) / TB];
        const double* v = fb.h_e_blk + (size_t)b * FAST_E_STRIDE;
        sl_energy& x = e[s];
        x.PE += v[0];
        x.contacts += (long long)v[8];
        x.KE += v[1];
        x.p[0] += v[2];
        x.p[1] += v[3];
        x.p[2] += v[4];
        x.L[0] += v[5];
        x.L[1] += v[6];
        x.L[2] += v[7];
        if (v[9] > 0) x.nonfinite = 1;
    }
    if (out)
        for (int s = 0; s < ns; s++) out[s] = e[s];
    return 0;
}

// One step launched directly.
int direct_step(FastBuffers& fb, bool write, bool drift_next, hipStream_t st, std::string* err)
{
    int rc;
    if (!fb.drifted && (rc = launch_drift_sys(fb, st, err))) return rc;
    fb.drifted = false;
    if ((rc = launch_tiles_step(fb, write, st, err))) return rc;
    return launch_ball_step(fb, write, true, drift_next, st, err);
}

// A run of `len` steps that starts drifted and ends drifted (the last may be a
// write step), captured once per (length, write, state parity) and replayed.
int graph_segment(FastBuffers& fb, int len, bool write_last, hipStream_t st, std::string* err)
{
    const FastBuffers::GraphKey key{len, write_last ? 1 : 0, fb.cur};
    auto it = fb.graphs.find(key);
    if (it == fb.graphs.end()) {
        if (fb.graphs.size() >= 16) {
            // earlier segments may still be running: destroy their graphs once idle
            if (hipStreamSynchronize(st) != hipSuccess) return err_hip(err, "synchronize");
            fast_clear_graphs(fb);
        }
        const int cur0 = fb.cur;
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return err_hip(err, "begin graph capture");
        int rc = 0;
        for (int k = 0; k < len && !rc; k++) {
            const bool w = write_last && k == len - 1;
            rc = launch_tiles_step(fb, w, st, err);
            if (!rc) rc = launch_ball_step(fb, w, true, true, st, err);
        }
        const hipError_t ec = hipStreamEndCapture(st, &g);
        fb.cur = cur0;   // capture recorded the launches; nothing ran yet
        fb.drifted = true;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess || !g) return err_hip(err, "end graph capture");
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess || !ex) return err_hip(err, "instantiate graph");
        it = fb.graphs.emplace(key, ex).first;
    }
    if (hipGraphLaunch(it->second, st) != hipSuccess) return err_hip(err, "graph launch");
    if (len & 1) fb.cur = 1 - fb.cur;
    fb.drifted = true;
    return 0;
}

// Morton code of a point already scaled to [0, 2^21) per axis.
uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    return spread(x) | spread(y) << 1 | spread(z) << 2;
}

// Engine slot -> canonical slot: each system's balls sorted by the Morton code of
// their positions (ties by canonical index); padding slots map to themselves.
std::vector<int> morton_perm(const FastLayout& L, const std::vector<double>& x, const std::vector<double>& y,
                             const std::vector<double>& z, int n_pad)
{
    std::vector<int> perm(n_pad);
    std::iota(perm.begin(), perm.end(), 0);
    for (int s = 0; s < L.n_sys; s++) {
        const int o = L.off[s], n = L.n[s];
        double lo[3] = {x[o], y[o], z[o]}, hi[3] = {x[o], y[o], z[o]};
        for (int k = 0; k < n; k++) {
            const double p[3] = {x[o + k], y[o + k], z[o + k]};
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], p[d]);
                hi[d] = std::max(hi[d], p[d]);
            }
        }
        double ext = 0;
        for (int d = 0; d < 3; d++) ext = std::max(ext, hi[d] - lo[d]);
        const double sc = ext > 0 ? 2097151.0 / ext : 0.0;
        std::vector<std::pair<uint64_t, int>> key(n);
        for (int k = 0; k < n; k++) {
            const uint32_t q0 = (uint32_t)((x[o + k] - lo[0]) * sc);
            const uint32_t q1 = (uint32_t)((y[o + k] - lo[1]) * sc);
            const uint32_t q2 = (uint32_t)((z[o + k] - lo[2]) * sc);
            key[k] = {morton3(q0, q1, q2), o + k};
        }
        std::sort(key.begin(), key.end());
        for (int k = 0; k < n; k++) perm[o + k] = key[k].second;
    }
    return perm;
}

// Refresh the engine's current set from the canonical set (dynamic fields).
int load_from_canon(FastBuffers& fb, State& canon, hipStream_t st, std::string* err)
{
    k_gather<<<(fb.lay.n_used + 255) / 256, 256, 0, st>>>(fb.eng[fb.cur], canon, fb.perm, fb.lay.n_used, PX, RAD);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "gather state");
    k_pad_sys<<<(fb.n_pad + 255) / 256, 256, 0, st>>>(fb.eng[fb.cur], fb.aux, fb.blk_sys, fb.sys_off, fb.sys_n,
                                                      fb.lay.n_used, fb.n_pad);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "padding");
    fb.eng_valid = true;
    fb.drifted = false;
    return 0;
}

int ensure_engine(FastBuffers& fb, State& canon, hipStream_t st, std::string* err)
{
    if (fb.eng_valid) return 0;
    return load_from_canon(fb, canon, st, err);
}

}  // namespace

// A profiling tool loaded through rocprofiler-sdk (rocprofv3) is named in
// ROCP_TOOL_LIBRARIES of the process it profiles.  Runs under rocprofv3 --kernel-trace
// that replayed captured step graphs ended three times in host-side SIGSEGVs inside the
// runtime, below hipGraphLaunch, while the same runs without the tool pass; graph
// executables here are only destroyed on an idle stream and re-captured after every
// reconfiguration.  With a tool attached the engines therefore launch every kernel
// directly (the kernels and their order are the same).
bool tool_attached()
{
    static const bool on = [] {
        const char* v = std::getenv("ROCP_TOOL_LIBRARIES");
        return v != nullptr && *v != 0;
    }();
    return on;
}

// Four wavefronts (three workgroups per CU) when there are more 64-ball blocks than CUs
// and every system has at most 80 blocks: many small systems, as in a collision
// ensemble, where the block count, not the slots per ball, sets the time.  Otherwise
// eight: a lone system, or large ones, whose balls have many slots to load and more
// partner masks to keep.  The results are the same bit for bit.
int fast_ball_waves(const FastBuffers& fb)
{
    if (fb.forced_ball_waves == 4 || fb.forced_ball_waves == 8) return fb.forced_ball_waves;
    return fb.lay.n_used / TB > fb.n_cu && fb.lay.nb_max <= 80 ? 4 : 8;
}

FastLayout fast_layout(int n_sys, const int* n_per_sys)
{
    FastLayout L;
    L.n_sys = n_sys;
    int off = 0;
    for (int s = 0; s < n_sys; s++) {
        const int n = n_per_sys[s];
        const int nb = (n + TB - 1) / TB;
        L.n.push_back(n);
        L.off.push_back(off);
        L.nb.push_back(nb);
        L.nb_max = std::max(L.nb_max, nb);
        off += nb * TB;
    }
    L.n_used = off;
    return L;
}

int fast_max_tile_blocks(int n_pad)
{
    const long long nb = (n_pad + TB - 1) / TB;
    return (int)std::min<long long>(nb * (nb + 1) / 2, 1 << 30);
}

int fast_alloc(FastBuffers& fb, int n_pad, int max_sys)
{
    fb.n_pad = n_pad;
    fb.max_sys = max_sys;
    const int nblocks = (n_pad + TB - 1) / TB;
    if (hipMalloc(&fb.eng_base, sizeof(double) * 2 * (size_t)NFIELD * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.perm, sizeof(int) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.q, sizeof(double) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.g, sizeof(double) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.q32, sizeof(float) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.aux.g32, sizeof(float) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.bnd, sizeof(double4) * 2 * nblocks) != hipSuccess) return -1;
    if (hipMalloc(&fb.xl, sizeof(float4) * n_pad) != hipSuccess) return -1;
    if (hipMalloc(&fb.gq_blk, sizeof(float) * nblocks) != hipSuccess) return -1;
    if (hipMalloc(&fb.queue, sizeof(int) * QPAD * (QSEG + 1)) != hipSuccess) return -1;
    if (hipMemset(fb.queue, 0, sizeof(int) * QPAD * (QSEG + 1)) != hipSuccess) return -1;
    if (hipMalloc(&fb.blk_sys, sizeof(int) * nblocks) != hipSuccess) return -1;
    if (hipMalloc(&fb.sys_off, sizeof(int) * max_sys) != hipSuccess) return -1;
    if (hipMalloc(&fb.sys_n, sizeof(int) * max_sys) != hipSuccess) return -1;
    if (hipMalloc(&fb.sys_nb, sizeof(int) * max_sys) != hipSuccess) return -1;
    if (hipMalloc(&fb.sysP, sizeof(Phys) * max_sys) != hipSuccess) return -1;
    if (hipMalloc(&fb.flags, sizeof(int) * 4) != hipSuccess) return -1;
    const int eblocks = (n_pad + BPB - 1) / BPB;
    if (hipMalloc(&fb.e_blk, sizeof(double) * FAST_E_STRIDE * eblocks) != hipSuccess) return -1;
    if (hipMalloc(&fb.blk_ctr, sizeof(unsigned long long) * 2 * eblocks) != hipSuccess) return -1;
    fb.cap_blocks = eblocks;
    if (hipHostMalloc(&fb.h_e_blk, sizeof(double) * FAST_E_STRIDE * eblocks, 0) != hipSuccess) return -1;
    if (hipMemset(fb.eng_base, 0, sizeof(double) * 2 * (size_t)NFIELD * n_pad) != hipSuccess) return -1;
    if (hipMemset(fb.flags, 0, sizeof(int) * 4) != hipSuccess) return -1;
    for (int k = 0; k < 2; k++)
        for (int f = 0; f < NFIELD; f++) fb.eng[k].f[f] = fb.eng_base + ((size_t)k * NFIELD + f) * n_pad;
    // radius, mass and moment of inertia are shared by both engine sets
    for (int f = RAD; f < NFIELD; f++) fb.eng[1].f[f] = fb.eng[0].f[f];
    return 0;
}

void fast_clear_graphs(FastBuffers& fb)
{
    for (auto& kv : fb.graphs) (void)hipGraphExecDestroy(kv.second);
    fb.graphs.clear();
}

void fast_free(FastBuffers& fb)
{
    fast_clear_graphs(fb);
    void* ptrs[] = {fb.eng_base, fb.perm, fb.aux.q, fb.aux.g, fb.aux.q32, fb.aux.g32, fb.bnd, fb.xl, fb.gq_blk, fb.queue, fb.blk_sys,
                    fb.sys_off, fb.sys_n, fb.sys_nb, fb.sysP, fb.tiles, fb.part, fb.nbm, fb.flags,
                    fb.blk_ctr, fb.e_blk, fb.pe_tile};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (fb.h_e_blk) (void)hipHostFree(fb.h_e_blk);
    if (fb.h_pe_tile) (void)hipHostFree(fb.h_pe_tile);
    fb = FastBuffers{};
}

int fast_set_phys(FastBuffers& fb, State& canon, const Phys* P, int precision, hipStream_t st, std::string* err)
{
    const FastLayout& L = fb.lay;
    fb.P.assign(P, P + L.n_sys);
    fb.precision = precision;
    if (hipMemcpyAsync(fb.sysP, fb.P.data(), sizeof(Phys) * L.n_sys, hipMemcpyHostToDevice, st) != hipSuccess)
        return err_hip(err, "physics table");
    // units of the mixed path: median radius and mass of the balls in use
    std::vector<double> R(L.n_used), m(L.n_used);
    if (hipMemcpyAsync(R.data(), canon.f[RAD], sizeof(double) * L.n_used, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(m.data(), canon.f[MASS], sizeof(double) * L.n_used, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return err_hip(err, "read radii");
    std::vector<double> Rv, mv;
    for (int s = 0; s < L.n_sys; s++)
        for (int b = 0; b < L.n[s]; b++) {
            Rv.push_back(R[L.off[s] + b]);
            mv.push_back(m[L.off[s] + b]);
        }
    const size_t h = Rv.size() / 2;
    std::nth_element(Rv.begin(), Rv.begin() + h, Rv.end());
    std::nth_element(mv.begin(), mv.begin() + h, mv.end());
    fb.L0 = Rv[h];
    fb.M0 = mv[h];
    // mixed path: force = Kf (q_i q_j / (s t)^2 + g_i g_j / d^3) (r / L0) in units of L0, M0
    const double Kv = (P[0].Ha / 6 * 64) / fb.L0, Kg = P[0].G * fb.M0 * fb.M0 / (fb.L0 * fb.L0);
    fb.Kf = Kv > 0 ? Kv : (Kg > 0 ? Kg : 1.0);
    const double C = P[0].Ha / 6 * 64;
    k_aux<<<(L.n_used + 255) / 256, 256, 0, st>>>(fb.eng[0], fb.aux, L.n_used, std::sqrt(C), std::sqrt(P[0].G),
                                                  1.0 / fb.L0, 1.0 / fb.M0, std::sqrt(Kv / fb.Kf),
                                                  std::sqrt(Kg / fb.Kf));
    for (int k = 0; k < 2; k++)
        k_pad_sys<<<(fb.n_pad + 255) / 256, 256, 0, st>>>(fb.eng[k], fb.aux, fb.blk_sys, fb.sys_off, fb.sys_n,
                                                          L.n_used, fb.n_pad);
    if (L.n_used > 0)
        k_gq_blk<<<L.n_used / TB, TB, 0, st>>>(fb.aux, fb.blk_sys, fb.sys_off, fb.sys_n, fb.gq_blk);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "per-ball constants");
    fast_clear_graphs(fb);
    return 0;
}

int fast_configure(FastBuffers& fb, State& canon, const FastLayout& lay, const Phys* P, int precision,
                   hipStream_t st, std::string* err)
{
    if (lay.n_used > fb.n_pad || lay.n_sys > fb.max_sys || lay.n_sys < 1) {
        if (err) *err = "fast path: layout exceeds the allocated capacity";
        return SL_E_ARG;
    }
    const bool same = fb.lay.n_sys == lay.n_sys && fb.lay.n == lay.n;
    // captured graphs hold the buffer addresses and tile counts of the old layout: drop them
    // (the stream is idle here) and the tuned split before any buffer is replaced
    fast_clear_graphs(fb);
    fb.lay = lay;
    fb.cur = 0;
    fb.drifted = false;
    if (!same) {
        fb.split = 0;
        fb.n_tiles = 0;
        const int nblocks = lay.n_used / TB;
        fb.blk_sys_h.assign(std::max(nblocks, 1), 0);
        std::vector<int4> tiles;
        fb.tile_sys_h.clear();
        // The diagonal tiles of every system first (half a sweep each, and most of the
        // per-pair near tests), then the others by band, I - J = 1, 2, ...: blocks close in
        // the space-filling-curve order are close in space, so the tiles that need the
        // per-pair near test come early and the launch ends on far tiles.  (Diagonal first
        // was measured 0.7% faster than row order.)  The order only schedules the tiles;
        // each tile's sums and the slots they land in do not depend on it.  Each system's
        // tiles keep the same relative order whether it runs alone or in a batch.
        for (int s = 0; s < lay.n_sys; s++) {
            const int b0 = lay.off[s] / TB;
            for (int k = 0; k < lay.nb[s]; k++) fb.blk_sys_h[b0 + k] = s;
        }
        int nb_top = 0;
        for (int s = 0; s < lay.n_sys; s++) nb_top = std::max(nb_top, lay.nb[s]);
        for (int d = 0; d < nb_top; d++)
            for (int s = 0; s < lay.n_sys; s++) {
                const int b0 = lay.off[s] / TB;
                for (int I = d; I < lay.nb[s]; I++) {
                    tiles.push_back(make_int4(b0 + I, b0 + I - d, s, b0));
                    fb.tile_sys_h.push_back(s);
                }
            }
        const int nt = (int)tiles.size();
        if (nt > fb.cap_tiles) {
            if (fb.tiles) (void)hipFree(fb.tiles);
            if (fb.pe_tile) (void)hipFree(fb.pe_tile);
            if (fb.h_pe_tile) (void)hipHostFree(fb.h_pe_tile);
            fb.tiles = nullptr;
            fb.pe_tile = nullptr;
            fb.h_pe_tile = nullptr;
            fb.cap_tiles = 0;
            if (hipMalloc(&fb.tiles, sizeof(int4) * nt) != hipSuccess ||
                hipMalloc(&fb.pe_tile, sizeof(double) * nt) != hipSuccess ||
                hipHostMalloc(&fb.h_pe_tile, sizeof(double) * nt, 0) != hipSuccess)
                return err_hip(err, "tile table allocation");
            fb.cap_tiles = nt;
        }
        if (lay.nb_max > fb.part_slots) {
            if (fb.part) (void)hipFree(fb.part);
            if (fb.nbm) (void)hipFree(fb.nbm);
            fb.part = nullptr;
            fb.nbm = nullptr;
            fb.part_slots = 0;
            if (hipMalloc(&fb.part, sizeof(double) * (size_t)lay.nb_max * 3 * fb.n_pad) != hipSuccess ||
                hipMalloc(&fb.nbm, sizeof(unsigned long long) * (size_t)lay.nb_max * fb.n_pad) != hipSuccess)
                return err_hip(err, "partial buffer allocation");
            fb.part_slots = lay.nb_max;
        }
        if (hipMemcpyAsync(fb.tiles, tiles.data(), sizeof(int4) * nt, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fb.blk_sys, fb.blk_sys_h.data(), sizeof(int) * fb.blk_sys_h.size(), hipMemcpyHostToDevice,
                           st) != hipSuccess ||
            hipMemcpyAsync(fb.sys_off, lay.off.data(), sizeof(int) * lay.n_sys, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fb.sys_n, lay.n.data(), sizeof(int) * lay.n_sys, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fb.sys_nb, lay.nb.data(), sizeof(int) * lay.n_sys, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return err_hip(err, "layout tables");
        fb.n_tiles = nt;
        fb.split = 0;   // re-tune for the new sizes
    }
    if (launch_zero_u64(fb.blk_ctr, 2 * fb.cap_blocks, st) != hipSuccess ||
        launch_zero_i32(fb.flags, 4, st) != hipSuccess)
        return err_hip(err, "clear counters");
    // Morton order of the current positions, per system
    const int nu = lay.n_used;
    std::vector<double> x(nu), y(nu), z(nu);
    if (hipMemcpyAsync(x.data(), canon.f[PX], sizeof(double) * nu, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(y.data(), canon.f[PY], sizeof(double) * nu, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(z.data(), canon.f[PZ], sizeof(double) * nu, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return err_hip(err, "read positions");
    const std::vector<int> perm = morton_perm(lay, x, y, z, fb.n_pad);
    if (hipMemcpyAsync(fb.perm, perm.data(), sizeof(int) * fb.n_pad, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return err_hip(err, "permutation");
    k_gather<<<(nu + 255) / 256, 256, 0, st>>>(fb.eng[0], canon, fb.perm, nu, RAD, NFIELD);
    if (hipGetLastError() != hipSuccess) return err_hip(err, "gather constants");
    int rc = fast_set_phys(fb, canon, P, precision, st, err);
    if (rc) return rc;
    rc = load_from_canon(fb, canon, st, err);
    if (rc) return rc;
    fb.canon_valid = true;
    return 0;
}

int fast_sync_to(FastBuffers& fb, State& canon, hipStream_t st, std::string* err)
{
    if (fb.canon_valid) return 0;
    k_scatter<<<(fb.lay.n_used + 255) / 256, 256, 0, st>>>(canon, fb.eng[fb.cur], fb.perm, fb.lay.n_used, PX, RAD);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return err_hip(err, "copy state");
    fb.canon_valid = true;
    return 0;
}

void fast_canon_changed(FastBuffers& fb)
{
    fb.eng_valid = false;
    fb.canon_valid = true;
}

int fast_init(FastBuffers& fb, State& canon, sl_energy* e_sys, hipStream_t st, std::string* err)
{
    int rc = ensure_engine(fb, canon, st, err);
    if (rc) return rc;
    if (launch_copy_prev(fb.eng[fb.cur], fb.lay.n_used, st) != hipSuccess) return err_hip(err, "copy positions");
    fb.drifted = false;
    fb.canon_valid = false;
    rc = launch_bounds(fb, st, err);
    if (!rc) rc = launch_tiles_step(fb, true, st, err);
    if (!rc) rc = launch_ball_step(fb, true, false, false, st, err);
    if (!rc) rc = read_energies(fb, e_sys, st, err);
    return rc;
}

int fast_steps(FastBuffers& fb, State& canon, int n_steps, long long* step_count, int write_every,
               sl_energy* e_writes, int cap, hipStream_t st, std::string* err)
{
    int rc = ensure_engine(fb, canon, st, err);
    if (rc) return rc;
    if (n_steps > 0) fb.canon_valid = false;
    const int ns = fb.lay.n_sys;
    const bool graphs = fb.use_graphs && !fb.hook.before && !tool_attached();
    int nw = 0;
    int k = 0;
    auto record = [&](void) -> int {
        sl_energy* dst = (e_writes && nw < cap) ? e_writes + (size_t)nw * ns : nullptr;
        const int r = read_energies(fb, dst, st, err);
        nw++;
        return r;
    };
    while (k < n_steps) {
        const long long s = *step_count + 1;
        const bool write = write_every > 0 && (s % write_every) == 0;
        // steps until (and including) the next write step, within this call
        int until = n_steps - k;
        if (write_every > 0) {
            const long long next = ((s + write_every - 1) / write_every) * write_every;
            until = (int)std::min<long long>(until, next - s + 1);
        }
        // graph segments need a drifted start and must not include the call's last step;
        // long runs are cut into segments of at most GRAPH_MAX_STEPS, replayed repeatedly
        const int seg = std::min(std::min(until, n_steps - 1 - k), GRAPH_MAX_STEPS);
        if (graphs && fb.drifted && split_ready(fb) && seg >= 2) {
            const bool write_last = write_every > 0 && ((s + seg - 1) % write_every) == 0;
            rc = graph_segment(fb, seg, write_last, st, err);
            if (rc) return rc;
            *step_count += seg;
            k += seg;
            if (write_last && (rc = record())) return rc;
            continue;
        }
        rc = direct_step(fb, write, k + 1 < n_steps, st, err);
        if (rc) return rc;
        *step_count = s;
        k++;
        if (write && (rc = record())) return rc;
    }
    if (hipStreamSynchronize(st) != hipSuccess) return err_hip(err, "synchronize");
    return nw;
}

hipError_t fast_reset_counters(FastBuffers& fb, hipStream_t st)
{
    if (!fb.blk_ctr) return hipSuccess;
    return launch_zero_u64(fb.blk_ctr, 2 * fb.cap_blocks, st);
}

int fast_read_counters(FastBuffers& fb, hipStream_t st, long long* contacts, long long* near_pairs)
{
    // running sums of the 64-ball blocks in use (each pair is counted by its ball A)
    std::vector<unsigned long long> h(2 * (size_t)(fb.lay.n_used / BPB));
    if (h.empty()) {
        if (contacts) *contacts = 0;
        if (near_pairs) *near_pairs = 0;
        return 0;
    }
    if (hipMemcpyAsync(h.data(), fb.blk_ctr, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return SL_E_HIP;
    long long cc = 0, cn = 0;
    for (size_t b = 0; b < h.size(); b += 2) {
        cc += (long long)h[b];
        cn += (long long)h[b + 1];
    }
    if (contacts) *contacts = cc;
    if (near_pairs) *near_pairs = cn;
    return 0;
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
    (void)hipFree(d);
    return rc;
}
