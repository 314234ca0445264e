// Device kernels of the fast path (included by sl_kernels_fast.hip).
//
//   k_tiles_*  SPLIT wavefronts per 64x64 tile (I, J) of a system's pair triangle,
//              I >= J, each sweeping 64/SPLIT of the tile's diagonals.  Lane l
//              holds ball i = 64 I + l in registers.  The 64 balls of block J are
//              staged in LDS twice (slots 0..63 and 64..127), so at sweep step s lane
//              l reads slot l + s -- ball j = (l + s) mod 64 -- at base + s*stride,
//              with no index arithmetic.  The reaction force on j is accumulated in
//              a register that travels with j: after every step the accumulators
//              rotate one lane down (DPP wave_rol:1), so lane l always holds the sum
//              of the j it meets next; no LDS atomics or cross-lane reductions.
//              A pair closer than S + h_min (touching, or in the band where the
//              reference clamps the vdW gap) contributes nothing to the sweep: it is
//              only recorded, as a bit of the partner masks of both balls.  The near
//              test's compare is already the wavefront's lane mask: a step with any
//              near lane stores that mask in LDS (a uniform branch, rarely taken), and
//              each lane's partner bits are rebuilt from the stored steps after the
//              sweep, so the sweep itself spends no vector instructions on them.  Each
//              tile writes its i-side and j-side force sums and partner masks once,
//              to slot J of the balls of block I and slot I of the balls of block J
//              (system-local block numbers), so every (ball, slot) is written by
//              exactly one tile per step and nothing needs clearing.
//   k_ball     one workgroup per 64-ball block, eight or four wavefronts with lane
//              l working for ball l: the far-field partials of every eighth slot
//              (one slot class) are summed by one wavefront (Kahan-compensated in
//              the mixed mode), which counts the ball's partner-mask bits; the block
//              numbers its near pairs ball by ball and evaluates them with the
//              reference arithmetic, one per thread, in rounds of one entry per
//              thread; then wavefront 0 adds the near results of
//              each ball in entry order, forms acc and aacc, applies the closing
//              kick, the write-step energies and the next step's drift (writing the
//              other state set).  A near pair is evaluated once for each of its two
//              balls, by the blocks that hold them.
// Every sum has a fixed order at a given tile split, so results are reproducible run
// to run; the k_ball shape does not change them.
#pragma once

#include "sl_fast_pair.hpp"
#include "sl_internal.hpp"
#include "sl_pair_ref.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace sl {
namespace fastk {

constexpr int TB = FAST_TILE;
constexpr int NW = 4;                 // wavefronts per block
#ifndef SL_TILE_UNROLL
#define SL_TILE_UNROLL 2              // pairs of the tile sweep unrolled per wavefront
#endif
constexpr int NT = TB * NW;

// Lane l receives the value held by lane (l + 1) mod 64 (DPP wave_rol:1).
__device__ __forceinline__ float rot1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xf, 0xf, false));
}
__device__ __forceinline__ double rot1(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x134, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x134, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// Lane l receives the value held by lane (l + k) mod 64 (ds_bpermute).
__device__ __forceinline__ float rotk(float v, int k)
{
    const int src = ((int)(threadIdx.x % TB) + k) & (TB - 1);
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(v)));
}
__device__ __forceinline__ double rotk(double v, int k)
{
    const int src = ((int)(threadIdx.x % TB) + k) & (TB - 1);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, (int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ unsigned long long rotl64(unsigned long long x, int k)
{
    return k ? (x << k) | (x >> (64 - k)) : x;
}

template <int N>
__device__ __forceinline__ double block_sum(double v, double* red)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = red[t] + red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

struct TileArgs {
    const double* px;
    const double* py;
    const double* pz;
    const double* R;
    FastAux aux;
    const Phys* sysP;       // per system
    const double4* bnd;     // per 64-ball block: bounding box of the centres (lo with max R, hi)
    const float4* xl;       // mixed: per ball, position relative to its block's box centre and R, in units of L0
    const float* gq_blk;    // mixed: per 64-ball block, the largest g/q of its balls (scaled weights)
    int n_pad;
    int n_tiles;
    const int4* tiles;      // {I, J, system, first block of the system}
    float L0inv, L0inv2;    // mixed: 1 / L0 and 1 / L0^2 (scaled units)
    double Kf;              // mixed: force = Kf * (scaled binary32 factor) * (r / L0)
    double* part;           // [slot][xyz][n_pad]; the mixed mode stores binary32 sums, Kf left out,
                            // in the first half of the same buffer ([slot][xyz][n_pad] floats)
    unsigned long long* nbm;   // [slot][n_pad]: near-partner masks
    double* pe_tile;        // per tile: far-field pair energy (write steps)
    int* queue;             // k_tiles_q: next group per segment, workgroups done (all 0 between launches)
};

// Sum over the 64 lanes of a wavefront as a fixed tree (deterministic).
__device__ __forceinline__ double wave_sum(double v)
{
    for (int m = 1; m < TB; m <<= 1) v = v + __shfl_xor(v, m, TB);
    return v;
}

__device__ __forceinline__ unsigned long long wave_or(unsigned long long v)
{
    for (int m = 1; m < TB; m <<= 1) v |= __shfl_xor(v, m, TB);
    return v;
}

__device__ __forceinline__ int wave_isum(int v)
{
    for (int m = 1; m < TB; m <<= 1) v += __shfl_xor(v, m, TB);
    return v;
}

// Bounding box of the valid balls held one per lane (every lane gets the result):
// lo = (min x, min y, min z, max R), hi = (max x, max y, max z, 0).
__device__ __forceinline__ void wave_aabb(double x, double y, double z, double R, bool valid, double4& lo,
                                          double4& hi)
{
    const double inf = __builtin_huge_val();
    double lx = valid ? x : inf, ly = valid ? y : inf, lz = valid ? z : inf;
    double hx = valid ? x : -inf, hy = valid ? y : -inf, hz = valid ? z : -inf;
    double rm = valid ? R : 0.0;
    for (int m = 1; m < TB; m <<= 1) {
        lx = fmin(lx, __shfl_xor(lx, m, TB));
        ly = fmin(ly, __shfl_xor(ly, m, TB));
        lz = fmin(lz, __shfl_xor(lz, m, TB));
        hx = fmax(hx, __shfl_xor(hx, m, TB));
        hy = fmax(hy, __shfl_xor(hy, m, TB));
        hz = fmax(hz, __shfl_xor(hz, m, TB));
        rm = fmax(rm, __shfl_xor(rm, m, TB));
    }
    lo = make_double4(lx, ly, lz, rm);
    hi = make_double4(hx, hy, hz, 0.0);
}

// A tile is far when the gap between its two blocks' bounding boxes exceeds the two
// largest radii plus h_min (with a 1e-9 relative margin): every pair in it is then
// farther apart than S + h_min by far more than the rounding of the per-pair test,
// so the sweep needs no near test.  Diagonal tiles always get the test.
__device__ __forceinline__ bool tile_is_far(const TileArgs& a, int I, int J, double h_min)
{
    if (I == J) return false;
    const double4 li = a.bnd[2 * I], ui = a.bnd[2 * I + 1], lj = a.bnd[2 * J], uj = a.bnd[2 * J + 1];
    const double gx = fmax(0.0, fmax(lj.x - ui.x, li.x - uj.x));
    const double gy = fmax(0.0, fmax(lj.y - ui.y, li.y - uj.y));
    const double gz = fmax(0.0, fmax(lj.z - ui.z, li.z - uj.z));
    const double lim = (li.w + lj.w + h_min) * (1 + 1e-9);
    return gx * gx + gy * gy + gz * gz > lim * lim;
}

// vdW energy of a far pair from s = d^2 - S^2, t = d^2 - (Ri-Rj)^2.
__device__ __forceinline__ double far_energy(double Ha, double Ri, double Rj, double s, double t)
{
    const double two = 2 * Ri * Rj;
    return -Ha / 6 * (two / s + two / t + ::log(s / t));
}

// Sweep range of part h of SPLIT: off-diagonal tiles sweep s = 0..63; diagonal
// tiles s = 1..32 (each unordered pair of the block once; s = 32 for the lower
// half of the lanes only).
template <int SPLIT>
__device__ __forceinline__ void sweep_range(bool diag, int h, int& s0, int& s1)
{
    if (!diag) {
        s0 = h * (TB / SPLIT);
        s1 = (h + 1) * (TB / SPLIT);
    } else {
        s0 = 1 + h * ((TB / 2) / SPLIT);
        s1 = 1 + (h + 1) * ((TB / 2) / SPLIT);
    }
}

__device__ __forceinline__ int4 fetch_tile(const TileArgs& a, int t, bool live)
{
    return live ? a.tiles[t] : make_int4(0, 0, 0, 0);
}

struct JRec64 {
    double x, y, z, R, q, g;
};
// The mixed path's record: the binary64 geometry, and the binary32 factors it uses as
// they are (radius in units of L0, cohesion and gravity weights), so no pair converts them.
struct JRecMix {
    double x, y, z, R;
    float Rs, q, g, pad;
};
static_assert(sizeof(JRecMix) == sizeof(JRec64), "both records share the LDS staging area");
// The mixed path's staging for a far tile (every pair farther apart than S + h_min): all
// binary32, ball j's block-local position shifted by the offset between the two blocks'
// reference points, so a pair's displacement is one binary32 subtraction per axis.  Three
// arrays rather than one record, so that lane l's reads at step s (slot l + s) are 16 or 4
// bytes apart from its neighbours': every LDS bank once per pass (a 32-byte record puts
// two lanes on each bank of a 16-byte read and eight on each of a 4-byte one).
struct JSoA32 {
    float4 p[2 * TB];   // x, y, z in units of L0 relative to block I's reference point; R / L0
    float q[2 * TB];    // cohesion weight
    float g[2 * TB];    // gravity weight
};
static_assert(sizeof(JSoA32) <= sizeof(JRec64) * 2 * TB, "all records share the LDS staging area");

// Reference point of a 64-ball block: the centre of its bounding box (lo, hi).
__device__ __forceinline__ double3 block_ref(double4 lo, double4 hi)
{
    return make_double3(0.5 * (lo.x + hi.x), 0.5 * (lo.y + hi.y), 0.5 * (lo.z + hi.z));
}

// The pair sweep of every tile: the far-field force of each pair, with the per-pair
// near test only in tiles the bounding-box test cannot rule out.  A near pair
// (closer than S + h_min) contributes nothing here; it is recorded as a bit of the
// partner masks of both balls, for k_ball.  SL_PREC_MIXED evaluates the factor in
// binary32, in units of L0 (median radius) and M0 (median mass), as one sum for vdW and
// gravity: f = Kf (q_i q_j / (s t)^2 + g_i g_j / d^3) with per-ball weights
// q = sqrt(Kv/Kf) (R/L0)^3, g = sqrt(Kg/Kf) m/M0, accumulated in binary32 per tile.  In
// far tiles of non-write steps the geometry is binary32 too (block-local coordinates);
// elsewhere it is binary64, rounded once the near test is done.  Each tile writes its i-side and
// j-side sums and masks once, to slot J of the balls of block I and slot I of the
// balls of block J, and its far-pair energy on write steps.
template <int PREC, bool WRITE, bool GRAV, int SPLIT>
__device__ __forceinline__ void tile_pass(const TileArgs& a, int tbase)
{
    constexpr int TPB = NW / SPLIT;
    constexpr size_t SJ_BYTES = sizeof(JRec64) * TPB * 2 * TB;
    constexpr size_t COMB_BYTES = sizeof(double) * NW * 6 * TB;
    __shared__ __attribute__((aligned(16))) unsigned char lds[SJ_BYTES > COMB_BYTES ? SJ_BYTES : COMB_BYTES];
    JRec64 (*sj)[2 * TB] = reinterpret_cast<JRec64 (*)[2 * TB]>(lds);
    double (*comb)[6][TB] = reinterpret_cast<double (*)[6][TB]>(lds);
    __shared__ double pe_w[NW];
    __shared__ unsigned long long im[TPB][TB], jm[TPB][TB];
    __shared__ unsigned long long bal[NW][TB];   // per wavefront: near lanes of step s

    // the wavefront index is uniform: in a scalar register, the tile and sweep indices
    // derived from it are too
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB), l = threadIdx.x % TB;
    const int tl = w / SPLIT, h = w % SPLIT;
    const int t = tbase + tl;
    const bool live = t < a.n_tiles;
    const int4 tt = fetch_tile(a, t, live);
    const int I = tt.x, J = tt.y, sys = tt.z, blk0 = tt.w;
    const Phys& P = a.sysP[sys];
    const bool diag = I == J;
    const bool test = live && !tile_is_far(a, I, J, P.h_min);
    const int i = I * TB + l;
    const int j0 = J * TB + l;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i], Ri = a.R[i];
    const double qi = PREC == SL_PREC_MIXED ? (double)a.aux.q32[i] : a.aux.q[i];
    const double gi = PREC == SL_PREC_MIXED ? (double)a.aux.g32[i] : a.aux.g[i];
    // mixed precision, far tile, no energies: the whole pair in binary32
    const bool f32 = PREC == SL_PREC_MIXED && !WRITE && live && !test;
    if (live && h == 0) {
        if (f32) {
            const double3 ci = block_ref(a.bnd[2 * I], a.bnd[2 * I + 1]);
            const double3 cj = block_ref(a.bnd[2 * J], a.bnd[2 * J + 1]);
            const double L0inv = (double)a.L0inv;
            const float4 xj = a.xl[j0];
            const float4 pj = make_float4(xj.x + (float)((cj.x - ci.x) * L0inv),
                                          xj.y + (float)((cj.y - ci.y) * L0inv),
                                          xj.z + (float)((cj.z - ci.z) * L0inv), xj.w);
            const float qj = a.aux.q32[j0], gj = a.aux.g32[j0];
            JSoA32* sm = reinterpret_cast<JSoA32*>(sj[tl]);
            sm->p[l] = pj;
            sm->p[l + TB] = pj;
            sm->q[l] = qj;
            sm->q[l + TB] = qj;
            sm->g[l] = gj;
            sm->g[l + TB] = gj;
        } else if constexpr (PREC == SL_PREC_MIXED) {
            JRecMix r;
            r.x = a.px[j0];
            r.y = a.py[j0];
            r.z = a.pz[j0];
            r.R = a.R[j0];
            r.Rs = (float)a.R[j0] * a.L0inv;
            r.q = a.aux.q32[j0];
            r.g = a.aux.g32[j0];
            r.pad = 0.f;
            JRecMix* sm = reinterpret_cast<JRecMix*>(sj[tl]);
            sm[l] = r;
            sm[l + TB] = r;
        } else {
            JRec64 r;
            r.x = a.px[j0];
            r.y = a.py[j0];
            r.z = a.pz[j0];
            r.R = a.R[j0];
            r.q = a.aux.q[j0];
            r.g = a.aux.g[j0];
            sj[tl][l] = r;
            sj[tl][l + TB] = r;
        }
        im[tl][l] = 0;
        jm[tl][l] = 0;
    }
    __syncthreads();

    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, pe = 0;
    unsigned long long any = 0;         // bit s: some lane met a near pair at step s
    if (live) {
        const JRec64* rec = &sj[tl][l];
        const double Ri4 = 4 * Ri;
        const double h_min = P.h_min;
        const double hmin2 = h_min * h_min, two_hmin = 2 * h_min;
        int s0, s1;
        sweep_range<SPLIT>(diag, h, s0, s1);
        if (PREC == SL_PREC_MIXED && f32) {
            // far tile: displacement, distance, the vdW and gravity factor and the sums
            // all in binary32, in units of L0 and M0 (no near test: the box test cleared
            // every pair)
            const float4 xi4 = a.xl[i];
            const float qif = a.aux.q32[i], gif = a.aux.g32[i];
            const float r4i = 4.0f * xi4.w;
            float fax = 0.f, fay = 0.f, faz = 0.f, fbx = 0.f, fby = 0.f, fbz = 0.f;
            const JSoA32* soa = reinterpret_cast<const JSoA32*>(sj[tl]);
            const float4* rp32 = soa->p + l;
            const float* rq32 = soa->q + l;
            const float* rg32 = soa->g + l;
            // Gravity is left out of a tile whose every pair has it below 2^-27 of the vdW
            // term: g_i g_j / d^3 over q_i q_j / (s t)^2 is at most (g/q)_I (g/q)_J dmax^5,
            // because s, t <= d^2 <= dmax^2 (dmax: the largest distance between the blocks'
            // boxes).  Added by the FMA it would be under half an ulp of the factor, so
            // leaving it out changes no bit.
            bool grav = GRAV;
            if (GRAV) {
                const double4 li = a.bnd[2 * I], ui = a.bnd[2 * I + 1], lj = a.bnd[2 * J], uj = a.bnd[2 * J + 1];
                const double ex = fmax(uj.x - li.x, ui.x - lj.x), ey = fmax(uj.y - li.y, ui.y - lj.y);
                const double ez = fmax(uj.z - li.z, ui.z - lj.z);
                const double dmax = __builtin_sqrt(ex * ex + ey * ey + ez * ez) * (double)a.L0inv;
                const double d2m = dmax * dmax;
                const double bound = (double)a.gq_blk[I] * (double)a.gq_blk[J] * d2m * d2m * dmax;
                grav = !(bound < 0x1p-27);
            }
            auto pair32 = [&](int s, auto gv) {
                const float4 rp = rp32[s];
                const float dx = rp.x - xi4.x, dy = rp.y - xi4.y, dz = rp.z - xi4.z;
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float S = xi4.w + rp.w;
                const float sv = __builtin_fmaf(-S, S, d2);
                const float tv = __builtin_fmaf(r4i, rp.w, sv);
                const float st = sv * tv;
                float f = (qif * rq32[s]) * __builtin_amdgcn_rcpf(st * st);
                if constexpr (decltype(gv)::value) {
                    const float rq = __builtin_amdgcn_rsqf(d2);
                    f = __builtin_fmaf(gif * rg32[s], rq * rq * rq, f);
                }
                fax = __builtin_fmaf(f, dx, fax);
                fay = __builtin_fmaf(f, dy, fay);
                faz = __builtin_fmaf(f, dz, faz);
                fbx = rot1(__builtin_fmaf(-f, dx, fbx));
                fby = rot1(__builtin_fmaf(-f, dy, fby));
                fbz = rot1(__builtin_fmaf(-f, dz, fbz));
            };
            {
                constexpr int NS = TB / SPLIT;
                if (GRAV && grav) {
#pragma unroll SL_TILE_UNROLL
                    for (int k = 0; k < NS; k++) pair32(s0 + k, std::true_type{});
                } else {
#pragma unroll SL_TILE_UNROLL
                    for (int k = 0; k < NS; k++) pair32(s0 + k, std::false_type{});
                }
            }
            const int back = (TB - (s1 & (TB - 1))) & (TB - 1);
            if (back) {
                fbx = rotk(fbx, back);
                fby = rotk(fby, back);
                fbz = rotk(fbz, back);
            }
            // binary32 sums, exact in binary64 (Kf is applied in k_ball)
            ax = fax; ay = fay; az = faz;
            bx = fbx; by = fby; bz = fbz;
        } else if constexpr (PREC == SL_PREC_MIXED) {
            const float qif = (float)qi, gif = (float)gi, L0inv2 = a.L0inv2, L0inv = a.L0inv;
            const float r4i = 4.0f * (float)Ri * L0inv;
            float fax = 0.f, fay = 0.f, faz = 0.f, fbx = 0.f, fby = 0.f, fbz = 0.f;
            const JRecMix* recm = reinterpret_cast<const JRecMix*>(sj[tl]) + l;
            auto pair = [&](int s, bool act, auto ntest) {
                const JRecMix& r = recm[s];
                const double dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
                const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                const double S = Ri + r.R;
                const double sv = __builtin_fma(-S, S, d2);
                const float s32 = (float)sv * L0inv2;
                const float t32 = __builtin_fmaf(r4i, r.Rs, s32);
                const float st = s32 * t32;
                float f = (qif * r.q) * __builtin_amdgcn_rcpf(st * st);
                if (GRAV) {
                    const float rq = __builtin_amdgcn_rsqf((float)d2 * L0inv2);
                    f = __builtin_fmaf(gif * r.g, rq * rq * rq, f);
                }
                bool use = true;
                if constexpr (decltype(ntest)::value) {
                    const bool near = sv < __builtin_fma(S, two_hmin, hmin2);
                    use = act && !near;
                    f = use ? f : 0.f;
                    const unsigned long long b = __ballot(act && near);
                    if (b) {
                        if (l == 0) bal[w][s] = b;
                        any |= 1ull << s;
                    }
                }
                // the unit scale on the factor once, not on each displacement
                f = f * L0inv;
                const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
                fax = __builtin_fmaf(f, dxf, fax);
                fay = __builtin_fmaf(f, dyf, fay);
                faz = __builtin_fmaf(f, dzf, faz);
                fbx = rot1(__builtin_fmaf(-f, dxf, fbx));
                fby = rot1(__builtin_fmaf(-f, dyf, fby));
                fbz = rot1(__builtin_fmaf(-f, dzf, fbz));
                if (WRITE && use) {
                    const double tv = __builtin_fma(Ri4, r.R, sv);
                    double u = far_energy(P.Ha, Ri, r.R, sv, tv);
                    if (GRAV) u = u - (a.aux.g[i] * a.aux.g[J * TB + ((l + s) & (TB - 1))]) * rsq_f64(d2);
                    pe = pe + u;
                }
            };
            auto sweep = [&](auto ntest) {
                if (!diag) {
                    constexpr int NS = TB / SPLIT;
                    if constexpr (WRITE) {
                        for (int k = 0; k < NS; k++) pair(s0 + k, true, ntest);
                    } else {
#pragma unroll SL_TILE_UNROLL
                        for (int k = 0; k < NS; k++) pair(s0 + k, true, ntest);
                    }
                } else {
                    for (int s = s0; s < s1 - 1; s++) pair(s, true, ntest);
                    pair(s1 - 1, s1 - 1 < TB / 2 || l < TB / 2, ntest);
                }
            };
            if (test)
                sweep(std::true_type{});
            else
                sweep(std::false_type{});
            const int back = (TB - (s1 & (TB - 1))) & (TB - 1);
            if (back) {
                fbx = rotk(fbx, back);
                fby = rotk(fby, back);
                fbz = rotk(fbz, back);
            }
            // binary32 sums, exact in binary64 (Kf is applied in k_ball)
            ax = fax; ay = fay; az = faz;
            bx = fbx; by = fby; bz = fbz;
        } else {
            auto pair = [&](int s, bool act, auto ntest) {
                const JRec64& r = rec[s];
                const double dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
                const double Rj = r.R;
                const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                const double S = Ri + Rj;
                const double sv = __builtin_fma(-S, S, d2);      // d^2 - S^2
                const double tv = __builtin_fma(Ri4, Rj, sv);    // d^2 - (Ri - Rj)^2
                const double st = sv * tv;
                double f = (qi * r.q) * rcp_f64(st * st);
                double rinv = 0;
                if (GRAV) {
                    rinv = rsq_f64(d2);
                    f = __builtin_fma(gi * r.g, rinv * rinv * rinv, f);
                }
                bool use = true;
                if constexpr (decltype(ntest)::value) {
                    const bool near = sv < __builtin_fma(S, two_hmin, hmin2);   // (S + h_min)^2 - S^2
                    use = act && !near;
                    f = use ? f : 0.0;
                    const unsigned long long b = __ballot(act && near);
                    if (b) {
                        if (l == 0) bal[w][s] = b;
                        any |= 1ull << s;
                    }
                }
                ax = __builtin_fma(f, dx, ax);
                ay = __builtin_fma(f, dy, ay);
                az = __builtin_fma(f, dz, az);
                bx = rot1(__builtin_fma(-f, dx, bx));
                by = rot1(__builtin_fma(-f, dy, by));
                bz = rot1(__builtin_fma(-f, dz, bz));
                if (WRITE && use) {
                    double u = far_energy(P.Ha, Ri, Rj, sv, tv);
                    if (GRAV) u = u - (gi * r.g) * rinv;
                    pe = pe + u;
                }
            };
            auto sweep = [&](auto ntest) {
                if (!diag) {
                    constexpr int NS = TB / SPLIT;
                    if constexpr (WRITE) {
                        for (int k = 0; k < NS; k++) pair(s0 + k, true, ntest);
                    } else {
#pragma unroll SL_TILE_UNROLL
                        for (int k = 0; k < NS; k++) pair(s0 + k, true, ntest);
                    }
                } else {
                    for (int s = s0; s < s1 - 1; s++) pair(s, true, ntest);
                    pair(s1 - 1, s1 - 1 < TB / 2 || l < TB / 2, ntest);
                }
            };
            if (test)
                sweep(std::true_type{});
            else
                sweep(std::false_type{});
            const int back = (TB - (s1 & (TB - 1))) & (TB - 1);
            if (back) {
                bx = rotk(bx, back);
                by = rotk(by, back);
                bz = rotk(bz, back);
            }
        }
        // this lane's near steps, from the wavefront's recorded steps
        unsigned long long nmask = 0;
        for (unsigned long long z = any; z; z &= z - 1) {
            const int s = __ffsll((long long)z) - 1;
            if ((bal[w][s] >> l) & 1ull) nmask |= 1ull << s;
        }
        if (nmask) {
            atomicOr(&im[tl][l], rotl64(nmask, l));
            unsigned long long m = nmask;
            while (m) {
                const int s = __ffsll((long long)m) - 1;
                m &= m - 1;
                atomicOr(&jm[tl][(l + s) & (TB - 1)], 1ull << l);
            }
        }
        if (WRITE) {
            const double ws = wave_sum(pe);
            if (l == 0) pe_w[w] = ws;
        }
    }
    __syncthreads();
    if (live) {
        comb[w][0][l] = ax;
        comb[w][1][l] = ay;
        comb[w][2][l] = az;
        comb[w][3][l] = bx;
        comb[w][4][l] = by;
        comb[w][5][l] = bz;
    }
    __syncthreads();
    if (!(live && h == 0)) return;
    const int w0 = tl * SPLIT;
    const size_t np = (size_t)a.n_pad;
    const size_t si = (size_t)(J - blk0), sj_ = (size_t)(I - blk0);
    const unsigned long long mi = im[tl][l], mj = jm[tl][l];
    // The mixed mode's partials are binary32: the wavefronts' sums (binary32 values, held
    // exactly in comb) are added in binary32 and stored in the first half of the buffer.
    using PT = std::conditional_t<PREC == SL_PREC_MIXED, float, double>;
    PT v[6];
    for (int k = 0; k < 6; k++) {
        PT acc = (PT)comb[w0][k][l];
        for (int q = 1; q < SPLIT; q++) acc = acc + (PT)comb[w0 + q][k][l];
        v[k] = acc;
    }
    PT* part = reinterpret_cast<PT*>(a.part);
    if (diag) {
        // the i side and the j side of the diagonal tile land on the same balls
        part[(si * 3 + 0) * np + i] = v[0] + v[3];
        part[(si * 3 + 1) * np + i] = v[1] + v[4];
        part[(si * 3 + 2) * np + i] = v[2] + v[5];
        a.nbm[si * np + i] = mi | mj;
    } else {
        part[(si * 3 + 0) * np + i] = v[0];
        part[(si * 3 + 1) * np + i] = v[1];
        part[(si * 3 + 2) * np + i] = v[2];
        part[(sj_ * 3 + 0) * np + j0] = v[3];
        part[(sj_ * 3 + 1) * np + j0] = v[4];
        part[(sj_ * 3 + 2) * np + j0] = v[5];
        a.nbm[si * np + i] = mi;
        a.nbm[sj_ * np + j0] = mj;
    }
    if (WRITE && l == 0) {
        double pt = pe_w[w0];
        for (int q = 1; q < SPLIT; q++) pt = pt + pe_w[w0 + q];
        a.pe_tile[t] = pt;
    }
}

// One launch, one workgroup per TPB tiles of the list.
template <int PREC, bool WRITE, bool GRAV, int SPLIT>
__global__ __launch_bounds__(NT) void k_tiles(TileArgs a)
{
    tile_pass<PREC, WRITE, GRAV, SPLIT>(a, blockIdx.x * (NW / SPLIT));
}

// The same tiles from a work queue: as many workgroups as fit on the device at once, each
// taking the next group of TPB tiles until none is left.  The list is cut into QSEG
// segments with a counter each, QPAD ints apart (device-scope atomics on one cache line
// serialise, some 10-30 ns apiece); workgroup b starts on segment b mod QSEG and, when
// that is empty, takes from the others in turn.  Every tile runs the same code as in
// k_tiles, so the results are the same bits.  queue[QPAD s] is segment s's next group,
// queue[QPAD QSEG] counts the workgroups done; the last one rewinds them all for the next
// launch, after every workgroup has taken its last item.
constexpr int QSEG = 8;
constexpr int QPAD = 64;   // 256 bytes between counters
template <int PREC, bool WRITE, bool GRAV, int SPLIT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WRITE ? 1 : 6))) void k_tiles_q(TileArgs a)
{
    constexpr int TPB = NW / SPLIT;
    __shared__ int next;
    const int groups = (a.n_tiles + TPB - 1) / TPB;
    const int seg0 = blockIdx.x % QSEG;
    for (int k = 0; k < QSEG; k++) {
        const int seg = (seg0 + k) % QSEG;
        const int g0 = (int)((long long)groups * seg / QSEG), g1 = (int)((long long)groups * (seg + 1) / QSEG);
        for (;;) {
            if (threadIdx.x == 0) next = g0 + atomicAdd(a.queue + QPAD * seg, 1);
            __syncthreads();
            const int g = __builtin_amdgcn_readfirstlane(next);   // uniform: scalar registers
            __syncthreads();   // every thread has read `next` before it is written again
            if (g >= g1) break;
            tile_pass<PREC, WRITE, GRAV, SPLIT>(a, g * TPB);
            __syncthreads();   // the pass's LDS is reused by the next item
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(a.queue + QPAD * QSEG, 1) == (int)gridDim.x - 1)
            for (int q = 0; q <= QSEG; q++) atomicExch(a.queue + QPAD * q, 0);
    }
}

__device__ __forceinline__ D3 load3(const State& c, int f, int b)
{
    return d3(c.f[f][b], c.f[f + 1][b], c.f[f + 2][b]);
}

// Result of one (ball, partner) near evaluation, for the ball: force, torque, and the
// pair's energy and touching flag when the ball is A (counted once per pair).
struct NearRes {
    double f[6];
    double pe;
    int contact, is_a;
};

// Velocities, spins and previous positions of a pair's two balls, read only when
// the pair touches.
struct NearStateG {
    const State& s;
    int A, B;
    __device__ D3 vA() const { return d3(s.f[VX][A], s.f[VY][A], s.f[VZ][A]); }
    __device__ D3 vB() const { return d3(s.f[VX][B], s.f[VY][B], s.f[VZ][B]); }
    __device__ D3 wA() const { return d3(s.f[WX][A], s.f[WY][A], s.f[WZ][A]); }
    __device__ D3 wB() const { return d3(s.f[WX][B], s.f[WY][B], s.f[WZ][B]); }
    __device__ double old_dist() const
    {
        return norm(d3(s.f[QX][B], s.f[QY][B], s.f[QZ][B]) - d3(s.f[QX][A], s.f[QY][A], s.f[QZ][A]));
    }
};

// Evaluate the near pair of ball b and partner p with the reference arithmetic in the
// reference orientation: A is the ball with the larger index in the caller's
// numbering (the friction terms are not symmetric in A and B).
template <bool WRITE>
__device__ __forceinline__ void eval_entry(const State& c, const int* perm, const Phys& P, int b, int p, NearRes& r)
{
    const bool is_a = perm[b] > perm[p];
    const int A = is_a ? b : p, B = is_a ? p : b;
    PairIn in;
    in.pA = load3(c, PX, A);
    in.pB = load3(c, PX, B);
    in.RA = c.f[RAD][A];
    in.RB = c.f[RAD][B];
    in.mA = c.f[MASS][A];
    in.mB = c.f[MASS][B];
    NearStateG ns{c, A, B};
    PairOut po;
    pair_ref(P, WRITE, in, ns, po);
    const D3 F = is_a ? po.F : -po.F;
    const D3 T = is_a ? po.tA : po.tB;
    r.f[0] = F.x; r.f[1] = F.y; r.f[2] = F.z;
    r.f[3] = T.x; r.f[4] = T.y; r.f[5] = T.z;
    r.pe = is_a ? po.pe : 0.0;
    r.contact = is_a ? po.contact : 0;
    r.is_a = is_a;
}

struct BallArgs {
    State cur, nxt;
    const Phys* sysP;
    const int* blk_sys;
    const int* sys_off;
    const int* sys_n;
    const int* sys_nb;
    int n_used, n_pad;
    const double* part;     // [slot][xyz][n_pad]; mixed: binary32, to be scaled by Kf
    double Kf;
    const unsigned long long* nbm;
    const int* perm;    // engine slot -> the caller's ball index (fixes near-pair orientation)
    int kick, drift_next;
    double4* bnd;       // per 64-ball block: bounding box of the drifted positions (lo, hi)
    float4* xl;         // per ball: drifted position relative to its block's box centre, R (units of L0)
    double L0inv;
    double* e_blk;      // FAST_E_STRIDE doubles per 64-ball block (write steps)
    unsigned long long* blk_ctr;   // per 64-ball block: running sums of touching and near pairs
    int* flags;         // [1]: a non-finite state appeared
};

// Block-local binary32 position of ball l of a block whose bounding box is (lo, hi); padding
// lanes get far, distinct, finite places (their weights are 0, so they add nothing).
__device__ __forceinline__ float4 local_pos(double x, double y, double z, double R, bool valid, double4 lo,
                                            double4 hi, double L0inv, int l)
{
    if (!valid) return make_float4(1e6f + 2048.f * (float)l, 1e6f, 1e6f, 0.f);
    const double3 c = block_ref(lo, hi);
    return make_float4((float)((x - c.x) * L0inv), (float)((y - c.y) * L0inv), (float)((z - c.z) * L0inv),
                       (float)(R * L0inv));
}

constexpr int BPB = TB;              // balls per k_ball block
constexpr int KB_C = 8;              // slot classes K = c, c + KB_C, ...: they fix the order of every sum
constexpr int SLOTS_IN_FLIGHT = 10;  // slots per lane loaded together: 40 loads, one batch for systems of up to 80 blocks
// k_ball workgroups have 8 wavefronts (one slot class each: the fewest dependent loads
// per wavefront, for a lone system of few blocks) or 4 (two classes each, 161 VGPRs and
// 48 KB of LDS: three workgroups share a CU, for launches of more blocks than CUs).
// The results are the same bit for bit.
constexpr int kb_threads(int nwv) { return nwv * TB; }

// One 64-ball block per workgroup; lane l of every wavefront works for ball l.  The
// slots of the block fall into KB_C classes, K = c, c + KB_C, ...; wavefront w takes the
// classes c = w, w + NWV, ...
//
// 1. For each of its classes a wavefront sums the far-field partials in slot order
//    (reads of 64 consecutive doubles; in the mixed mode binary32 sums scaled by Kf
//    and added with Kahan compensation), counts
//    the near-partner bits and keeps the first MC nonzero partner masks of each ball
//    in LDS (more than that: they are read again).
// 2. The block's near entries are numbered ball by ball (then by class, slot and bit),
//    and all threads evaluate them, one round of NWV * 64 at a time, with the reference
//    arithmetic and orientation.  Lane l of wavefront 0 adds its ball's results in
//    entry order.  Each near pair is so evaluated once from each of its balls,
//    identically, and no sum needs an atomic.
// 3. Wavefront 0 combines the KB_C class sums in a fixed order and finishes the step
//    for the 64 balls (coalesced stores).
template <bool COMP, bool WRITE, int NWV>
__global__ __launch_bounds__(kb_threads(NWV)) void k_ball(BallArgs a)
{
    static_assert(NWV == 8 || NWV == 4, "8 or 4 wavefronts");
    constexpr int KB_W = NWV;                 // wavefronts
    constexpr int KB_T = kb_threads(NWV);     // threads, and near entries per round
    constexpr int KB_MC = NWV == 8 ? 4 : 2;   // partner masks kept per (class, ball)
    __shared__ double sf[KB_C][3][TB];       // far-field force sums per class
    __shared__ int s_cnt[KB_C][TB];          // near entries per (class, ball)
    __shared__ int s_off[KB_C][TB];          // first entry of (class, ball)
    __shared__ int s_nm[KB_C][TB];           // slots of (class, ball) with near partners
    __shared__ int s_total;
    __shared__ int ent[KB_T];                // entries of the current round
    __shared__ NearRes res[KB_T];
    __shared__ unsigned long long s_mk[KB_C][KB_MC][TB];   // nonzero partner masks ...
    __shared__ int s_mks[KB_C][KB_MC][TB];                  // ... and their slots
    // the wavefront index is uniform: in a scalar register, the slot indices derived from
    // it are too
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / TB), l = threadIdx.x % TB;
    const int i = blockIdx.x * TB + l;
    const int sys = a.blk_sys[blockIdx.x];
    const int off = a.sys_off[sys];
    const bool valid = (i - off) < a.sys_n[sys];
    const int ic = valid ? i : off;
    const int nblk = valid ? a.sys_nb[sys] : 0;
    const Phys& P = a.sysP[sys];
    const State& c = a.cur;
    const size_t np = (size_t)a.n_pad;
    // far-field partials: binary32 in the mixed mode (the force is Kf times the stored sum)
    using PT = std::conditional_t<COMP, float, double>;
    const double Kf = a.Kf;

    // wavefront 0 finishes the step: its own-state loads (and the block's running pair
    // counters) go out first
    unsigned long long ctr0 = 0, ctr1 = 0;
    if (w == 0 && l == 0) {
        ctr0 = a.blk_ctr[2 * blockIdx.x];
        ctr1 = a.blk_ctr[2 * blockIdx.x + 1];
    }
    D3 p0 = d3(0, 0, 0), h0 = p0, g0 = p0;
    double mi = 1, Ii = 1, Rb = 0;
    if (w == 0 && valid) {
        p0 = load3(c, PX, i);
        h0 = load3(c, HX, i);
        g0 = load3(c, GX, i);
        mi = c.f[MASS][i];
        Ii = c.f[MOI][i];
        Rb = c.f[RAD][i];
    }

#pragma unroll 1
    for (int cl = w; cl < KB_C; cl += KB_W) {
        double fx = 0, fy = 0, fz = 0, cx = 0, cy = 0, cz = 0;
        int cn = 0;                    // near entries of this (class, ball)
        int nm = 0;                    // slots of this (class, ball) with near partners
        for (int K0 = cl; K0 < nblk; K0 += SLOTS_IN_FLIGHT * KB_C) {
            PT q[SLOTS_IN_FLIGHT][3];
            unsigned long long mk[SLOTS_IN_FLIGHT];
#pragma unroll
            for (int u = 0; u < SLOTS_IN_FLIGHT; u++) {
                const int K = K0 + u * KB_C;
                const int Kc = K < nblk ? K : 0;
                const PT* pp = reinterpret_cast<const PT*>(a.part) + (size_t)Kc * 3 * np + ic;
                q[u][0] = pp[0];
                q[u][1] = pp[np];
                q[u][2] = pp[2 * np];
                mk[u] = K < nblk ? a.nbm[(size_t)Kc * np + ic] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < SLOTS_IN_FLIGHT; u++) {
                if (K0 + u * KB_C >= nblk) break;
                cn += __popcll(mk[u]);
                if (mk[u]) {
                    if (nm < KB_MC) {
                        s_mk[cl][nm][l] = mk[u];
                        s_mks[cl][nm][l] = K0 + u * KB_C;
                    }
                    nm++;
                }
                if (COMP) {
                    double y, t;
                    y = Kf * (double)q[u][0] - cx; t = fx + y; cx = (t - fx) - y; fx = t;
                    y = Kf * (double)q[u][1] - cy; t = fy + y; cy = (t - fy) - y; fy = t;
                    y = Kf * (double)q[u][2] - cz; t = fz + y; cz = (t - fz) - y; fz = t;
                } else {
                    fx = fx + q[u][0];
                    fy = fy + q[u][1];
                    fz = fz + q[u][2];
                }
            }
        }
        sf[cl][0][l] = fx - cx;
        sf[cl][1][l] = fy - cy;
        sf[cl][2][l] = fz - cz;
        s_cnt[cl][l] = cn;
        s_nm[cl][l] = nm;
    }
    __syncthreads();

    // number the block's near entries: ball-major, then class, slot and bit
    if (w == 0) {
        int tot = 0;
        for (int q = 0; q < KB_C; q++) tot += s_cnt[q][l];
        int incl = tot;
        for (int m = 1; m < TB; m <<= 1) {
            const int y = __shfl_up(incl, m, TB);
            if (l >= m) incl += y;
        }
        int o = incl - tot;
        for (int q = 0; q < KB_C; q++) {
            s_off[q][l] = o;
            o += s_cnt[q][l];
        }
        if (l == TB - 1) s_total = incl;
    }
    __syncthreads();
    const int total = s_total;

    double n[6] = {0, 0, 0, 0, 0, 0};
    double npe = 0;
    int ncon = 0, npair = 0;
    for (int r0 = 0; r0 < total; r0 += KB_T) {
        const int r1 = r0 + KB_T;
        // write this round's entries of (class cl, ball l) for the classes of wavefront w
#pragma unroll 1
        for (int cl = w; cl < KB_C; cl += KB_W) {
            const int cn = s_cnt[cl][l], nm = s_nm[cl][l];
            const int my0 = s_off[cl][l];   // (class, ball) entries are [my0, my0 + cn)
            if (cn && my0 < r1 && my0 + cn > r0) {
                int k = my0;
                auto put = [&](int K, unsigned long long m) {
                    while (m && k < r1) {
                        const int bit = __ffsll((long long)m) - 1;
                        m &= m - 1;
                        if (k >= r0) ent[k - r0] = l | (bit << 6) | (K << 12);
                        k++;
                    }
                };
                if (nm <= KB_MC) {
                    for (int q = 0; q < nm && k < r1; q++) put(s_mks[cl][q][l], s_mk[cl][q][l]);
                } else {
                    for (int K = cl; K < nblk && k < r1; K += KB_C) put(K, a.nbm[(size_t)K * np + ic]);
                }
            }
        }
        __syncthreads();
        const int t = threadIdx.x;
        if (r0 + t < total) {
            const int e = ent[t];
            const int el = e & (TB - 1), bit = (e >> 6) & (TB - 1), K = e >> 12;
            NearRes rr;
            eval_entry<WRITE>(c, a.perm, P, blockIdx.x * TB + el, off + K * TB + bit, rr);
            res[t] = rr;
        }
        __syncthreads();
        if (w == 0) {
            // ball l's entries in this round, in order
            int tot = 0;
            for (int q = 0; q < KB_C; q++) tot += s_cnt[q][l];
            const int b0 = s_off[0][l], b1 = b0 + tot;
            const int k0 = b0 > r0 ? b0 : r0, k1 = b1 < r1 ? b1 : r1;
            for (int k = k0; k < k1; k++) {
                const NearRes& rr = res[k - r0];
                for (int q = 0; q < 6; q++) n[q] = n[q] + rr.f[q];
                npe = npe + rr.pe;
                ncon += rr.contact;
                npair += rr.is_a;
            }
        }
        __syncthreads();
    }
    if (w != 0) return;

    // wavefront 0: combine in a fixed order and finish the step of ball l
    double F[3], T[3];
    for (int k = 0; k < 3; k++) {
        double s = sf[0][k][l], cmp = 0;
        for (int q = 1; q < KB_C; q++) {
            if (COMP) {
                const double y = sf[q][k][l] - cmp, t = s + y;
                cmp = (t - s) - y;
                s = t;
            } else {
                s = s + sf[q][k][l];
            }
        }
        F[k] = (s - cmp) + n[k];
        T[k] = n[3 + k];
    }

    double e[FAST_E_STRIDE] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // PE KE p L contacts nonfinite
    D3 pn = p0;   // positions the next pair pass will read
    if (valid) {
        const double px = p0.x, py = p0.y, pz = p0.z;
        const double ax = F[0] / mi, ay = F[1] / mi, az = F[2] / mi;
        const double bx = T[0] / Ii, by = T[1] / Ii, bz = T[2] / Ii;
        const double dt = P.dt;
        double vx, vy, vz, wx, wy, wz;
        if (a.kick) {
            vx = h0.x + .5 * ax * dt;
            vy = h0.y + .5 * ay * dt;
            vz = h0.z + .5 * az * dt;
            wx = g0.x + .5 * bx * dt;
            wy = g0.y + .5 * by * dt;
            wz = g0.z + .5 * bz * dt;
        } else {
            vx = c.f[VX][i]; vy = c.f[VY][i]; vz = c.f[VZ][i];
            wx = c.f[WX][i]; wy = c.f[WY][i]; wz = c.f[WZ][i];
        }
        const bool bad = !(isfinite(vx) && isfinite(vy) && isfinite(vz) && isfinite(wx) && isfinite(wy) &&
                           isfinite(wz) && isfinite(px) && isfinite(py) && isfinite(pz));
        if (bad) atomicOr(&a.flags[1], 1);
        if (WRITE) {
            const D3 v = d3(vx, vy, vz), wv = d3(wx, wy, wz);
            e[1] = .5 * mi * dot(v, v) + .5 * Ii * dot(wv, wv);
            const D3 mv = v * mi;
            e[2] = mv.x; e[3] = mv.y; e[4] = mv.z;
            const D3 L = cross(p0, v) * mi + wv * Ii;
            e[5] = L.x; e[6] = L.y; e[7] = L.z;
            e[9] = bad ? 1.0 : 0.0;
        }
        const State& o = a.nxt;
        o.f[VX][i] = vx; o.f[VY][i] = vy; o.f[VZ][i] = vz;
        o.f[WX][i] = wx; o.f[WY][i] = wy; o.f[WZ][i] = wz;
        o.f[AX][i] = ax; o.f[AY][i] = ay; o.f[AZ][i] = az;
        o.f[BX][i] = bx; o.f[BY][i] = by; o.f[BZ][i] = bz;
        if (a.drift_next) {
            const double hx = vx + .5 * ax * dt, hy = vy + .5 * ay * dt, hz = vz + .5 * az * dt;
            o.f[HX][i] = hx; o.f[HY][i] = hy; o.f[HZ][i] = hz;
            o.f[GX][i] = wx + .5 * bx * dt;
            o.f[GY][i] = wy + .5 * by * dt;
            o.f[GZ][i] = wz + .5 * bz * dt;
            o.f[QX][i] = px; o.f[QY][i] = py; o.f[QZ][i] = pz;
            pn = d3(px + hx * dt, py + hy * dt, pz + hz * dt);
            o.f[PX][i] = pn.x;
            o.f[PY][i] = pn.y;
            o.f[PZ][i] = pn.z;
        } else {
            o.f[HX][i] = h0.x; o.f[HY][i] = h0.y; o.f[HZ][i] = h0.z;
            o.f[GX][i] = g0.x; o.f[GY][i] = g0.y; o.f[GZ][i] = g0.z;
            o.f[QX][i] = c.f[QX][i]; o.f[QY][i] = c.f[QY][i]; o.f[QZ][i] = c.f[QZ][i];
            o.f[PX][i] = px; o.f[PY][i] = py; o.f[PZ][i] = pz;
        }
    }
    if (a.drift_next) {
        double4 lo, hi;
        wave_aabb(pn.x, pn.y, pn.z, Rb, valid, lo, hi);
        if (l == 0) {
            a.bnd[2 * blockIdx.x] = lo;
            a.bnd[2 * blockIdx.x + 1] = hi;
        }
        a.xl[i] = local_pos(pn.x, pn.y, pn.z, Rb, valid, lo, hi, a.L0inv, l);
    }
    if (WRITE) {
        // a 64-ball block lies inside one system, so its sums are per system
        for (int k = 1; k < FAST_E_STRIDE; k++) {
            if (k == 8) continue;
            const double s = wave_sum(e[k]);
            if (l == 0) a.e_blk[blockIdx.x * FAST_E_STRIDE + k] = s;
        }
    }
    // the block's near-pair energy and pair counts (each pair counted by its ball A)
    const double bpe = WRITE ? wave_sum(npe) : 0.0;
    const int bcc = wave_isum(ncon), bcn = wave_isum(npair);
    if (l == 0) {
        a.blk_ctr[2 * blockIdx.x] = ctr0 + (unsigned long long)bcc;
        a.blk_ctr[2 * blockIdx.x + 1] = ctr1 + (unsigned long long)bcn;
        if (WRITE) {
            a.e_blk[blockIdx.x * FAST_E_STRIDE] = bpe;
            a.e_blk[blockIdx.x * FAST_E_STRIDE + 8] = (double)bcc;
        }
    }
}

// Bounding boxes of the 64-ball blocks (one wavefront per block), for positions that
// did not come from k_ball's drift.
__global__ __launch_bounds__(TB) void k_bounds(State s, const int* blk_sys, const int* sys_off, const int* sys_n,
                                               double4* bnd, float4* xl, double L0inv)
{
    const int l = threadIdx.x;
    const int b = blockIdx.x * TB + l;
    const int sy = blk_sys[blockIdx.x];
    const bool valid = b - sys_off[sy] < sys_n[sy];
    const int bc = valid ? b : sys_off[sy];
    double4 lo, hi;
    const double x = s.f[PX][bc], y = s.f[PY][bc], z = s.f[PZ][bc], R = s.f[RAD][bc];
    wave_aabb(x, y, z, R, valid, lo, hi);
    if (l == 0) {
        bnd[2 * blockIdx.x] = lo;
        bnd[2 * blockIdx.x + 1] = hi;
    }
    xl[b] = local_pos(x, y, z, R, valid, lo, hi, L0inv, l);
}

// Start-of-call drift of the valid balls (each system with its own dt).
__global__ void k_drift_sys(State s, const int* blk_sys, const int* sys_off, const int* sys_n,
                            const Phys* sysP, int n_used)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_used) return;
    const int sy = blk_sys[b / TB];
    if (b - sys_off[sy] >= sys_n[sy]) return;
    const double dt = sysP[sy].dt;
    const double px = s.f[PX][b], py = s.f[PY][b], pz = s.f[PZ][b];
    s.f[QX][b] = px;
    s.f[QY][b] = py;
    s.f[QZ][b] = pz;
    const double hx = s.f[VX][b] + .5 * s.f[AX][b] * dt;
    const double hy = s.f[VY][b] + .5 * s.f[AY][b] * dt;
    const double hz = s.f[VZ][b] + .5 * s.f[AZ][b] * dt;
    s.f[HX][b] = hx;
    s.f[HY][b] = hy;
    s.f[HZ][b] = hz;
    s.f[GX][b] = s.f[WX][b] + .5 * s.f[BX][b] * dt;
    s.f[GY][b] = s.f[WY][b] + .5 * s.f[BY][b] * dt;
    s.f[GZ][b] = s.f[WZ][b] + .5 * s.f[BZ][b] * dt;
    s.f[PX][b] = px + hx * dt;
    s.f[PY][b] = py + hy * dt;
    s.f[PZ][b] = pz + hz * dt;
}

}  // namespace fastk
}  // namespace sl
