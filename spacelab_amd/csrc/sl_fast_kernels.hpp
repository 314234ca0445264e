// Device kernels of the fast path (included by sl_kernels_fast.hip).
//
//   k_tiles_*  one wavefront per 64x64 tile (I, J) of the pair triangle, I >= J.
//              Lane l holds ball i = 64 I + l in registers.  The 64 balls of block
//              J are staged in LDS twice (slots 0..63 and 64..127), so at sweep
//              step s lane l reads slot l + s -- ball j = (l + s) mod 64 -- at
//              address base + s*stride, with no index arithmetic.  The reaction
//              force on j is accumulated in a register that travels with j: after
//              every step the accumulators rotate one lane down (DPP wave_rol:1),
//              so lane l always holds the sum of the j it meets next.  Neither LDS
//              atomics nor cross-lane reductions are needed.  A pair closer than
//              S + h_min (touching, or in the band where the reference clamps the
//              vdW gap) contributes nothing here; it is staged in LDS (ballot
//              compaction) and appended to both balls' near lists in one flush.
//              Each tile writes its i-side and j-side force sums once, to
//              part[J][.][i] and part[I][.][j].
//   k_ball     eight lanes per ball: sum the tile partials (fixed order; Kahan in
//              the mixed mode), sort the ball's near list in LDS and evaluate its
//              near pairs with the reference arithmetic, then one lane forms acc
//              and aacc, applies the closing kick, the write-step energies and
//              the next step's drift.  It reads the current state set and writes
//              the other one, because near pairs read neighbours' state.
// Every sum has a fixed order, so results are reproducible run to run.
#pragma once

#include "sl_fast_pair.hpp"
#include "sl_internal.hpp"
#include "sl_pair_ref.hpp"

#pragma clang fp contract(off)

namespace sl {
namespace fastk {

constexpr int TB = FAST_TILE;
constexpr int WPB = FAST_TILES_PER_BLOCK;
constexpr int NT = TB * WPB;
constexpr int MAXNB = FAST_MAX_NEAR;
constexpr int NSTAGE = 64;           // near pairs staged per wave before a flush

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

__device__ __forceinline__ void near_append(int a, int b, int* nb_count, int* nb_list, int* flags)
{
    const int sa = atomicAdd(&nb_count[a], 1);
    if (sa < MAXNB)
        nb_list[(size_t)a * MAXNB + sa] = b;
    else
        atomicOr(&flags[0], 1);
}

struct TileArgs {
    const double* px;
    const double* py;
    const double* pz;
    const double* R;
    FastAux aux;
    int n_pad;
    int n_tiles;
    const int2* tiles;
    double h_min;
    double Ha;
    double L0inv2;          // 1 / L0^2 (mixed)
    double Kv, Kg;          // mixed: force = Kv * vdw_sum + Kg * grav_sum
    double* part;
    int* nb_count;
    int* nb_list;
    int* flags;
    double* pe_part;
};

// Near hits of the sweep are recorded branch-free as bits of a per-lane mask
// (bit s: the pair met at step s).  After the sweep each lane turns its bits into
// (i, j) pairs in the wave's LDS stage (slot from an LDS counter), and the whole
// wave then appends the staged pairs to the per-ball lists in parallel.  Pairs
// beyond the stage capacity are appended directly by their lane.
__device__ void flush_near(const TileArgs& a, unsigned long long mask, int i, int J, int2* stage,
                           int* stage_count)
{
    const int l = threadIdx.x % TB;
    if (l == 0) *stage_count = 0;
    __builtin_amdgcn_wave_barrier();
    while (mask) {
        const int s = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int j = J * TB + ((l + s) & (TB - 1));
        const int slot = atomicAdd(stage_count, 1);
        if (slot < NSTAGE) {
            stage[slot] = make_int2(i, j);
        } else {
            near_append(i, j, a.nb_count, a.nb_list, a.flags);
            near_append(j, i, a.nb_count, a.nb_list, a.flags);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    int count = *stage_count;
    count = count < NSTAGE ? count : NSTAGE;
    if (l < count) {
        const int2 p = stage[l];
        near_append(p.x, p.y, a.nb_count, a.nb_list, a.flags);
        near_append(p.y, p.x, a.nb_count, a.nb_list, a.flags);
    }
}

// vdW energy of a far pair from s = d^2 - S^2, t = d^2 - (Ri-Rj)^2.
__device__ __forceinline__ double far_energy(double Ha, double Ri, double Rj, double s, double t)
{
    const double two = 2 * Ri * Rj;
    return -Ha / 6 * (two / s + two / t + ::log(s / t));
}

struct JRec64 {
    double x, y, z, R, q, g;
};

template <bool WRITE, bool GRAV>
__global__ __launch_bounds__(NT) void k_tiles_f64(TileArgs a)
{
    __shared__ JRec64 sj[WPB][2 * TB];
    __shared__ int2 nst[WPB][NSTAGE];
    __shared__ int nstc[WPB];
    __shared__ double red[NT];

    const int w = threadIdx.x / TB, l = threadIdx.x % TB;
    const int t = blockIdx.x * WPB + w;
    const bool live = t < a.n_tiles;
    int I = 0, J = 0;
    if (live) {
        const int2 ij = a.tiles[t];
        I = ij.x;
        J = ij.y;
    }
    const int i = I * TB + l;
    const int j0 = J * TB + l;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i], Ri = a.R[i];
    const double qi = a.aux.q[i], gi = a.aux.g[i];
    const double Ri4 = 4 * Ri;
    {
        JRec64 r;
        r.x = a.px[j0];
        r.y = a.py[j0];
        r.z = a.pz[j0];
        r.R = a.R[j0];
        r.q = a.aux.q[j0];
        r.g = a.aux.g[j0];
        sj[w][l] = r;
        sj[w][l + TB] = r;
    }
    __syncthreads();

    const double h_min = a.h_min;
    const double hmin2 = h_min * h_min;
    const double two_hmin = 2 * h_min;
    double ax = 0, ay = 0, az = 0;      // force on i
    double bx = 0, by = 0, bz = 0;      // reaction on the j this lane meets next
    double pe = 0;
    const bool diag = I == J;
    const JRec64* rec = &sj[w][l];
    unsigned long long nmask = 0;   // bit s: near pair met at step s

    auto pair = [&](int s, bool act) {
        const JRec64& r = rec[s];
        const double dx = r.x - xi;
        const double dy = r.y - yi;
        const double dz = r.z - zi;
        const double Rj = r.R;
        const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        const double S = Ri + Rj;
        const double sv = __builtin_fma(-S, S, d2);          // d^2 - S^2
        const double thr = __builtin_fma(S, two_hmin, hmin2); // (S + h_min)^2 - S^2
        const double tv = __builtin_fma(Ri4, Rj, sv);        // d^2 - (Ri - Rj)^2
        const double st = sv * tv;
        double f = (qi * r.q) * rcp_f64(st * st);
        double rinv = 0;
        if (GRAV) {
            rinv = rsq_f64(d2);
            f = __builtin_fma(gi * r.g, rinv * rinv * rinv, f);
        }
        const bool near = sv < thr;
        const bool use = act && !near;
        f = use ? f : 0.0;
        ax = __builtin_fma(f, dx, ax);
        ay = __builtin_fma(f, dy, ay);
        az = __builtin_fma(f, dz, az);
        bx = rot1(__builtin_fma(-f, dx, bx));
        by = rot1(__builtin_fma(-f, dy, by));
        bz = rot1(__builtin_fma(-f, dz, bz));
        if (WRITE && use) {
            double u = far_energy(a.Ha, Ri, Rj, sv, tv);
            if (GRAV) u = u - (gi * r.g) * rinv;
            pe = pe + u;
        }
        nmask |= (act && near) ? (1ull << s) : 0ull;
    };

    if (live) {
        if (!diag) {
#pragma unroll (WRITE ? 1 : 8)
            for (int s = 0; s < TB; s++) pair(s, true);
        } else {
            // each unordered pair of the block once: offsets 1..31 for every lane,
            // offset 32 for the lower half only
#pragma unroll (WRITE ? 1 : 8)
            for (int s = 1; s < TB / 2; s++) pair(s, true);
            pair(TB / 2, l < TB / 2);
            // lane l now holds the reaction sum of slot l + 33: bring slot l home
            bx = rotk(bx, TB - (TB / 2 + 1));
            by = rotk(by, TB - (TB / 2 + 1));
            bz = rotk(bz, TB - (TB / 2 + 1));
        }
        if (__ballot(nmask != 0)) flush_near(a, nmask, i, J, &nst[w][0], &nstc[w]);
        const size_t np = (size_t)a.n_pad;
        if (diag) {
            // the i side and the j side of the diagonal tile land on the same balls
            a.part[((size_t)J * 3 + 0) * np + i] = ax + bx;
            a.part[((size_t)J * 3 + 1) * np + i] = ay + by;
            a.part[((size_t)J * 3 + 2) * np + i] = az + bz;
        } else {
            a.part[((size_t)J * 3 + 0) * np + i] = ax;
            a.part[((size_t)J * 3 + 1) * np + i] = ay;
            a.part[((size_t)J * 3 + 2) * np + i] = az;
            a.part[((size_t)I * 3 + 0) * np + j0] = bx;
            a.part[((size_t)I * 3 + 1) * np + j0] = by;
            a.part[((size_t)I * 3 + 2) * np + j0] = bz;
        }
    }
    if (WRITE) {
        const double tot = block_sum<NT>(pe, red);
        if (threadIdx.x == 0) a.pe_part[blockIdx.x] = tot;
    }
}

struct JRec32 {
    double x, y, z, R;
    float r, q, g, pad;
};

// Mixed precision: relative displacement, d^2, the near test and s = d^2 - S^2
// in binary64; the vdW and gravity factors and the per-tile force sums in
// binary32 (scaled units; separate sums for vdW and gravity, whose ratio can
// reach 1e-20); tile sums converted to binary64 forces when written.
template <bool WRITE, bool GRAV>
__global__ __launch_bounds__(NT) void k_tiles_mixed(TileArgs a)
{
    __shared__ JRec32 sj[WPB][2 * TB];
    __shared__ int2 nst[WPB][NSTAGE];
    __shared__ int nstc[WPB];
    __shared__ double red[NT];

    const int w = threadIdx.x / TB, l = threadIdx.x % TB;
    const int t = blockIdx.x * WPB + w;
    const bool live = t < a.n_tiles;
    int I = 0, J = 0;
    if (live) {
        const int2 ij = a.tiles[t];
        I = ij.x;
        J = ij.y;
    }
    const int i = I * TB + l;
    const int j0 = J * TB + l;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i], Ri = a.R[i];
    const float r4i = 4.0f * a.aux.r32[i], qi = a.aux.q32[i], gi = a.aux.g32[i];
    const double gi64 = a.aux.g[i];
    {
        JRec32 r;
        r.x = a.px[j0];
        r.y = a.py[j0];
        r.z = a.pz[j0];
        r.R = a.R[j0];
        r.r = a.aux.r32[j0];
        r.q = a.aux.q32[j0];
        r.g = a.aux.g32[j0];
        r.pad = 0.f;
        sj[w][l] = r;
        sj[w][l + TB] = r;
    }
    __syncthreads();

    const double h_min = a.h_min;
    const double hmin2 = h_min * h_min;
    const double two_hmin = 2 * h_min;
    const double L0inv2 = a.L0inv2;
    float vx = 0.f, vy = 0.f, vz = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;       // on i
    float wvx = 0.f, wvy = 0.f, wvz = 0.f, wgx = 0.f, wgy = 0.f, wgz = 0.f; // travelling j sums
    double pe = 0;
    const bool diag = I == J;
    const JRec32* rec = &sj[w][l];
    unsigned long long nmask = 0;   // bit s: near pair met at step s

    auto pair = [&](int s, bool act) {
        const JRec32& r = rec[s];
        const double dx = r.x - xi;
        const double dy = r.y - yi;
        const double dz = r.z - zi;
        const double Rj = r.R;
        const double d2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        const double S = Ri + Rj;
        const double sv = __builtin_fma(-S, S, d2);
        const double thr = __builtin_fma(S, two_hmin, hmin2);
        const bool near = sv < thr;
        const bool use = act && !near;
        const float s32 = (float)(sv * L0inv2);
        const float t32 = __builtin_fmaf(r4i, r.r, s32);
        const float st = s32 * t32;
        float fv = (qi * r.q) * __builtin_amdgcn_rcpf(st * st);
        fv = use ? fv : 0.f;
        const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
        vx = __builtin_fmaf(fv, dxf, vx);
        vy = __builtin_fmaf(fv, dyf, vy);
        vz = __builtin_fmaf(fv, dzf, vz);
        wvx = rot1(__builtin_fmaf(-fv, dxf, wvx));
        wvy = rot1(__builtin_fmaf(-fv, dyf, wvy));
        wvz = rot1(__builtin_fmaf(-fv, dzf, wvz));
        if (GRAV) {
            const float rif = __builtin_amdgcn_rsqf((float)d2);
            float fg = (gi * r.g) * (rif * rif * rif);
            fg = use ? fg : 0.f;
            gx = __builtin_fmaf(fg, dxf, gx);
            gy = __builtin_fmaf(fg, dyf, gy);
            gz = __builtin_fmaf(fg, dzf, gz);
            wgx = rot1(__builtin_fmaf(-fg, dxf, wgx));
            wgy = rot1(__builtin_fmaf(-fg, dyf, wgy));
            wgz = rot1(__builtin_fmaf(-fg, dzf, wgz));
        }
        if (WRITE && use) {
            // write-step energy in binary64 (not on the per-step path)
            const double tv = __builtin_fma(4 * Ri, Rj, sv);
            double u = far_energy(a.Ha, Ri, Rj, sv, tv);
            if (GRAV) u = u - (gi64 * a.aux.g[J * TB + ((l + s) & (TB - 1))]) * rsq_f64(d2);
            pe = pe + u;
        }
        nmask |= (act && near) ? (1ull << s) : 0ull;
    };

    if (live) {
        if (!diag) {
#pragma unroll (WRITE ? 1 : 8)
            for (int s = 0; s < TB; s++) pair(s, true);
        } else {
#pragma unroll (WRITE ? 1 : 8)
            for (int s = 1; s < TB / 2; s++) pair(s, true);
            pair(TB / 2, l < TB / 2);
            constexpr int K = TB - (TB / 2 + 1);
            wvx = rotk(wvx, K); wvy = rotk(wvy, K); wvz = rotk(wvz, K);
            if (GRAV) {
                wgx = rotk(wgx, K); wgy = rotk(wgy, K); wgz = rotk(wgz, K);
            }
        }
        if (__ballot(nmask != 0)) flush_near(a, nmask, i, J, &nst[w][0], &nstc[w]);
        const size_t np = (size_t)a.n_pad;
        const double Kv = a.Kv, Kg = a.Kg;
        double ax = Kv * (double)vx, ay = Kv * (double)vy, az = Kv * (double)vz;
        double bx = Kv * (double)wvx, by = Kv * (double)wvy, bz = Kv * (double)wvz;
        if (GRAV) {
            ax = ax + Kg * (double)gx;
            ay = ay + Kg * (double)gy;
            az = az + Kg * (double)gz;
            bx = bx + Kg * (double)wgx;
            by = by + Kg * (double)wgy;
            bz = bz + Kg * (double)wgz;
        }
        if (diag) {
            a.part[((size_t)J * 3 + 0) * np + i] = ax + bx;
            a.part[((size_t)J * 3 + 1) * np + i] = ay + by;
            a.part[((size_t)J * 3 + 2) * np + i] = az + bz;
        } else {
            a.part[((size_t)J * 3 + 0) * np + i] = ax;
            a.part[((size_t)J * 3 + 1) * np + i] = ay;
            a.part[((size_t)J * 3 + 2) * np + i] = az;
            a.part[((size_t)I * 3 + 0) * np + j0] = bx;
            a.part[((size_t)I * 3 + 1) * np + j0] = by;
            a.part[((size_t)I * 3 + 2) * np + j0] = bz;
        }
    }
    if (WRITE) {
        const double tot = block_sum<NT>(pe, red);
        if (threadIdx.x == 0) a.pe_part[blockIdx.x] = tot;
    }
}

struct NearState {
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

struct BallArgs {
    State cur, nxt;
    Phys P;
    int n, n_pad, nblk;
    const double* part;
    int* nb_count;
    const int* nb_list;
    int write, kick, drift_next, compensated;
    double* pe_part;    // this kernel's PE partials (already offset)
    double* e_part;
    unsigned long long* contacts;
    unsigned long long* near_total;
    int* nonfinite;
};

constexpr int LPB = 8;               // lanes per ball in k_ball
constexpr int BPB = 256 / LPB;       // balls per 256-thread block

// Sum over the LPB lanes of one ball as a fixed tree (deterministic).
__device__ __forceinline__ double lane_group_sum(double v)
{
    v = v + __shfl_xor(v, 1, LPB);
    v = v + __shfl_xor(v, 2, LPB);
    v = v + __shfl_xor(v, 4, LPB);
    return v;
}

__global__ __launch_bounds__(256) void k_ball(BallArgs a)
{
    __shared__ double red[256];
    __shared__ int lst_in[BPB][MAXNB];
    __shared__ int lst_sorted[BPB][MAXNB];
    const int g = threadIdx.x % LPB;
    const int bl = threadIdx.x / LPB;
    const int i = blockIdx.x * BPB + bl;
    const bool valid = i < a.n;
    const int ic = valid ? i : 0;
    const State& c = a.cur;
    const size_t np = (size_t)a.n_pad;
    double pe = 0;
    double e[7] = {0, 0, 0, 0, 0, 0, 0};
    unsigned ncontact = 0;

    // the ball's own state, loaded up front: nothing below depends on it until the
    // kick, and the loads cannot move across the barriers that follow
    const double hx0 = c.f[HX][ic], hy0 = c.f[HY][ic], hz0 = c.f[HZ][ic];
    const double gx0 = c.f[GX][ic], gy0 = c.f[GY][ic], gz0 = c.f[GZ][ic];
    const double vx0 = c.f[VX][ic], vy0 = c.f[VY][ic], vz0 = c.f[VZ][ic];
    const double wx0 = c.f[WX][ic], wy0 = c.f[WY][ic], wz0 = c.f[WZ][ic];
    const double qx0 = c.f[QX][ic], qy0 = c.f[QY][ic], qz0 = c.f[QZ][ic];
    const D3 pi = d3(c.f[PX][ic], c.f[PY][ic], c.f[PZ][ic]);
    const double Ri = c.f[RAD][ic], mi = c.f[MASS][ic], Ii = c.f[MOI][ic];

    // near list into LDS, then sorted by partner index (rank = entries below it)
    const int cnt_raw = valid ? a.nb_count[ic] : 0;
    const int cnt = cnt_raw < MAXNB ? cnt_raw : MAXNB;
    for (int u = g; u < cnt; u += LPB) lst_in[bl][u] = a.nb_list[(size_t)ic * MAXNB + u];
    __syncthreads();
    for (int u = g; u < cnt; u += LPB) {
        const int v = lst_in[bl][u];
        int below = 0;
        for (int x = 0; x < cnt; x++) below += lst_in[bl][x] < v;
        lst_sorted[bl][below] = v;
    }

    // far field: lane g sums the tile partials K = g, g+8, ...
    double fx = 0, fy = 0, fz = 0;
    if (valid) {
        if (a.compensated) {
            double cx = 0, cy = 0, cz = 0;   // Kahan compensation
#pragma unroll 4
            for (int K = g; K < a.nblk; K += LPB) {
                const double* p = a.part + (size_t)K * 3 * np + ic;
                double y, t;
                y = p[0] - cx; t = fx + y; cx = (t - fx) - y; fx = t;
                y = p[np] - cy; t = fy + y; cy = (t - fy) - y; fy = t;
                y = p[2 * np] - cz; t = fz + y; cz = (t - fz) - y; fz = t;
            }
            fx = fx - cx;
            fy = fy - cy;
            fz = fz - cz;
        } else {
#pragma unroll 4
            for (int K = g; K < a.nblk; K += LPB) {
                const double* p = a.part + (size_t)K * 3 * np + ic;
                fx = fx + p[0];
                fy = fy + p[np];
                fz = fz + p[2 * np];
            }
        }
    }
    __syncthreads();

    // near pairs: lane g takes ranks g, g+8, ... with the reference arithmetic
    D3 nF = d3(0, 0, 0), nT = d3(0, 0, 0);
    for (int r = g; r < cnt; r += LPB) {
        const int j = lst_sorted[bl][r];
        if (j < 0 || j >= a.n || j == i) continue;   // unreachable: entries are distinct partners
        const bool iA = i > j;
        const int A = iA ? i : j, B = iA ? j : i;
        const D3 pj = d3(c.f[PX][j], c.f[PY][j], c.f[PZ][j]);
        PairIn in;
        in.pA = iA ? pi : pj;
        in.pB = iA ? pj : pi;
        const double Rj = c.f[RAD][j], mj = c.f[MASS][j];
        in.RA = iA ? Ri : Rj;
        in.RB = iA ? Rj : Ri;
        in.mA = iA ? mi : mj;
        in.mB = iA ? mj : mi;
        PairOut o;
        NearState nsr{c, A, B};
        pair_ref(a.P, a.write && iA, in, nsr, o);
        if (iA) {
            nF = nF + o.F;
            if (o.contact) nT = nT + o.tA;
            pe = pe + o.pe;
            ncontact += o.contact;
        } else {
            nF = nF - o.F;
            if (o.contact) nT = nT + o.tB;
        }
    }
    fx = lane_group_sum(fx + nF.x);
    fy = lane_group_sum(fy + nF.y);
    fz = lane_group_sum(fz + nF.z);
    const double tx = lane_group_sum(nT.x), ty = lane_group_sum(nT.y), tz = lane_group_sum(nT.z);
    unsigned nnear = 0;
    if (valid && g == 0) {
        nnear = (unsigned)cnt_raw;
        a.nb_count[i] = 0;   // ready for the next step's tile pass
        const double ax = fx / mi, ay = fy / mi, az = fz / mi;
        const double bx = tx / Ii, by = ty / Ii, bz = tz / Ii;
        const double dt = a.P.dt;
        double vx, vy, vz, wx, wy, wz;
        if (a.kick) {
            vx = hx0 + .5 * ax * dt;
            vy = hy0 + .5 * ay * dt;
            vz = hz0 + .5 * az * dt;
            wx = gx0 + .5 * bx * dt;
            wy = gy0 + .5 * by * dt;
            wz = gz0 + .5 * bz * dt;
        } else {
            vx = vx0; vy = vy0; vz = vz0;
            wx = wx0; wy = wy0; wz = wz0;
        }
        const double px = pi.x, py = pi.y, pz = pi.z;
        if (!(isfinite(vx) && isfinite(vy) && isfinite(vz) && isfinite(wx) && isfinite(wy) &&
              isfinite(wz) && isfinite(px) && isfinite(py) && isfinite(pz)))
            atomicOr(a.nonfinite, 1);
        if (a.write) {
            const D3 v = d3(vx, vy, vz), w = d3(wx, wy, wz), p = d3(px, py, pz);
            e[0] = .5 * mi * dot(v, v) + .5 * Ii * dot(w, w);
            const D3 mv = v * mi;
            e[1] = mv.x; e[2] = mv.y; e[3] = mv.z;
            const D3 L = cross(p, v) * mi + w * Ii;
            e[4] = L.x; e[5] = L.y; e[6] = L.z;
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
            o.f[PX][i] = px + hx * dt;
            o.f[PY][i] = py + hy * dt;
            o.f[PZ][i] = pz + hz * dt;
        } else {
            o.f[HX][i] = hx0; o.f[HY][i] = hy0; o.f[HZ][i] = hz0;
            o.f[GX][i] = gx0; o.f[GY][i] = gy0; o.f[GZ][i] = gz0;
            o.f[QX][i] = qx0; o.f[QY][i] = qy0; o.f[QZ][i] = qz0;
            o.f[PX][i] = px; o.f[PY][i] = py; o.f[PZ][i] = pz;
        }
    }
    if (a.write) {
        const double tot = block_sum<256>(pe, red);
        if (threadIdx.x == 0) a.pe_part[blockIdx.x] = tot;
        for (int k = 0; k < 7; k++) {
            const double s = block_sum<256>(e[k], red);
            if (threadIdx.x == 0) a.e_part[blockIdx.x * 7 + k] = s;
        }
    }
    // per-block integer counts, one atomic each
    __shared__ unsigned cc, cn;
    if (threadIdx.x == 0) {
        cc = 0;
        cn = 0;
    }
    __syncthreads();
    if (ncontact) atomicAdd(&cc, ncontact);
    if (nnear) atomicAdd(&cn, nnear);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (cc) {
            atomicAdd(&a.contacts[0], (unsigned long long)cc);
            atomicAdd(&a.contacts[1], (unsigned long long)cc);
        }
        if (cn) atomicAdd(a.near_total, (unsigned long long)cn);
    }
}

}  // namespace fastk
}  // namespace sl
