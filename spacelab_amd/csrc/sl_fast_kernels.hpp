// Device kernels of the fast path (included by sl_kernels_fast.hip).
//
//   k_tiles_*  SPLIT wavefronts per 64x64 tile (I, J) of the pair triangle, I >= J,
//              each sweeping 64/SPLIT of the tile's diagonals.  Lane l holds ball
//              i = 64 I + l in registers.  The 64 balls of block J are staged in
//              LDS twice (slots 0..63 and 64..127), so at sweep step s lane l reads
//              slot l + s -- ball j = (l + s) mod 64 -- at base + s*stride, with no
//              index arithmetic.  The reaction force on j is accumulated in a
//              register that travels with j: after every step the accumulators
//              rotate one lane down (DPP wave_rol:1), so lane l always holds the
//              sum of the j it meets next; no LDS atomics or cross-lane reductions.
//              A pair closer than S + h_min (touching, or in the band where the
//              reference clamps the vdW gap) contributes nothing to the sweep; it is
//              recorded as a bit of a per-lane mask, and after the sweep each lane
//              evaluates its near pairs once with the reference arithmetic and
//              files the result (partner, force, torque, energy) in both balls'
//              near-entry lists.  The SPLIT sweeps of a tile are combined in LDS
//              in a fixed order, and each tile writes its i-side and j-side force
//              sums once, to part[J][.][i] and part[I][.][j].
//   k_ball     eight lanes per ball: sum the tile partials (fixed order; Kahan in
//              the mixed mode), sort the ball's near entries by partner index and
//              sum them in that order, then one lane forms acc and aacc, applies
//              the closing kick, the write-step energies and the next step's
//              drift (writing the other state set).
// Every sum has a fixed order, so results are reproducible run to run.
#pragma once

#include "sl_fast_pair.hpp"
#include "sl_internal.hpp"
#include "sl_pair_ref.hpp"

#pragma clang fp contract(off)

namespace sl {
namespace fastk {

constexpr int TB = FAST_TILE;
constexpr int NW = 4;                 // wavefronts per block
constexpr int NT = TB * NW;
constexpr int MAXNB = FAST_MAX_NEAR;

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

struct TileArgs {
    const double* px;
    const double* py;
    const double* pz;
    const double* R;
    FastAux aux;
    State cur;              // full state, read by the near-pair evaluation
    const Phys* sysP;       // per system
    int n_pad;
    int n_tiles;
    const int4* tiles;      // {I, J, system, first block of the system}
    double L0inv2;          // 1 / L0^2 (mixed)
    double Kv, Kg;          // mixed: force = Kv * vdw_sum + Kg * grav_sum
    double* part;
    int* nb_count;
    NearEntry* nb_ent;
    int* flags;
    double* pe_tile;
};

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

__device__ __forceinline__ void file_entry(const TileArgs& a, int self, int other, int contact,
                                           D3 F, D3 T, double pe)
{
    const int k = atomicAdd(&a.nb_count[self], 1);
    if (k < MAXNB) {
        NearEntry e;
        e.j = other;
        e.contact = contact;
        e.fx = F.x; e.fy = F.y; e.fz = F.z;
        e.tx = T.x; e.ty = T.y; e.tz = T.z;
        e.pe = pe;
        a.nb_ent[(size_t)self * MAXNB + k] = e;
    } else {
        atomicOr(&a.flags[0], 1);
    }
}

// Near pairs of this lane (bit s of mask: the pair met at step s): evaluate each
// once, in the reference's orientation (A = max index) and arithmetic, and file
// it with both balls.  The pair energy goes with A only, so it is counted once.
__device__ void eval_near(const TileArgs& a, const Phys& P, unsigned long long mask, int i, int J, int write)
{
    const int l = threadIdx.x % TB;
    const State& c = a.cur;
    while (mask) {
        const int s = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int j = J * TB + ((l + s) & (TB - 1));
        const int A = i > j ? i : j, B = i > j ? j : i;
        PairIn in;
        in.pA = d3(c.f[PX][A], c.f[PY][A], c.f[PZ][A]);
        in.pB = d3(c.f[PX][B], c.f[PY][B], c.f[PZ][B]);
        in.RA = c.f[RAD][A];
        in.RB = c.f[RAD][B];
        in.mA = c.f[MASS][A];
        in.mB = c.f[MASS][B];
        PairOut o;
        NearStateG ns{c, A, B};
        pair_ref(P, write != 0, in, ns, o);
        file_entry(a, A, B, o.contact, o.F, o.tA, o.pe);
        file_entry(a, B, A, o.contact ? 2 : 0, -o.F, o.tB, 0.0);
    }
}

// Sum over the 64 lanes of a wavefront as a fixed tree (deterministic).
__device__ __forceinline__ double wave_sum(double v)
{
    for (int m = 1; m < TB; m <<= 1) v = v + __shfl_xor(v, m, TB);
    return v;
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

struct JRec64 {
    double x, y, z, R, q, g;
};

template <bool WRITE, bool GRAV, int SPLIT>
__global__ __launch_bounds__(NT) void k_tiles_f64(TileArgs a)
{
    constexpr int TPB = NW / SPLIT;   // tiles per block
    __shared__ JRec64 sj[TPB][2 * TB];
    __shared__ double comb[NW][6][TB];
    __shared__ double pe_w[NW];

    const int w = threadIdx.x / TB, l = threadIdx.x % TB;
    const int tl = w / SPLIT, h = w % SPLIT;
    const int t = blockIdx.x * TPB + tl;
    const bool live = t < a.n_tiles;
    int I = 0, J = 0, sys = 0, blk0 = 0;
    if (live) {
        const int4 tt = a.tiles[t];
        I = tt.x;
        J = tt.y;
        sys = tt.z;
        blk0 = tt.w;
    }
    const Phys& P = a.sysP[sys];
    const int i = I * TB + l;
    const int j0 = J * TB + l;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i], Ri = a.R[i];
    const double qi = a.aux.q[i], gi = a.aux.g[i];
    const double Ri4 = 4 * Ri;
    if (h == 0) {
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
    __syncthreads();

    const double h_min = P.h_min;
    const double hmin2 = h_min * h_min;
    const double two_hmin = 2 * h_min;
    double ax = 0, ay = 0, az = 0;      // force on i
    double bx = 0, by = 0, bz = 0;      // reaction on the j this lane meets next
    double pe = 0;
    const bool diag = I == J;
    const JRec64* rec = &sj[tl][l];
    unsigned long long nmask = 0;       // bit s: near pair met at step s

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
            double u = far_energy(P.Ha, Ri, Rj, sv, tv);
            if (GRAV) u = u - (gi * r.g) * rinv;
            pe = pe + u;
        }
        nmask |= (act && near) ? (1ull << s) : 0ull;
    };

    if (live) {
        int s0, s1;
        sweep_range<SPLIT>(diag, h, s0, s1);
        if (!diag) {
#pragma unroll (WRITE ? 1 : 8)
            for (int s = s0; s < s1; s++) pair(s, true);
        } else {
#pragma unroll (WRITE ? 1 : 4)
            for (int s = s0; s < s1 - 1; s++) pair(s, true);
            pair(s1 - 1, s1 - 1 < TB / 2 || l < TB / 2);
        }
        // lane l holds the reaction sum of slot l + s1: bring slot l home
        const int back = (TB - (s1 & (TB - 1))) & (TB - 1);
        if (back) {
            bx = rotk(bx, back);
            by = rotk(by, back);
            bz = rotk(bz, back);
        }
        if (__ballot(nmask != 0)) eval_near(a, P, nmask, i, J, WRITE ? 1 : 0);
        if (WRITE) {
            const double ws = wave_sum(pe);
            if (l == 0) pe_w[w] = ws;
        }
        comb[w][0][l] = ax;
        comb[w][1][l] = ay;
        comb[w][2][l] = az;
        comb[w][3][l] = bx;
        comb[w][4][l] = by;
        comb[w][5][l] = bz;
    }
    __syncthreads();
    if (live && h == 0) {
        // the SPLIT sweeps of this tile, in a fixed order
        const int w0 = tl * SPLIT;
        double v[6];
        for (int k = 0; k < 6; k++) {
            double acc = comb[w0][k][l];
            for (int q = 1; q < SPLIT; q++) acc = acc + comb[w0 + q][k][l];
            v[k] = acc;
        }
        const size_t np = (size_t)a.n_pad;
        if (diag) {
            // the i side and the j side of the diagonal tile land on the same balls
            a.part[((size_t)(J - blk0) * 3 + 0) * np + i] = v[0] + v[3];
            a.part[((size_t)(J - blk0) * 3 + 1) * np + i] = v[1] + v[4];
            a.part[((size_t)(J - blk0) * 3 + 2) * np + i] = v[2] + v[5];
        } else {
            a.part[((size_t)(J - blk0) * 3 + 0) * np + i] = v[0];
            a.part[((size_t)(J - blk0) * 3 + 1) * np + i] = v[1];
            a.part[((size_t)(J - blk0) * 3 + 2) * np + i] = v[2];
            a.part[((size_t)(I - blk0) * 3 + 0) * np + j0] = v[3];
            a.part[((size_t)(I - blk0) * 3 + 1) * np + j0] = v[4];
            a.part[((size_t)(I - blk0) * 3 + 2) * np + j0] = v[5];
        }
    }
    if (WRITE && live && h == 0 && l == 0) {
        double pt = pe_w[tl * SPLIT];
        for (int q = 1; q < SPLIT; q++) pt = pt + pe_w[tl * SPLIT + q];
        a.pe_tile[t] = pt;
    }
}

struct JRec32 {
    double x, y, z, R;
    float r, q, g, pad;
};

// Mixed precision: relative displacement, d^2, the near test and s = d^2 - S^2
// in binary64; the vdW and gravity factors and the per-tile force sums in
// binary32 (scaled units; separate sums for vdW and gravity, whose ratio can
// reach 1e-20); tile sums converted to binary64 forces when combined.
template <bool WRITE, bool GRAV, int SPLIT>
__global__ __launch_bounds__(NT) void k_tiles_mixed(TileArgs a)
{
    constexpr int TPB = NW / SPLIT;
    __shared__ JRec32 sj[TPB][2 * TB];
    __shared__ double comb[NW][6][TB];
    __shared__ double pe_w[NW];

    const int w = threadIdx.x / TB, l = threadIdx.x % TB;
    const int tl = w / SPLIT, h = w % SPLIT;
    const int t = blockIdx.x * TPB + tl;
    const bool live = t < a.n_tiles;
    int I = 0, J = 0, sys = 0, blk0 = 0;
    if (live) {
        const int4 tt = a.tiles[t];
        I = tt.x;
        J = tt.y;
        sys = tt.z;
        blk0 = tt.w;
    }
    const Phys& P = a.sysP[sys];
    const int i = I * TB + l;
    const int j0 = J * TB + l;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i], Ri = a.R[i];
    const float r4i = 4.0f * a.aux.r32[i], qi = a.aux.q32[i], gi = a.aux.g32[i];
    const double gi64 = a.aux.g[i];
    if (h == 0) {
        JRec32 r;
        r.x = a.px[j0];
        r.y = a.py[j0];
        r.z = a.pz[j0];
        r.R = a.R[j0];
        r.r = a.aux.r32[j0];
        r.q = a.aux.q32[j0];
        r.g = a.aux.g32[j0];
        r.pad = 0.f;
        sj[tl][l] = r;
        sj[tl][l + TB] = r;
    }
    __syncthreads();

    const double h_min = P.h_min;
    const double hmin2 = h_min * h_min;
    const double two_hmin = 2 * h_min;
    const double L0inv2 = a.L0inv2;
    float vx = 0.f, vy = 0.f, vz = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;       // on i
    float wvx = 0.f, wvy = 0.f, wvz = 0.f, wgx = 0.f, wgy = 0.f, wgz = 0.f; // travelling j sums
    double pe = 0;
    const bool diag = I == J;
    const JRec32* rec = &sj[tl][l];
    unsigned long long nmask = 0;

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
            double u = far_energy(P.Ha, Ri, Rj, sv, tv);
            if (GRAV) u = u - (gi64 * a.aux.g[J * TB + ((l + s) & (TB - 1))]) * rsq_f64(d2);
            pe = pe + u;
        }
        nmask |= (act && near) ? (1ull << s) : 0ull;
    };

    if (live) {
        int s0, s1;
        sweep_range<SPLIT>(diag, h, s0, s1);
        if (!diag) {
#pragma unroll (WRITE ? 1 : 8)
            for (int s = s0; s < s1; s++) pair(s, true);
        } else {
#pragma unroll (WRITE ? 1 : 4)
            for (int s = s0; s < s1 - 1; s++) pair(s, true);
            pair(s1 - 1, s1 - 1 < TB / 2 || l < TB / 2);
        }
        const int back = (TB - (s1 & (TB - 1))) & (TB - 1);
        if (back) {
            wvx = rotk(wvx, back); wvy = rotk(wvy, back); wvz = rotk(wvz, back);
            if (GRAV) {
                wgx = rotk(wgx, back); wgy = rotk(wgy, back); wgz = rotk(wgz, back);
            }
        }
        if (__ballot(nmask != 0)) eval_near(a, P, nmask, i, J, WRITE ? 1 : 0);
        if (WRITE) {
            const double ws = wave_sum(pe);
            if (l == 0) pe_w[w] = ws;
        }
        const double Kv = a.Kv, Kg = a.Kg;
        double v[6] = {Kv * (double)vx, Kv * (double)vy, Kv * (double)vz,
                       Kv * (double)wvx, Kv * (double)wvy, Kv * (double)wvz};
        if (GRAV) {
            v[0] = v[0] + Kg * (double)gx;
            v[1] = v[1] + Kg * (double)gy;
            v[2] = v[2] + Kg * (double)gz;
            v[3] = v[3] + Kg * (double)wgx;
            v[4] = v[4] + Kg * (double)wgy;
            v[5] = v[5] + Kg * (double)wgz;
        }
        for (int k = 0; k < 6; k++) comb[w][k][l] = v[k];
    }
    __syncthreads();
    if (live && h == 0) {
        const int w0 = tl * SPLIT;
        double v[6];
        for (int k = 0; k < 6; k++) {
            double acc = comb[w0][k][l];
            for (int q = 1; q < SPLIT; q++) acc = acc + comb[w0 + q][k][l];
            v[k] = acc;
        }
        const size_t np = (size_t)a.n_pad;
        if (diag) {
            a.part[((size_t)(J - blk0) * 3 + 0) * np + i] = v[0] + v[3];
            a.part[((size_t)(J - blk0) * 3 + 1) * np + i] = v[1] + v[4];
            a.part[((size_t)(J - blk0) * 3 + 2) * np + i] = v[2] + v[5];
        } else {
            a.part[((size_t)(J - blk0) * 3 + 0) * np + i] = v[0];
            a.part[((size_t)(J - blk0) * 3 + 1) * np + i] = v[1];
            a.part[((size_t)(J - blk0) * 3 + 2) * np + i] = v[2];
            a.part[((size_t)(I - blk0) * 3 + 0) * np + j0] = v[3];
            a.part[((size_t)(I - blk0) * 3 + 1) * np + j0] = v[4];
            a.part[((size_t)(I - blk0) * 3 + 2) * np + j0] = v[5];
        }
    }
    if (WRITE && live && h == 0 && l == 0) {
        double pt = pe_w[tl * SPLIT];
        for (int q = 1; q < SPLIT; q++) pt = pt + pe_w[tl * SPLIT + q];
        a.pe_tile[t] = pt;
    }
}

struct BallArgs {
    State cur, nxt;
    const Phys* sysP;
    const int* blk_sys;
    const int* sys_off;
    const int* sys_n;
    const int* sys_nb;
    int n_used, n_pad;
    const double* part;
    int* nb_count;
    const NearEntry* nb_ent;
    int write, kick, drift_next, compensated;
    double* e_blk;      // FAST_E_STRIDE doubles per 32-ball block (write steps)
    int* flags;         // [1]: a non-finite state appeared
    unsigned long long* counters;
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
    __shared__ int key[BPB][MAXNB];
    __shared__ int order[BPB][MAXNB];
    const int g = threadIdx.x % LPB;
    const int bl = threadIdx.x / LPB;
    const int i = blockIdx.x * BPB + bl;
    const int sys = a.blk_sys[(i < a.n_used ? i : 0) / TB];
    const bool valid = i < a.n_used && (i - a.sys_off[sys]) < a.sys_n[sys];
    const int ic = valid ? i : 0;
    const int nblk = a.sys_nb[sys];
    const State& c = a.cur;
    const size_t np = (size_t)a.n_pad;
    double e[FAST_E_STRIDE] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // PE KE p L contacts nonfinite
    unsigned ncontact = 0;

    // the ball's own state, loaded up front (the loads cannot cross the barriers below)
    const double hx0 = c.f[HX][ic], hy0 = c.f[HY][ic], hz0 = c.f[HZ][ic];
    const double gx0 = c.f[GX][ic], gy0 = c.f[GY][ic], gz0 = c.f[GZ][ic];
    const double vx0 = c.f[VX][ic], vy0 = c.f[VY][ic], vz0 = c.f[VZ][ic];
    const double wx0 = c.f[WX][ic], wy0 = c.f[WY][ic], wz0 = c.f[WZ][ic];
    const double qx0 = c.f[QX][ic], qy0 = c.f[QY][ic], qz0 = c.f[QZ][ic];
    const double px = c.f[PX][ic], py = c.f[PY][ic], pz = c.f[PZ][ic];
    const double mi = c.f[MASS][ic], Ii = c.f[MOI][ic];

    // near entries: partner keys into LDS, order by partner index (rank = keys below)
    const int cnt_raw = valid ? a.nb_count[ic] : 0;
    const int cnt = cnt_raw < MAXNB ? cnt_raw : MAXNB;
    const NearEntry* ent = a.nb_ent + (size_t)ic * MAXNB;
    for (int u = g; u < cnt; u += LPB) key[bl][u] = ent[u].j;

    // far field: lane g sums the tile partials K = g, g+8, ... of the ball's system
    double fx = 0, fy = 0, fz = 0;
    if (valid) {
        if (a.compensated) {
            double cx = 0, cy = 0, cz = 0;   // Kahan compensation
#pragma unroll 4
            for (int K = g; K < nblk; K += LPB) {
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
            for (int K = g; K < nblk; K += LPB) {
                const double* p = a.part + (size_t)K * 3 * np + ic;
                fx = fx + p[0];
                fy = fy + p[np];
                fz = fz + p[2 * np];
            }
        }
    }
    __syncthreads();
    for (int u = g; u < cnt; u += LPB) {
        const int v = key[bl][u];
        int below = 0;
        for (int x = 0; x < cnt; x++) below += key[bl][x] < v;
        order[bl][below] = u;
    }
    __syncthreads();

    // near contributions in partner order: lane g takes ranks g, g+8, ...
    double nfx = 0, nfy = 0, nfz = 0, ntx = 0, nty = 0, ntz = 0, pe = 0;
    for (int r = g; r < cnt; r += LPB) {
        const NearEntry en = ent[order[bl][r]];
        nfx = nfx + en.fx;
        nfy = nfy + en.fy;
        nfz = nfz + en.fz;
        ntx = ntx + en.tx;
        nty = nty + en.ty;
        ntz = ntz + en.tz;
        pe = pe + en.pe;
        ncontact += en.contact == 1;
    }
    fx = lane_group_sum(fx + nfx);
    fy = lane_group_sum(fy + nfy);
    fz = lane_group_sum(fz + nfz);
    const double tx = lane_group_sum(ntx), ty = lane_group_sum(nty), tz = lane_group_sum(ntz);
    pe = lane_group_sum(pe);
    ncontact = (unsigned)lane_group_sum((double)ncontact);
    unsigned nnear = 0;
    if (valid && g == 0) {
        nnear = (unsigned)cnt_raw;
        a.nb_count[i] = 0;   // ready for the next step's tile pass
        const double ax = fx / mi, ay = fy / mi, az = fz / mi;
        const double bx = tx / Ii, by = ty / Ii, bz = tz / Ii;
        const double dt = a.sysP[sys].dt;
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
        const bool bad = !(isfinite(vx) && isfinite(vy) && isfinite(vz) && isfinite(wx) && isfinite(wy) &&
                           isfinite(wz) && isfinite(px) && isfinite(py) && isfinite(pz));
        if (bad) atomicOr(&a.flags[1], 1);
        if (a.write) {
            const D3 v = d3(vx, vy, vz), w = d3(wx, wy, wz), p = d3(px, py, pz);
            e[0] = pe;
            e[1] = .5 * mi * dot(v, v) + .5 * Ii * dot(w, w);
            const D3 mv = v * mi;
            e[2] = mv.x; e[3] = mv.y; e[4] = mv.z;
            const D3 L = cross(p, v) * mi + w * Ii;
            e[5] = L.x; e[6] = L.y; e[7] = L.z;
            e[8] = (double)ncontact;
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
        // a 32-ball block lies inside one 64-aligned system, so its sums are per system
        for (int k = 0; k < FAST_E_STRIDE; k++) {
            const double s = block_sum<256>(e[k], red);
            if (threadIdx.x == 0) a.e_blk[blockIdx.x * FAST_E_STRIDE + k] = s;
        }
    }
    // profiling counters: one atomic per block each
    __shared__ unsigned cc, cn;
    if (threadIdx.x == 0) {
        cc = 0;
        cn = 0;
    }
    __syncthreads();
    if (g == 0 && ncontact) atomicAdd(&cc, ncontact);
    if (nnear) atomicAdd(&cn, nnear);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (cc) atomicAdd(&a.counters[0], (unsigned long long)cc);
        if (cn) atomicAdd(&a.counters[1], (unsigned long long)cn);
    }
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
