// Ordered-mode kernels: bitwise the reference's single-core arithmetic.
//
// The reference adds pair contributions into acc[i] and aacc[i] in ascending
// partner index j (its triangular loop visits (A=i, B<i) first, then (A>i, B=i),
// ColliderSingleCore.cpp:100-103).  Floating-point addition is not associative,
// so the GPU reproduces that order with one sequential accumulator per (ball,
// component), while the expensive part -- evaluating the pair -- runs on every
// lane of the block:
//
//   block = 256 threads, BI balls i, j swept in chunks of CJ (64 for BI = 16, else 128):
//     produce: thread (il, jq) evaluates pairs (i0+il, j0+jq+JQ*k) in the
//              reference orientation (A = max(i,j), B = min(i,j)) and writes
//              the signed contribution for ball i into LDS;
//     consume: 6*BI threads, one per (ball, component), add the CJ chunk
//              contributions in ascending j into their running sums.
//   Far pairs contribute +0.0 to aacc and j == i contributes +0.0 to both; an
//   accumulator that starts at +0.0 under round-to-nearest never becomes -0.0,
//   so these additions leave it bit-for-bit unchanged.
//
// PE is summed per thread and per block in a fixed order (deterministic, but not
// the reference's single running sum, so it agrees to rounding only).
#include "sl_common.hpp"
#include "sl_pair_ref.hpp"

#pragma clang fp contract(off)

namespace sl {

namespace {

struct GlobalContact {
    const State& s;
    int A, B;
    __device__ D3 vA() const { return d3(s.f[VX][A], s.f[VY][A], s.f[VZ][A]); }
    __device__ D3 vB() const { return d3(s.f[VX][B], s.f[VY][B], s.f[VZ][B]); }
    __device__ D3 wA() const { return d3(s.f[WX][A], s.f[WY][A], s.f[WZ][A]); }
    __device__ D3 wB() const { return d3(s.f[WX][B], s.f[WY][B], s.f[WZ][B]); }
    // distance of the pair in the previous pair pass, recomputed from the
    // positions that pass used with the same operations (|x_B - x_A|)
    __device__ double old_dist() const
    {
        const D3 qa = d3(s.f[QX][A], s.f[QY][A], s.f[QZ][A]);
        const D3 qb = d3(s.f[QX][B], s.f[QY][B], s.f[QZ][B]);
        return norm(qb - qa);
    }
};

template <int N>
__device__ __forceinline__ double block_sum(double v, double* red)
{
    // fixed-shape tree over the block: deterministic
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

}  // namespace

// First pass (ColliderSingleCore.cpp:82-97) plus the hysteresis state: the
// positions before the drift are the ones the previous pair pass used.
__global__ void k_drift(State s, int n, double dt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
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

// Positions the step-0 pass uses are also its hysteresis reference (k = kout
// for a pair overlapping at step 0).
__global__ void k_copy_prev(State s, int n)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    s.f[QX][b] = s.f[PX][b];
    s.f[QY][b] = s.f[PY][b];
    s.f[QZ][b] = s.f[PZ][b];
}

// Limited to 128 VGPRs (four wavefronts per SIMD; a 20-byte spill) for the 4- and 8-ball
// blocks: 8% faster at 5000 spheres, 6.5% at 10,000 (the 16-ball block is held to three by
// its LDS anyway).
template <int BI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_pairs_ordered(State s, Phys P, int n, int write,
                                                       double* __restrict__ pe_part,
                                                       unsigned long long* __restrict__ contacts)
{
    constexpr int NT = 256;
    constexpr int CJ = BI == 16 ? 64 : 128;   // partners per chunk (one barrier pair each)
    constexpr int JQ = NT / BI;            // threads sharing one ball i
    constexpr int CS = CJ * BI + 8;         // component stride, padded against bank conflicts
    __shared__ double cont[6 * CS];
    __shared__ double red[NT];

    const int t = threadIdx.x;
    const int il = t % BI;
    const int jq = t / BI;
    const int i = blockIdx.x * BI + il;
    const bool iv = i < n;
    const int ic = iv ? i : 0;
    const D3 pi = d3(s.f[PX][ic], s.f[PY][ic], s.f[PZ][ic]);
    const double Ri = s.f[RAD][ic], mi = s.f[MASS][ic], Ii = s.f[MOI][ic];

    // consumer role: (ball il, component cc) for t < 6*BI
    const int cc = t / BI;
    double run = 0.0;
    double pe = 0.0;
    unsigned cnt = 0;

    for (int j0 = 0; j0 < n; j0 += CJ) {
#pragma unroll 1
        for (int k = 0; k < CJ / JQ; k++) {
            const int jl = jq + JQ * k;
            const int j = j0 + jl;
            double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
            if (iv && j < n && j != i) {
                const bool iA = i > j;
                const int A = iA ? i : j, B = iA ? j : i;
                const D3 pj = d3(s.f[PX][j], s.f[PY][j], s.f[PZ][j]);
                const double Rj = s.f[RAD][j], mj = s.f[MASS][j];
                PairIn in;
                in.pA = iA ? pi : pj;
                in.pB = iA ? pj : pi;
                in.RA = iA ? Ri : Rj;
                in.RB = iA ? Rj : Ri;
                in.mA = iA ? mi : mj;
                in.mB = iA ? mj : mi;
                PairOut o;
                GlobalContact gc{s, A, B};
                pair_ref(P, write && iA, in, gc, o);
                if (iA) {
                    // acc[A] += F/m[A]; aacc[A] += tA/moi[A]   (:187, :249)
                    const D3 a = o.F / mi;
                    c0 = a.x; c1 = a.y; c2 = a.z;
                    if (o.contact) {
                        const D3 aa = o.tA / Ii;
                        c3 = aa.x; c4 = aa.y; c5 = aa.z;
                    }
                    pe = pe + o.pe;
                    cnt += o.contact;
                } else {
                    // acc[B] -= F/m[B]; aacc[B] += tB/moi[B]   (:188, :250)
                    const D3 a = o.F / mi;
                    c0 = -a.x; c1 = -a.y; c2 = -a.z;
                    if (o.contact) {
                        const D3 aa = o.tB / Ii;
                        c3 = aa.x; c4 = aa.y; c5 = aa.z;
                    }
                }
            }
            const int o_ = jl * BI + il;
            cont[0 * CS + o_] = c0;
            cont[1 * CS + o_] = c1;
            cont[2 * CS + o_] = c2;
            cont[3 * CS + o_] = c3;
            cont[4 * CS + o_] = c4;
            cont[5 * CS + o_] = c5;
        }
        __syncthreads();
        if (t < 6 * BI) {
            const double* src = cont + cc * CS + il;
            const int m = (n - j0) < CJ ? (n - j0) : CJ;
            for (int jl = 0; jl < m; jl++) run = run + src[jl * BI];
        }
        __syncthreads();
    }

    if (t < 6 * BI && iv) {
        const int fld = cc < 3 ? (AX + cc) : (BX + cc - 3);
        s.f[fld][i] = run;
    }
    if (write) {
        const double tot = block_sum<NT>(pe, red);
        if (t == 0) pe_part[blockIdx.x] = tot;
    }
    // contact count (integer, order-independent)
    __shared__ unsigned ccount;
    if (t == 0) ccount = 0;
    __syncthreads();
    if (cnt) atomicAdd(&ccount, cnt);
    __syncthreads();
    // contacts[0]: the running total; contacts[1]: this write step's touching pairs
    if (t == 0 && ccount) {
        atomicAdd(contacts, (unsigned long long)ccount);
        if (write) atomicAdd(contacts + 1, (unsigned long long)ccount);
    }
}

// Third pass (ColliderSingleCore.cpp:263-287): closing half kick and, on write
// steps, per-block partial sums of KE, momentum and angular momentum.
__global__ __launch_bounds__(256) void k_kick(State s, int n, double dt, int write, int kick,
                                              double* __restrict__ e_part,
                                              int* __restrict__ nonfinite)
{
    __shared__ double red[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    double e[7] = {0, 0, 0, 0, 0, 0, 0};
    if (b < n) {
        double vx, vy, vz, wx, wy, wz;
        if (kick) {
            vx = s.f[HX][b] + .5 * s.f[AX][b] * dt;
            vy = s.f[HY][b] + .5 * s.f[AY][b] * dt;
            vz = s.f[HZ][b] + .5 * s.f[AZ][b] * dt;
            wx = s.f[GX][b] + .5 * s.f[BX][b] * dt;
            wy = s.f[GY][b] + .5 * s.f[BY][b] * dt;
            wz = s.f[GZ][b] + .5 * s.f[BZ][b] * dt;
            s.f[VX][b] = vx;
            s.f[VY][b] = vy;
            s.f[VZ][b] = vz;
            s.f[WX][b] = wx;
            s.f[WY][b] = wy;
            s.f[WZ][b] = wz;
        } else {
            vx = s.f[VX][b]; vy = s.f[VY][b]; vz = s.f[VZ][b];
            wx = s.f[WX][b]; wy = s.f[WY][b]; wz = s.f[WZ][b];
        }
        const double px = s.f[PX][b], py = s.f[PY][b], pz = s.f[PZ][b];
        if (!(isfinite(vx) && isfinite(vy) && isfinite(vz) && isfinite(wx) && isfinite(wy) &&
              isfinite(wz) && isfinite(px) && isfinite(py) && isfinite(pz)))
            atomicOr(nonfinite, 1);
        if (write) {
            const double m = s.f[MASS][b], I = s.f[MOI][b];
            const D3 v = d3(vx, vy, vz), w = d3(wx, wy, wz), p = d3(px, py, pz);
            e[0] = .5 * m * dot(v, v) + .5 * I * dot(w, w);      // :282-283
            const D3 mv = v * m;                                   // :284
            e[1] = mv.x; e[2] = mv.y; e[3] = mv.z;
            const D3 L = cross(p, v) * m + w * I;                  // :285
            e[4] = L.x; e[5] = L.y; e[6] = L.z;
        }
    }
    if (write) {
        for (int c = 0; c < 7; c++) {
            const double tot = block_sum<256>(e[c], red);
            if (threadIdx.x == 0) e_part[blockIdx.x * 7 + c] = tot;
        }
    }
}

// ---- launchers --------------------------------------------------------------

hipError_t launch_drift(const State& s, int n, double dt, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_drift<<<(n + 255) / 256, 256, 0, st>>>(s, n, dt);
    return hipGetLastError();
}

hipError_t launch_copy_prev(const State& s, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_copy_prev<<<(n + 255) / 256, 256, 0, st>>>(s, n);
    return hipGetLastError();
}

int ordered_blocks(int n, int* bi)
{
    // the largest i-block that still gives about 940 blocks: the pass is latency-bound
    // at a few wavefronts per SIMD, and smaller blocks raise that (measured best: 4 balls
    // up to ~7,500 spheres, 8 up to ~15,000, 16 above; 343 vs 458 us per 5000-sphere step)
    const int b = n >= 16 * 940 ? 16 : n >= 8 * 940 ? 8 : 4;
    if (bi) *bi = b;
    return (n + b - 1) / b;
}

hipError_t launch_pairs_ordered(const State& s, const Phys& P, int n, int write, double* pe_part,
                                unsigned long long* contacts, hipStream_t st)
{
    if (n <= 1) return hipSuccess;
    int bi = 0;
    const int nb = ordered_blocks(n, &bi);
    if (bi == 16)
        k_pairs_ordered<16><<<nb, 256, 0, st>>>(s, P, n, write, pe_part, contacts);
    else if (bi == 8)
        k_pairs_ordered<8><<<nb, 256, 0, st>>>(s, P, n, write, pe_part, contacts);
    else
        k_pairs_ordered<4><<<nb, 256, 0, st>>>(s, P, n, write, pe_part, contacts);
    return hipGetLastError();
}

int kick_blocks(int n) { return (n + 255) / 256; }

hipError_t launch_kick(const State& s, int n, double dt, int write, int kick, double* e_part,
                       int* nonfinite, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_kick<<<kick_blocks(n), 256, 0, st>>>(s, n, dt, write, kick, e_part, nonfinite);
    return hipGetLastError();
}

}  // namespace sl
