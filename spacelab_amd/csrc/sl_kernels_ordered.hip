// Ordered-mode kernels: bitwise the reference's single-core arithmetic.
//
// The reference adds pair contributions into acc[i] and aacc[i] in ascending
// partner index j (its triangular loop visits (A=i, B<i) first, then (A>i, B=i),
// ColliderSingleCore.cpp:100-103).  Floating-point addition is not associative,
// so the GPU reproduces that order with one sequential accumulator per (ball,
// component), while the expensive part -- evaluating the pair -- runs on every
// producer lane of the workgroup (k_pairs_ordered below: BI balls per workgroup,
// partners in chunks, evaluation and summation of consecutive chunks overlapped).
// Far pairs contribute +0.0 to aacc and j == i contributes +0.0 to both; an
// accumulator that starts at +0.0 under round-to-nearest never becomes -0.0,
// so these additions leave it bit-for-bit unchanged and are skipped.
//
// PE is summed per ball A over its partners B < A in ascending B, and the per-ball sums
// over A in ascending order by the caller: the same for every launch shape, but grouped
// per ball rather than the reference's single running sum, so it agrees to rounding only.
#include "sl_common.hpp"
#include "sl_pair_ref.hpp"

#include <type_traits>

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

// One touching pair (i, j) in the reference's contact arm (ColliderSingleCore.cpp:119-204),
// in its orientation A = max(i, j): ball i's acc contribution to c3[0], c3[cs], c3[2 cs], its
// aacc contribution likewise to a3, and the pair energy (write steps, i = A) to *pe.  The
// state is one array of fields, field f at base + f stride (how every caller lays it out),
// passed as its base so that the caller has no field table to spill for the call.
__device__ __forceinline__ void contact_pair_body(const double* base, long stride, const Phys& P, int write, int i,
                                                  int j, double ymi, double* c3, double* a3, int cs, double* pe)
{
    State s;
    for (int f = 0; f < NFIELD; f++) s.f[f] = const_cast<double*>(base + f * stride);
    const bool iA = i > j;
    const D3 pi = d3(s.f[PX][i], s.f[PY][i], s.f[PZ][i]);
    const D3 pj = d3(s.f[PX][j], s.f[PY][j], s.f[PZ][j]);
    const double Ri = s.f[RAD][i], mi = s.f[MASS][i], Ii = s.f[MOI][i];
    const double Rj = s.f[RAD][j], mj = s.f[MASS][j];
    PairIn in;
    in.pA = iA ? pi : pj;
    in.pB = iA ? pj : pi;
    in.RA = iA ? Ri : Rj;
    in.RB = iA ? Rj : Ri;
    in.mA = iA ? mi : mj;
    in.mB = iA ? mj : mi;
    PairOut o;
    GlobalContact gcx{s, iA ? i : j, iA ? j : i};
    pair_ref(P, write && iA, in, gcx, o);
    // acc[A] += F/m[A] or acc[B] -= F/m[B]; aacc[A] += tA/moi[A], aacc[B] += tB/moi[B] (:187-188, 249-250)
    D3 a = d3(div_rcp(o.F.x, mi, ymi), div_rcp(o.F.y, mi, ymi), div_rcp(o.F.z, mi, ymi));
    if (!iA) a = -a;
    const D3 aa = (iA ? o.tA : o.tB) / Ii;
    c3[0] = a.x;
    c3[cs] = a.y;
    c3[2 * cs] = a.z;
    a3[0] = aa.x;
    a3[cs] = aa.y;
    a3[2 * cs] = aa.z;
    *pe = iA ? o.pe : 0.0;
}

// A partner's position, radius and mass as the far-field arm reads them (48 bytes: three
// 16-byte LDS reads).
struct PartnerRec {
    double x, y, z, R, m, pad;
};

// One workgroup owns BI balls and sweeps every partner j in chunks of CJ.  NPW wavefronts
// evaluate pairs (producers); the running sums are kept by consumer lanes, one per (ball,
// quantity): acc x, y, z, aacc x, y, z and, on write steps, the pair energy of the ball as
// A.  With DED the consumer lanes are a wavefront of their own; otherwise they are the first
// 7 BI lanes of wavefront 0, which then also produces.
//   produce (chunk k): thread (il, jq) evaluates the pairs (i0 + il, j0 + jq + JQ q) in the
//     reference orientation and writes ball i's signed acc contribution (for a touching
//     pair also its aacc contribution and a bit of the chunk's contact mask; on write
//     steps the pair energy when i is A) to LDS buffer k mod 2;
//   consume (chunk k - 1, at the same time): each consumer lane adds its ball's CJ
//     contributions of its quantity in ascending j (aacc: only at the masked positions).
// One barrier per chunk separates the two buffers' roles.  Every sum has the same order
// whatever BI, NPW, CJ and DED are, so all launch shapes give the same bits (the energies
// included: per-ball sums, added over the balls in ascending order by the caller).
// The far arm divides by dist and by m_i through correctly rounded reciprocals
// (div_rcp): the same quotients as the reference's divisions, for five FMA-pipe
// instructions each instead of a full division.
__device__ __attribute__((noinline)) void contact_pair(const double* base, long stride, const Phys& P, int write,
                                                       int i, int j, double ymi, double* c3, double* a3, int cs,
                                                       double* pe)
{
    contact_pair_body(base, stride, P, write, i, j, ymi, c3, a3, cs, pe);
}

// CALL: the contact arm out of line (its registers weigh only on the calls) or inline.
// WPE: the wavefronts per SIMD the register budget is held to (0: the compiler's choice).
// UNR: how many of a thread's far-field pairs per chunk the sweep unrolls (the sums' order
// is fixed by the contribution buffer, so it changes no bits).
template <int BI, int NPW, int CJ, bool DED, bool CALL, int WPE, int UNR = 1>
__global__ __launch_bounds__((NPW + (DED ? 1 : 0)) * 64) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_pairs_ordered(State s, Phys P, int n, int write, double* __restrict__ pe_ball,
                     unsigned long long* __restrict__ contacts)
{
    constexpr int NP = NPW * 64;            // producer threads
    constexpr int JQ = NP / BI;             // producer threads per ball
    constexpr int KP = CJ / JQ;             // pairs per producer thread and chunk
    constexpr int MW = (CJ + 63) / 64;      // contact-mask words per ball and chunk
    constexpr int CS = CJ * BI + 4;         // quantity stride (staggers the consumer's banks)
    constexpr int NQ = 7;                   // acc xyz, aacc xyz, pair energy
    static_assert(NP % BI == 0 && CJ % JQ == 0 && KP >= 1, "chunk shape");
    static_assert(NQ * BI <= 64, "one consumer lane per (ball, quantity)");
    constexpr int NT = (NPW + (DED ? 1 : 0)) * 64;   // threads of the workgroup
    constexpr int NL = (CJ + NT - 1) / NT;           // partner records each thread stages per chunk
    __shared__ double cb[2][NQ * CS];
    __shared__ unsigned long long cmask[2][BI][MW];
    __shared__ PartnerRec sp[2][CJ];                 // the partners of chunks k and k + 1
    __shared__ unsigned ccount;
    __shared__ double sgm[BI];                       // G m_i of the workgroup's balls

    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t / 64);
    if (t < 2 * BI * MW) (&cmask[0][0][0])[t] = 0ull;
    if (t == 0) ccount = 0;
    if (t < BI) sgm[t] = blockIdx.x * BI + t < n ? P.G * s.f[MASS][blockIdx.x * BI + t] : 0.0;

    // producer role
    const int il = t % BI;
    const int jq = (t / BI) % JQ;
    const int i = blockIdx.x * BI + il;
    const bool iv = w < NPW && i < n;
    const int ic = i < n ? i : 0;
    const D3 pi = d3(s.f[PX][ic], s.f[PY][ic], s.f[PZ][ic]);
    const double Ri = s.f[RAD][ic], mi = s.f[MASS][ic];
    const double ymi = 1.0 / mi;             // divisions by m_i through div_rcp
    // ball i's factors of the reference's products when it is A (2 Ra, 4 Ra, Ha/6*64 Ra^3, G mA)
    const double two_Ri = 2 * Ri, four_Ri = 4 * Ri;
    const double Hc = P.Ha / 6 * 64;
    const double prefA = Hc * Ri * Ri * Ri;
    const double Gmi = P.G * mi;
    unsigned cnt = 0;

    // consumer role: lane c of the consumer wavefront, (ball c % BI, quantity c / BI)
    constexpr int CW = DED ? NPW : 0;       // the consumer wavefront
    const int c = t - CW * 64;
    const int cl = c % BI, cq = c / BI;
    const bool consumer = w == CW && c < NQ * BI;
    double run = 0.0;

    // Partner records are staged in LDS a chunk ahead: the loads of chunk k + 1 are issued
    // before chunk k is evaluated and stored after it, so their latency is hidden.
    const int nch = (n + CJ - 1) / CJ;
    PartnerRec nx[NL];
    auto load_chunk = [&](int kc) {
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int jl = t + u * NT;
            const int j = kc * CJ + jl;
            if (jl < CJ && j < n)
                nx[u] = PartnerRec{s.f[PX][j], s.f[PY][j], s.f[PZ][j], s.f[RAD][j], s.f[MASS][j], 0.0};
        }
    };
    auto store_chunk = [&](int kc) {
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int jl = t + u * NT;
            if (jl < CJ && kc * CJ + jl < n) sp[kc & 1][jl] = nx[u];
        }
    };
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int k = 0; k <= nch; k++) {
        const int b = k & 1;
        if (k + 1 < nch) load_chunk(k + 1);
        if (w == CW && k > 0) {
            // sums of chunk k - 1
            const int bb = b ^ 1;
            const int m = (n - (k - 1) * CJ) < CJ ? (n - (k - 1) * CJ) : CJ;
            if (consumer) {
                const double* src = &cb[bb][cq * CS + cl];
                if (cq < 3 || (cq == 6 && write)) {
                    int jl = 0;
                    for (; jl + 8 <= m; jl += 8) {
                        double v[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) v[u] = src[(jl + u) * BI];
#pragma unroll
                        for (int u = 0; u < 8; u++) run = run + v[u];
                    }
                    for (; jl < m; jl++) run = run + src[jl * BI];
                } else if (cq < 6) {
                    // aacc: only touching pairs contribute (the others add +0.0, a no-op on
                    // a sum that starts at +0.0), in ascending j
                    for (int wd = 0; wd < MW; wd++) {
                        unsigned long long mk = cmask[bb][cl][wd];
                        while (mk) {
                            const int bit = __ffsll((long long)mk) - 1;
                            mk &= mk - 1;
                            run = run + src[(wd * 64 + bit) * BI];
                        }
                    }
                }
            }
            // every lane of this wavefront has read the masks: clear them for chunk k + 1
            __builtin_amdgcn_wave_barrier();
            if (c >= 0 && c < BI * MW) (&cmask[bb][0][0])[c] = 0ull;
        }
        if (w < NPW && k < nch) {
            const int j0 = k * CJ;
            const int i0 = blockIdx.x * BI;
            unsigned pend = 0;   // pairs q of this chunk that touch: evaluated after the sweep
            // One far-field pair (i, j) in the reference's orientation A = max(i, j): ball i
            // is A when IA (ColliderSingleCore.cpp:104-110, 211-227).  Writes ball i's acc
            // contribution (and its energy on write steps), or marks the pair touching.
            auto far_pair = [&](auto ia, int q) {
                constexpr bool IA = decltype(ia)::value;
                const int jl = jq + JQ * q;
                const int j = j0 + jl;
                double c0 = 0, c1 = 0, c2 = 0, u = 0;
                if (iv && j < n && j != i) {
                    const PartnerRec& pr = sp[b][jl];
                    const D3 pj = d3(pr.x, pr.y, pr.z);
                    const double Rj = pr.R, mj = pr.m;
                    const double S = IA ? Ri + Rj : Rj + Ri;
                    const D3 r = IA ? pj - pi : pi - pj;               // x_B - x_A
                    const double dist = norm(r);
                    const double overlap = S - dist;
                    if (overlap > 0) {
                        pend |= 1u << q;
                    } else {
                        double h = ::fabs(overlap);
                        if (h < P.h_min) h = P.h_min;
                        const double h2 = h * h;
                        // vdw_factor with Ra = R_A, Rb = R_B: 2 Ra h, 2 Rb h, 4 Ra Rb and
                        // Ha/6*64 Ra Ra Ra Rb Rb Rb in the reference's product order
                        const double twoRah = IA ? two_Ri * h : (2 * Rj) * h;
                        const double twoRbh = IA ? (2 * Rj) * h : two_Ri * h;
                        const double sv = h2 + twoRah + twoRbh;
                        const double tv = sv + (IA ? four_Ri * Rj : (4 * Rj) * Ri);
                        const double quot = (IA ? h + Ri + Rj : h + Rj + Ri) / (sv * sv * tv * tv);
                        const double fv = IA ? prefA * Rj * Rj * Rj * quot : Hc * Rj * Rj * Rj * Ri * Ri * Ri * quot;
                        const double yd = 1.0 / dist;
                        const D3 rh = d3(div_rcp(r.x, dist, yd), div_rcp(r.y, dist, yd), div_rcp(r.z, dist, yd));
                        D3 F = rh * fv;
                        if (P.far_gravity) {
                            // Under a register cap G m_i is read from LDS at each use (an atomic
                            // load, which the compiler keeps in the loop): held in a register it
                            // is spilled, and the spill's reload waits for every load in flight,
                            // the next chunk's partners included.
                            double gmi = Gmi;
                            if constexpr (WPE > 0)
                                gmi = __hip_atomic_load(&sgm[il], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            F = F + rh * ((IA ? gmi * mj : P.G * mj * mi) / (dist * dist));
                        }
                        // acc[A] += F/m[A]  or  acc[B] -= F/m[B]   (:187-188, 249-250)
                        c0 = div_rcp(F.x, mi, ymi);
                        c1 = div_rcp(F.y, mi, ymi);
                        c2 = div_rcp(F.z, mi, ymi);
                        if (!IA) {
                            c0 = -c0;
                            c1 = -c1;
                            c2 = -c2;
                        } else if (write) {                            // :228-240
                            u = vdw_energy(P.Ha, Ri, Rj, S, h);
                            if (P.far_gravity) u = u + -P.G * mi * mj / dist;
                        }
                    }
                }
                const int o_ = jl * BI + il;
                cb[b][0 * CS + o_] = c0;
                cb[b][1 * CS + o_] = c1;
                cb[b][2 * CS + o_] = c2;
                if (write) cb[b][6 * CS + o_] = u;
            };
            if (j0 + CJ <= i0) {
                // every partner of the chunk precedes every ball of the block: i is A
#pragma unroll UNR
                for (int q = 0; q < KP; q++) far_pair(std::true_type{}, q);
            } else if (j0 >= i0 + BI) {
                // every partner follows: i is B
#pragma unroll UNR
                for (int q = 0; q < KP; q++) far_pair(std::false_type{}, q);
            } else {
#pragma unroll 1
                for (int q = 0; q < KP; q++) {
                    if (i > j0 + jq + JQ * q)
                        far_pair(std::true_type{}, q);
                    else
                        far_pair(std::false_type{}, q);
                }
            }
            // touching pairs: the reference's contact arm in full (:119-204), out of line so
            // that its registers do not weigh on the far-field sweep
            while (pend) {
                const int q = __ffs(pend) - 1;
                pend &= pend - 1;
                const int jl = jq + JQ * q;
                const int o_ = jl * BI + il;
                double pc = 0;
                if constexpr (CALL)
                    contact_pair(s.f[0], s.f[1] - s.f[0], P, write, i, j0 + jl, ymi, &cb[b][o_], &cb[b][3 * CS + o_],
                                 CS, &pc);
                else
                    contact_pair_body(s.f[0], s.f[1] - s.f[0], P, write, i, j0 + jl, ymi, &cb[b][o_],
                                      &cb[b][3 * CS + o_], CS, &pc);
                if (write) cb[b][6 * CS + o_] = pc;
                atomicOr(&cmask[b][il][jl >> 6], 1ull << (jl & 63));
                if (i > j0 + jl) cnt++;
            }
        }
        // chunk k + 1's partners into the buffer chunk k - 1 used (all reads of it ended
        // before the previous barrier)
        if (k + 1 < nch) store_chunk(k + 1);
        __syncthreads();
    }

    if (consumer) {
        const int ib = blockIdx.x * BI + cl;
        if (ib < n) {
            if (cq < 3)
                s.f[AX + cq][ib] = run;
            else if (cq < 6)
                s.f[BX + cq - 3][ib] = run;
            else if (write)
                pe_ball[ib] = run;
        }
    }
    // contact count (integer, order-independent)
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

// The closing kick of a step without energies (k_kick with kick = 1, write = 0) and the
// next step's drift (k_drift) in one pass over the balls: the velocities the kick writes
// are the ones the drift reads, and the accelerations are the same pair pass's, so every
// value is computed by the same operations as in the two launches.
__global__ void k_kick_drift(State s, int n, double dt, int* __restrict__ nonfinite)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const double ax = s.f[AX][b], ay = s.f[AY][b], az = s.f[AZ][b];
    const double bx = s.f[BX][b], by = s.f[BY][b], bz = s.f[BZ][b];
    // kick (ColliderSingleCore.cpp:263-287)
    const double vx = s.f[HX][b] + .5 * ax * dt;
    const double vy = s.f[HY][b] + .5 * ay * dt;
    const double vz = s.f[HZ][b] + .5 * az * dt;
    const double wx = s.f[GX][b] + .5 * bx * dt;
    const double wy = s.f[GY][b] + .5 * by * dt;
    const double wz = s.f[GZ][b] + .5 * bz * dt;
    s.f[VX][b] = vx;
    s.f[VY][b] = vy;
    s.f[VZ][b] = vz;
    s.f[WX][b] = wx;
    s.f[WY][b] = wy;
    s.f[WZ][b] = wz;
    const double px = s.f[PX][b], py = s.f[PY][b], pz = s.f[PZ][b];
    if (!(isfinite(vx) && isfinite(vy) && isfinite(vz) && isfinite(wx) && isfinite(wy) &&
          isfinite(wz) && isfinite(px) && isfinite(py) && isfinite(pz)))
        atomicOr(nonfinite, 1);
    // drift of the next step (ColliderSingleCore.cpp:82-97)
    s.f[QX][b] = px;
    s.f[QY][b] = py;
    s.f[QZ][b] = pz;
    const double hx = vx + .5 * ax * dt;
    const double hy = vy + .5 * ay * dt;
    const double hz = vz + .5 * az * dt;
    s.f[HX][b] = hx;
    s.f[HY][b] = hy;
    s.f[HZ][b] = hz;
    s.f[GX][b] = wx + .5 * bx * dt;
    s.f[GY][b] = wy + .5 * by * dt;
    s.f[GZ][b] = wz + .5 * bz * dt;
    s.f[PX][b] = px + hx * dt;
    s.f[PY][b] = py + hy * dt;
    s.f[PZ][b] = pz + hz * dt;
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

// Launch shapes of the ordered pair pass: {balls per workgroup, producer wavefronts, partners
// per chunk, consumer wavefront of its own}.  All give the same results bit for bit.
namespace {
struct OrdShape {
    int bi, npw, cj, ded;
};
constexpr OrdShape ORD_SHAPES[] = {
    {2, 3, 192, 1}, {4, 3, 48, 1}, {4, 4, 64, 0}, {8, 4, 32, 0}, {2, 2, 128, 0},
    {4, 3, 96, 1},  {2, 2, 128, 0}, {2, 4, 128, 0}, {2, 4, 128, 0},
};
constexpr int N_ORD_SHAPES = (int)(sizeof(ORD_SHAPES) / sizeof(ORD_SHAPES[0]));
}  // namespace

int ordered_shapes() { return N_ORD_SHAPES; }

// The shape used when none is chosen (the allocation-free entry point, and until the
// tuner has run), from timings on MI355X, both with a consumer wavefront of their own and
// registers held to 3 wavefronts per SIMD: up to about 1500 balls 2 balls per workgroup and
// chunks of 192 partners; above, 4 balls and chunks of 96 (the fastest shape from 2,200 to
// 65,536 balls).
int ordered_default_shape(int n) { return n < 1500 ? 0 : 5; }

int ordered_blocks(int n, int shape)
{
    const int bi = ORD_SHAPES[(shape >= 0 && shape < N_ORD_SHAPES) ? shape : 0].bi;
    return (n + bi - 1) / bi;
}

hipError_t launch_pairs_ordered(const State& s, const Phys& P, int n, int write, double* pe_ball,
                                unsigned long long* contacts, hipStream_t st, int shape)
{
    if (n <= 1) return hipSuccess;
    if (shape < 0 || shape >= N_ORD_SHAPES) shape = 0;
    const int nb = ordered_blocks(n, shape);
    // template arguments: balls, producer wavefronts, partners per chunk, dedicated consumer
    // wavefront, contact arm out of line, wavefronts per SIMD the registers are held to,
    // unroll of the far-field loop (case 5: both of a thread's pairs per chunk, 1% faster at
    // 5,000 balls than one at a time)
    switch (shape) {
    case 0: k_pairs_ordered<2, 3, 192, true, false, 3><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 1: k_pairs_ordered<4, 3, 48, true, false, 3><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 2: k_pairs_ordered<4, 4, 64, false, false, 3><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 3: k_pairs_ordered<8, 4, 32, false, false, 3><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 4: k_pairs_ordered<2, 2, 128, false, false, 0><<<nb, 128, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 5: k_pairs_ordered<4, 3, 96, true, false, 3, 2><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 6: k_pairs_ordered<2, 2, 128, false, true, 0><<<nb, 128, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    case 7: k_pairs_ordered<2, 4, 128, false, false, 3><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    default: k_pairs_ordered<2, 4, 128, false, false, 0><<<nb, 256, 0, st>>>(s, P, n, write, pe_ball, contacts); break;
    }
    return hipGetLastError();
}

int kick_blocks(int n) { return (n + 255) / 256; }

hipError_t launch_kick_drift(const State& s, int n, double dt, int* nonfinite, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_kick_drift<<<kick_blocks(n), 256, 0, st>>>(s, n, dt, nonfinite);
    return hipGetLastError();
}

hipError_t launch_kick(const State& s, int n, double dt, int write, int kick, double* e_part,
                       int* nonfinite, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_kick<<<kick_blocks(n), 256, 0, st>>>(s, n, dt, write, kick, e_part, nonfinite);
    return hipGetLastError();
}

}  // namespace sl
