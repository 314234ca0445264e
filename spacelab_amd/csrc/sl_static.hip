// Allocation-free stepping: the ordered (bitwise) step on state held in the caller's
// device memory, with the step's scratch -- per-ball pair energies, per-block kinetic
// partials, the touching-pair counters and the non-finite flag -- in a fixed buffer of
// this library on each device.
// Nothing here allocates or frees device memory, so the call can run where hipMalloc
// may not: under HIP graph capture, in tight serving loops, in harnesses that forbid it.
//
// The buffer is a file-scope __device__ array.  In host code its name only stands in for
// the symbol; the device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept, and only that address is handed to kernels.  Every call
// on a GPU shares that GPU's buffer and leaves its contents behind: the counters are
// cleared by kernels at the start of each call (not by hipMemsetAsync, which captured
// graphs have replayed wrongly), and calls that overlap on one GPU race on it.
#include "spacelab/sl_api.h"

#include "sl_common.hpp"
#include "sl_internal.hpp"

#include <cmath>
#include <mutex>

namespace sl {
namespace {

// Scratch layout, in doubles from the start of the buffer.  Sizes cover the largest
// system the entry point accepts, SL_STATIC_MAX_BALLS balls.
constexpr int ST_MAX = SL_STATIC_MAX_BALLS;
constexpr int ST_PE_SLOTS = ST_MAX + 64;                // pair energies, one per ball
constexpr int ST_E_SLOTS = 7 * ((ST_MAX + 255) / 256 + 8);   // kick blocks x {KE, p, L}
constexpr int ST_PE_OFF = 0;
constexpr int ST_E_OFF = ST_PE_OFF + ST_PE_SLOTS;
constexpr int ST_CNT_OFF = ST_E_OFF + ST_E_SLOTS;        // 2 x u64: contacts (running, this write step)
constexpr int ST_FLAG_OFF = ST_CNT_OFF + 2;              // int: non-finite state seen
constexpr int ST_DOUBLES = ST_FLAG_OFF + 2;               // 540 KB

// the caller's state layout, as the header documents it, is the kernels' own
static_assert(NFIELD == SL_NFIELD && PX == SL_F_POS && VX == SL_F_VEL && WX == SL_F_W && AX == SL_F_ACC &&
                  BX == SL_F_AACC && HX == SL_F_VELH && GX == SL_F_WH && QX == SL_F_POS_PREV && RAD == SL_F_R &&
                  MASS == SL_F_M && MOI == SL_F_MOI,
              "field order of sl_step_ordered_static");

// Energies of a write step from the partials, summed in the order the context path sums
// them on the host (ascending ball or block index from +0.0), so the values are the same.
__global__ void k_energy_out(const double* __restrict__ pe_part, int n_pe, const double* __restrict__ e_part,
                             int n_e, const unsigned long long* __restrict__ cnt, const int* __restrict__ flag,
                             double* __restrict__ out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < n_pe; b++) v[0] += pe_part[b];
    for (int b = 0; b < n_e; b++)
        for (int k = 0; k < 7; k++) v[1 + k] += e_part[7 * b + k];
    for (int k = 0; k < 8; k++) out[k] = v[k];
    out[8] = (double)cnt[1];
    out[9] = (double)(*flag);
}

}  // namespace
}  // namespace sl

// The scratch buffer (one copy per GPU the library runs on, held until the process ends).
static __device__ __align__(256) double scratch[sl::ST_DOUBLES];

namespace sl {
namespace {

constexpr int ST_MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[ST_MAX_DEVICES] = {};   // device address of each GPU's copy of the buffer

// Device address of the current device's buffer, looked up on the first call there.
hipError_t scratch_address(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= ST_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *out = static_cast<double*>(g_ptr[dev]);
    return hipSuccess;
}

}  // namespace
}  // namespace sl

using namespace sl;

extern "C" int sl_step_ordered_static(const sl_params* p, int n, int stride, double* dev_state,
                                      long long first_step, int n_steps, int write_every,
                                      double* dev_energy, void* stream)
{
    if (!p || !dev_state) return SL_E_ARG;
    if (n < 1 || stride < n || n_steps < 0 || first_step < 0) return SL_E_ARG;
    if (n > SL_STATIC_MAX_BALLS) return SL_E_CAPACITY;
    if (!(p->dt > 0) || !std::isfinite(p->dt) || !(p->kin >= 0) || !(p->kout >= 0) || !(p->h_min > 0))
        return SL_E_ARG;
    if (p->precision != SL_PREC_FP64) return SL_E_ARG;
    double* s = nullptr;
    if (scratch_address(&s) != hipSuccess) return SL_E_HIP;
    double* pe_part = s + ST_PE_OFF;
    double* e_part = s + ST_E_OFF;
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(s + ST_CNT_OFF);
    int* flag = reinterpret_cast<int*>(s + ST_FLAG_OFF);
    const int shape = ordered_default_shape(n);
    const int n_pe = n > 1 ? n : 0;   // per-ball pair energies
    const int n_e = kick_blocks(n);
    if (n_pe > ST_PE_SLOTS || 7 * n_e > ST_E_SLOTS) return SL_E_CAPACITY;

    hipStream_t st = static_cast<hipStream_t>(stream);
    State S;
    for (int f = 0; f < NFIELD; f++) S.f[f] = dev_state + (size_t)f * stride;
    Phys P;
    P.G = p->G;
    P.Ha = p->Ha;
    P.h_min = p->h_min;
    P.u_s = p->u_s;
    P.u_r = p->u_r;
    P.kin = p->kin;
    P.kout = p->kout;
    P.dt = p->dt;
    P.far_gravity = p->far_gravity;

    // the buffer keeps the previous call's contents: clear what must start at zero
    if (launch_zero_u64(cnt, 2, st) != hipSuccess || launch_zero_i32(flag, 1, st) != hipSuccess) return SL_E_HIP;
    int nw = 0;
    for (int k = 1; k <= n_steps; k++) {
        const long long step = first_step + k;
        const bool write = write_every > 0 && step % write_every == 0;
        if (launch_drift(S, n, p->dt, st) != hipSuccess) return SL_E_HIP;
        if (write && launch_zero_u64(cnt + 1, 1, st) != hipSuccess) return SL_E_HIP;
        if (launch_pairs_ordered(S, P, n, write ? 1 : 0, pe_part, cnt, st, shape) != hipSuccess) return SL_E_HIP;
        if (launch_kick(S, n, p->dt, write ? 1 : 0, 1, e_part, flag, st) != hipSuccess) return SL_E_HIP;
        if (write) {
            if (dev_energy) {
                k_energy_out<<<1, 1, 0, st>>>(pe_part, n_pe, e_part, n_e, cnt, flag,
                                               dev_energy + (size_t)nw * SL_STATIC_E_STRIDE);
                if (hipGetLastError() != hipSuccess) return SL_E_HIP;
            }
            nw++;
        }
    }
    return nw;
}
