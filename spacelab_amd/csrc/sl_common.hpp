// Shared host/device definitions for the spacelab_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sl {

// Per-ball fields, each a device array of n_pad doubles (structure of arrays).
// Positions, velocities, spins etc. are split per component so that a wave
// reads 64 consecutive doubles of one component per load.
enum Field : int {
    PX, PY, PZ,        // position (after this step's drift once the step has begun)
    VX, VY, VZ,        // velocity at the last full step (the reference's vel)
    WX, WY, WZ,        // spin at the last full step (w)
    AX, AY, AZ,        // acceleration (acc)
    BX, BY, BZ,        // angular acceleration (aacc)
    HX, HY, HZ,        // half-step velocity (velh)
    GX, GY, GZ,        // half-step spin (wh)
    QX, QY, QZ,        // positions used by the previous pair pass (hysteresis state)
    RAD, MASS, MOI,    // radius, mass, moment of inertia
    NFIELD
};

struct State {
    double* f[NFIELD];
};

// Physics constants of one system as the kernels read them.
struct Phys {
    double G, Ha, h_min, u_s, u_r, kin, kout, dt;
    int far_gravity;
};

// Energy partials per block: PE, KE, p.xyz, L.xyz
constexpr int NE = 8;

}  // namespace sl
