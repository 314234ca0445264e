// C ABI of batched ensembles: many independent Ball_groups advanced by the fast
// engine as one set of launches per step (sl_kernels_fast.hip).  Each system keeps
// the per-step semantics of sim_one_step (ColliderSingleCore.cpp:78-288) with its
// own kin, kout and dt.
#include "spacelab/sl_api.h"

#include "sl_common.hpp"
#include "sl_internal.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace sl;

struct sl_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    int max_sys = 0, max_balls = 0, n_pad = 0;
    sl_params p{};
    double* base = nullptr;   // NFIELD * n_pad doubles
    State st{};
    FastBuffers fb{};
    FastLayout lay;
    std::vector<Phys> P;
    std::vector<int> n_sys_balls;
    int n_total = 0;
    long long step = 0;
    bool uploaded = false;
    bool forces_valid = false;
    std::string err;
};

namespace {

int berr(sl_batch* b, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

bool positive(double x) { return x > 0 && std::isfinite(x); }

void release(sl_batch* b)
{
    fast_free(b->fb);
    if (b->base) (void)hipFree(b->base);
    b->base = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

// Per-system physics from the common parameters and optional per-system arrays.
int build_phys(sl_batch* b, int n_sys, const double* kin, const double* kout, const double* dt,
               std::vector<Phys>& out)
{
    out.assign(n_sys, Phys{});
    for (int s = 0; s < n_sys; s++) {
        Phys& P = out[s];
        const Phys* prev = (int)b->P.size() == n_sys ? &b->P[s] : nullptr;
        P.G = b->p.G;
        P.Ha = b->p.Ha;
        P.h_min = b->p.h_min;
        P.u_s = b->p.u_s;
        P.u_r = b->p.u_r;
        P.far_gravity = b->p.far_gravity;
        P.kin = kin ? kin[s] : prev ? prev->kin : b->p.kin;
        P.kout = kout ? kout[s] : prev ? prev->kout : b->p.kout;
        P.dt = dt ? dt[s] : prev ? prev->dt : b->p.dt;
        if (!positive(P.dt)) return berr(b, SL_E_ARG, "system %d: dt must be > 0 (got %g)", s, P.dt);
        if (!(P.kin >= 0) || !(P.kout >= 0) || !std::isfinite(P.kin) || !std::isfinite(P.kout))
            return berr(b, SL_E_ARG, "system %d: kin/kout must be >= 0", s);
    }
    return SL_OK;
}

}  // namespace

extern "C" {

int sl_batch_create(sl_batch** out, int max_systems, int max_balls, const sl_params* p)
{
    if (!out) return SL_E_ARG;
    *out = nullptr;
    if (!p || max_systems < 1 || max_balls < 1 || max_balls > (1 << 26) || max_systems > max_balls)
        return SL_E_ARG;
    if (p->order != SL_ORDER_FAST) return SL_E_ARG;
    if (p->precision != SL_PREC_FP64 && p->precision != SL_PREC_MIXED) return SL_E_ARG;
    if (!(p->h_min > 0)) return SL_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SL_E_NOGPU;
    sl_batch* b = new (std::nothrow) sl_batch();
    if (!b) return SL_E_ARG;
    hipDeviceProp_t prop;
    if (hipGetDevice(&b->dev) != hipSuccess || hipGetDeviceProperties(&prop, b->dev) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete b;
        return SL_E_NOGPU;
    }
    b->p = *p;
    b->fb.n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    b->max_sys = max_systems;
    b->max_balls = max_balls;
    // every system starts on a 64-slot boundary: at most 63 padding slots each
    const long long slots = (long long)max_balls + 63LL * max_systems;
    b->n_pad = (int)((slots + 255) / 256 * 256);
    auto fail = [&](int code) {
        release(b);
        delete b;
        return code;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(SL_E_HIP);
    if (hipMalloc(&b->base, sizeof(double) * (size_t)NFIELD * b->n_pad) != hipSuccess) return fail(SL_E_HIP);
    if (hipMemset(b->base, 0, sizeof(double) * (size_t)NFIELD * b->n_pad) != hipSuccess) return fail(SL_E_HIP);
    for (int f = 0; f < NFIELD; f++) b->st.f[f] = b->base + (size_t)f * b->n_pad;
    if (fast_alloc(b->fb, b->n_pad, max_systems) != 0) return fail(SL_E_HIP);
    *out = b;
    return SL_OK;
}

int sl_batch_upload(sl_batch* b, int n_sys, const int* n_per_sys, const double* pos3, const double* vel3,
                    const double* w3, const double* R, const double* m, const double* moi, const double* kin,
                    const double* kout, const double* dt)
{
    if (!b) return SL_E_ARG;
    if (n_sys < 1 || n_sys > b->max_sys) return berr(b, SL_E_ARG, "n_sys=%d outside [1, %d]", n_sys, b->max_sys);
    if (!n_per_sys || !pos3 || !vel3 || !w3 || !R || !m || !moi) return berr(b, SL_E_ARG, "required array is NULL");
    long long total = 0;
    for (int s = 0; s < n_sys; s++) {
        if (n_per_sys[s] < 1) return berr(b, SL_E_ARG, "system %d has %d balls", s, n_per_sys[s]);
        total += n_per_sys[s];
    }
    if (total > b->max_balls) return berr(b, SL_E_ARG, "%lld balls exceed the capacity %d", total, b->max_balls);
    const int n = (int)total;
    for (int i = 0; i < n; i++) {
        if (!positive(R[i]) || !positive(m[i]) || !positive(moi[i]))
            return berr(b, SL_E_ARG, "ball %d: R, m and moi must be > 0", i);
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(pos3[3 * i + k]) || !std::isfinite(vel3[3 * i + k]) || !std::isfinite(w3[3 * i + k]))
                return berr(b, SL_E_ARG, "non-finite position/velocity/spin at ball %d", i);
    }
    b->uploaded = false;
    b->forces_valid = false;
    b->P.clear();
    std::vector<Phys> P;
    int rc = build_phys(b, n_sys, kin, kout, dt, P);
    if (rc) return rc;
    const FastLayout L = fast_layout(n_sys, n_per_sys);
    const size_t np = (size_t)b->n_pad;
    std::vector<double> h((size_t)NFIELD * np, 0.0);
    int src = 0;
    for (int s = 0; s < n_sys; s++)
        for (int k = 0; k < L.n[s]; k++, src++) {
            const size_t d = (size_t)L.off[s] + k;
            for (int x = 0; x < 3; x++) {
                h[(size_t)(PX + x) * np + d] = pos3[3 * src + x];
                h[(size_t)(QX + x) * np + d] = pos3[3 * src + x];
                h[(size_t)(VX + x) * np + d] = vel3[3 * src + x];
                h[(size_t)(HX + x) * np + d] = vel3[3 * src + x];
                h[(size_t)(WX + x) * np + d] = w3[3 * src + x];
                h[(size_t)(GX + x) * np + d] = w3[3 * src + x];
            }
            h[(size_t)RAD * np + d] = R[src];
            h[(size_t)MASS * np + d] = m[src];
            h[(size_t)MOI * np + d] = moi[src];
        }
    if (hipMemcpyAsync(b->base, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return berr(b, SL_E_HIP, "state upload failed");
    rc = fast_configure(b->fb, b->st, L, P.data(), b->p.precision, b->stream, &b->err);
    if (rc) return rc;
    b->lay = L;
    b->P = P;
    b->n_sys_balls.assign(n_per_sys, n_per_sys + n_sys);
    b->n_total = n;
    b->step = 0;
    b->uploaded = true;
    b->forces_valid = false;
    return SL_OK;
}

int sl_batch_set_system_params(sl_batch* b, const double* kin, const double* kout, const double* dt)
{
    if (!b) return SL_E_ARG;
    if (!b->uploaded) return berr(b, SL_E_STATE, "set_system_params before upload");
    std::vector<Phys> P;
    int rc = build_phys(b, b->lay.n_sys, kin, kout, dt, P);
    if (rc) return rc;
    rc = fast_sync_to(b->fb, b->st, b->stream, &b->err);
    if (rc) return rc;
    rc = fast_set_phys(b->fb, b->st, P.data(), b->p.precision, b->stream, &b->err);
    if (rc) return rc;
    b->P = P;
    return SL_OK;
}

int sl_batch_init_forces(sl_batch* b, sl_energy* e_sys)
{
    if (!b) return SL_E_ARG;
    if (!b->uploaded) return berr(b, SL_E_STATE, "init before upload");
    const int rc = fast_init(b->fb, b->st, e_sys, b->stream, &b->err);
    if (rc) return rc;
    b->forces_valid = true;
    return SL_OK;
}

int sl_batch_step(sl_batch* b, int n_steps, int write_every, sl_energy* e, int e_cap)
{
    if (!b) return SL_E_ARG;
    if (!b->uploaded) return berr(b, SL_E_STATE, "step before upload");
    if (!b->forces_valid) return berr(b, SL_E_STATE, "no forces: call sl_batch_init_forces");
    if (n_steps < 0) return berr(b, SL_E_ARG, "n_steps < 0");
    return fast_steps(b->fb, b->st, n_steps, &b->step, write_every, e, e ? e_cap : 0, b->stream, &b->err);
}

int sl_batch_download(sl_batch* b, double* pos3, double* vel3, double* w3, double* acc3, double* aacc3)
{
    if (!b) return SL_E_ARG;
    if (!b->uploaded) return berr(b, SL_E_STATE, "download before upload");
    const int rc = fast_sync_to(b->fb, b->st, b->stream, &b->err);
    if (rc) return rc;
    const size_t np = (size_t)b->n_pad;
    const int nf = BZ + 1;   // positions through angular accelerations
    std::vector<double> h((size_t)nf * np);
    if (hipMemcpyAsync(h.data(), b->base, sizeof(double) * h.size(), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return berr(b, SL_E_HIP, "state download failed");
    int dst = 0;
    for (int s = 0; s < b->lay.n_sys; s++)
        for (int k = 0; k < b->lay.n[s]; k++, dst++) {
            const size_t d = (size_t)b->lay.off[s] + k;
            for (int x = 0; x < 3; x++) {
                if (pos3) pos3[3 * dst + x] = h[(size_t)(PX + x) * np + d];
                if (vel3) vel3[3 * dst + x] = h[(size_t)(VX + x) * np + d];
                if (w3) w3[3 * dst + x] = h[(size_t)(WX + x) * np + d];
                if (acc3) acc3[3 * dst + x] = h[(size_t)(AX + x) * np + d];
                if (aacc3) aacc3[3 * dst + x] = h[(size_t)(BX + x) * np + d];
            }
        }
    return SL_OK;
}

int sl_batch_set_option(sl_batch* b, int option, int value)
{
    if (!b) return SL_E_ARG;
    switch (option) {
    case SL_OPT_GRAPHS:
        b->fb.use_graphs = value != 0;
        fast_clear_graphs(b->fb);
        return SL_OK;
    case SL_OPT_TILE_SPLIT:
        if (value != 0 && value != 1 && value != 2 && value != 4)
            return berr(b, SL_E_ARG, "tile split must be 0 (autotune), 1, 2 or 4 (got %d)", value);
        b->fb.forced_split = value;
        b->fb.split = 0;
        fast_clear_graphs(b->fb);
        return SL_OK;
    case SL_OPT_TILE_QUEUE:
        if (value < 0 || value > 2)
            return berr(b, SL_E_ARG, "tile queue must be 0 (autotune), 1 (per workgroup) or 2 (work queue) (got %d)", value);
        b->fb.forced_queue = value;
        b->fb.split = 0;
        fast_clear_graphs(b->fb);
        return SL_OK;
    case SL_OPT_BALL_WAVES:
        if (value != 0 && value != 4 && value != 8)
            return berr(b, SL_E_ARG, "ball waves must be 0 (by size), 4 or 8 (got %d)", value);
        b->fb.forced_ball_waves = value;
        fast_clear_graphs(b->fb);
        return SL_OK;
    default:
        return berr(b, SL_E_ARG, "unknown option %d", option);
    }
}

int sl_batch_get_option(sl_batch* b, int option, int* value)
{
    if (!b || !value) return SL_E_ARG;
    switch (option) {
    case SL_OPT_GRAPHS:
        *value = b->fb.use_graphs ? 1 : 0;
        return SL_OK;
    case SL_OPT_TILE_SPLIT:
        *value = b->fb.split;
        return SL_OK;
    case SL_OPT_TILE_QUEUE:
        *value = b->fb.split == 0 ? 0 : (b->fb.queue_launch ? 2 : 1);
        return SL_OK;
    case SL_OPT_BALL_WAVES:
        *value = fast_ball_waves(b->fb);
        return SL_OK;
    default:
        return berr(b, SL_E_ARG, "unknown option %d", option);
    }
}

long long sl_batch_step_count(const sl_batch* b) { return b ? b->step : -1; }

const char* sl_batch_last_error(const sl_batch* b) { return b ? b->err.c_str() : "NULL batch"; }

void sl_batch_destroy(sl_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    release(b);
    delete b;
}

}  // extern "C"
