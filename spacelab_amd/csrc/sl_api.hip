// C ABI of the spacelab_amd library: contexts, state transfer, stepping.
//
// The step sequence mirrors sim_looper/sim_one_step (ColliderSingleCore.cpp:
// 298-331, 78-288): drift -> pair pass -> closing kick (+ energies on write
// steps).  All kernels of a context run in order on its one stream.  Energies
// are read back only on write steps.
#include "spacelab/sl_api.h"

#include "sl_common.hpp"
#include "sl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

using namespace sl;

struct sl_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    int n_max = 0, n = 0, n_pad = 0;
    sl_params p{};
    double* base = nullptr;     // NFIELD * n_pad doubles
    State st{};
    double* pe_part = nullptr;  // per-ball pair energies (write steps)
    double* e_part = nullptr;   // per-block KE/p/L partials
    unsigned long long* cnt = nullptr;  // [0] contacts total, [1] contacts this write step, [2..3] tuner trials
    int ord_shape = -1;         // launch shape of the ordered pair pass (-1: not chosen yet)
    int forced_ord_shape = 0;   // 1 + shape forces it
    std::map<int, hipGraphExec_t> ord_graphs;   // captured runs of ordered non-write steps, by length
    int* nonfinite = nullptr;
    double* h_pe = nullptr;     // pinned
    double* h_e = nullptr;      // pinned
    unsigned long long* h_cnt = nullptr;
    int* h_nonfinite = nullptr;
    FastBuffers fb{};
    long long step = 0;
    bool uploaded = false;
    bool forces_valid = false;
    std::string err;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Mark { int kind; size_t e0, e1; };
    std::vector<Mark> marks;
    size_t prof_e0 = 0;          // open event of the launch being profiled
    bool prof_open = false;
    sl_profile acc_prof{};
};

namespace {

int set_err(sl_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err((c), SL_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

int check_params(sl_ctx* c, const sl_params* p)
{
    if (!p) return set_err(c, SL_E_ARG, "params is NULL");
    if (!(p->dt > 0) || !std::isfinite(p->dt)) return set_err(c, SL_E_ARG, "dt must be > 0 (got %g)", p->dt);
    if (!(p->kin >= 0) || !(p->kout >= 0)) return set_err(c, SL_E_ARG, "kin/kout must be >= 0");
    if (!(p->h_min > 0)) return set_err(c, SL_E_ARG, "h_min must be > 0");
    if (p->precision != SL_PREC_FP64 && p->precision != SL_PREC_MIXED)
        return set_err(c, SL_E_ARG, "unknown precision %d", p->precision);
    if (p->order != SL_ORDER_ORDERED && p->order != SL_ORDER_FAST)
        return set_err(c, SL_E_ARG, "unknown order %d", p->order);
    if (p->precision == SL_PREC_MIXED && p->order == SL_ORDER_ORDERED)
        return set_err(c, SL_E_ARG, "mixed precision requires SL_ORDER_FAST");
    return SL_OK;
}

Phys phys_of(const sl_params& p)
{
    Phys P;
    P.G = p.G;
    P.Ha = p.Ha;
    P.h_min = p.h_min;
    P.u_s = p.u_s;
    P.u_r = p.u_r;
    P.kin = p.kin;
    P.kout = p.kout;
    P.dt = p.dt;
    P.far_gravity = p.far_gravity;
    return P;
}

hipEvent_t next_event(sl_ctx* c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Record an event pair around a launch when profiling.
struct ProfScope {
    sl_ctx* c;
    int kind;
    size_t e0 = 0;
    bool on = false;
    ProfScope(sl_ctx* c_, int kind_) : c(c_), kind(kind_)
    {
        if (c->prof) {
            hipEvent_t e = next_event(c);
            if (e && hipEventRecord(e, c->stream) == hipSuccess) {
                e0 = c->ev_used - 1;
                on = true;
            }
        }
    }
    ~ProfScope()
    {
        if (on) {
            hipEvent_t e = next_event(c);
            if (e && hipEventRecord(e, c->stream) == hipSuccess) c->marks.push_back({kind, e0, c->ev_used - 1});
        }
    }
};

int flush_profile(sl_ctx* c)
{
    if (c->marks.empty()) {
        c->ev_used = 0;
        return SL_OK;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (const auto& mk : c->marks) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[mk.e0], c->ev_pool[mk.e1]));
        if (mk.kind == 0) {
            c->acc_prof.pair_ms += ms;
            c->acc_prof.pair_launches++;
        } else {
            c->acc_prof.ball_ms += ms;
            c->acc_prof.ball_launches++;
        }
    }
    c->marks.clear();
    c->ev_used = 0;
    return SL_OK;
}

int free_ctx_buffers(sl_ctx* c)
{
    if (c->base) hipFree(c->base);
    if (c->pe_part) hipFree(c->pe_part);
    if (c->e_part) hipFree(c->e_part);
    if (c->cnt) hipFree(c->cnt);
    if (c->nonfinite) hipFree(c->nonfinite);
    if (c->h_pe) hipHostFree(c->h_pe);
    if (c->h_e) hipHostFree(c->h_e);
    if (c->h_cnt) hipHostFree(c->h_cnt);
    if (c->h_nonfinite) hipHostFree(c->h_nonfinite);
    fast_free(c->fb);
    c->base = nullptr;
    return SL_OK;
}

// Launch hooks of the fast engine while profiling: an event pair around each launch.
void hook_before(void* u, int)
{
    sl_ctx* c = static_cast<sl_ctx*>(u);
    c->prof_open = false;
    hipEvent_t e = next_event(c);
    if (e && hipEventRecord(e, c->stream) == hipSuccess) {
        c->prof_e0 = c->ev_used - 1;
        c->prof_open = true;
    }
}

void hook_after(void* u, int kind)
{
    sl_ctx* c = static_cast<sl_ctx*>(u);
    if (!c->prof_open) return;
    c->prof_open = false;
    hipEvent_t e = next_event(c);
    if (e && hipEventRecord(e, c->stream) == hipSuccess) c->marks.push_back({kind, c->prof_e0, c->ev_used - 1});
}

// Read back and sum the partials of a write step (fixed order: deterministic).
int read_energy(sl_ctx* c, int n_pe_blocks, int n_e_blocks, sl_energy* out, bool has_pe)
{
    if (has_pe && n_pe_blocks > 0)
        HIPCHK(c, hipMemcpyAsync(c->h_pe, c->pe_part, sizeof(double) * n_pe_blocks, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_e, c->e_part, sizeof(double) * 7 * n_e_blocks, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->cnt, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_nonfinite, c->nonfinite, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sl_energy e{};
    if (has_pe)
        for (int b = 0; b < n_pe_blocks; b++) e.PE += c->h_pe[b];
    for (int b = 0; b < n_e_blocks; b++) {
        e.KE += c->h_e[7 * b + 0];
        e.p[0] += c->h_e[7 * b + 1];
        e.p[1] += c->h_e[7 * b + 2];
        e.p[2] += c->h_e[7 * b + 3];
        e.L[0] += c->h_e[7 * b + 4];
        e.L[1] += c->h_e[7 * b + 5];
        e.L[2] += c->h_e[7 * b + 6];
    }
    e.contacts = (long long)c->h_cnt[1];
    e.nonfinite = *c->h_nonfinite;
    if (out) *out = e;
    return SL_OK;
}

// One ordered pair pass and per-ball pass for the current positions.
// Captured runs of ordered steps hold the parameters, the launch shape and the state's
// size in their kernel arguments: they are dropped (on an idle stream) whenever one of
// those changes.
void clear_ord_graphs(sl_ctx* c)
{
    if (c->ord_graphs.empty()) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->ord_graphs)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    c->ord_graphs.clear();
}

constexpr int ORD_GRAPH_MAX_STEPS = 128;

// len ordered steps without energies, as one replay of a graph captured on first use:
// drift, then per step the pair pass and the closing kick, each kick but the last fused
// with the next step's drift (the same operations as the direct launches, so the same
// bits, with one launch fewer per step).
int ordered_run(sl_ctx* c, int len)
{
    hipGraphExec_t ex = nullptr;
    auto it = c->ord_graphs.find(len);
    if (it != c->ord_graphs.end()) {
        ex = it->second;
    } else {
        const Phys P = phys_of(c->p);
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return set_err(c, SL_E_HIP, "ordered step capture");
        bool ok = launch_drift(c->st, c->n, c->p.dt, c->stream) == hipSuccess;
        for (int k = 0; k < len && ok; k++)
            ok = launch_pairs_ordered(c->st, P, c->n, 0, c->pe_part, c->cnt, c->stream, c->ord_shape) == hipSuccess &&
                 (k + 1 < len ? launch_kick_drift(c->st, c->n, c->p.dt, c->nonfinite, c->stream)
                              : launch_kick(c->st, c->n, c->p.dt, 0, 1, c->e_part, c->nonfinite, c->stream)) == hipSuccess;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (!ok || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return set_err(c, SL_E_HIP, "ordered step capture");
        }
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess || !ex) return set_err(c, SL_E_HIP, "ordered step graph instantiation");
        c->ord_graphs[len] = ex;
    }
    HIPCHK(c, hipGraphLaunch(ex, c->stream));
    return SL_OK;
}

int pair_pass(sl_ctx* c, bool write, bool kick, int* n_pe_blocks)
{
    const Phys P = phys_of(c->p);
    if (write) HIPCHK(c, launch_zero_u64(c->cnt + 1, 1, c->stream));
    {
        ProfScope ps(c, 0);
        HIPCHK(c, launch_pairs_ordered(c->st, P, c->n, write ? 1 : 0, c->pe_part, c->cnt, c->stream,
                                       c->ord_shape >= 0 ? c->ord_shape : ordered_default_shape(c->n)));
    }
    // one sphere: no pair pass ran and no PE partial was written
    *n_pe_blocks = c->n > 1 ? c->n : 0;   // per-ball pair energies
    {
        ProfScope ps(c, 1);
        HIPCHK(c, launch_kick(c->st, c->n, c->p.dt, write ? 1 : 0, kick ? 1 : 0, c->e_part, c->nonfinite, c->stream));
    }
    return SL_OK;
}

// Launch shape of the ordered pair pass for this system: a forced one, a stored choice, or
// each shape timed on the live (drifted) state -- one warm-up and three timed launches,
// energies off, counters into the trial slots -- with the fastest stored.  Every shape
// gives the same bits, so the choice changes only the time.  The trial launches write the
// same accelerations the step's own pass then writes again.
int tune_ordered(sl_ctx* c)
{
    if (c->forced_ord_shape > 0) {
        c->ord_shape = c->forced_ord_shape - 1;
        return SL_OK;
    }
    const std::string key = "ordered:n" + std::to_string(c->n) + ":grav" + std::to_string(c->p.far_gravity);
    if (const int cached = tune_cache_get(key); cached >= 1 && cached <= ordered_shapes()) {
        c->ord_shape = cached - 1;
        return SL_OK;
    }
    const Phys P = phys_of(c->p);
    const int ns = ordered_shapes();
    std::vector<float> t(ns, 1e30f);
    // two events of the pool, handed back on every way out
    const size_t ev_mark = c->ev_used;
    struct EvRelease {
        sl_ctx* c;
        size_t mark;
        ~EvRelease() { c->ev_used = mark; }
    } release{c, ev_mark};
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    if (!e0 || !e1) return set_err(c, SL_E_HIP, "tuner events");
    for (int s = 0; s < ns; s++) {
        HIPCHK(c, launch_pairs_ordered(c->st, P, c->n, 0, c->pe_part, c->cnt + 2, c->stream, s));
        HIPCHK(c, hipEventRecord(e0, c->stream));
        for (int r = 0; r < 3; r++) HIPCHK(c, launch_pairs_ordered(c->st, P, c->n, 0, c->pe_part, c->cnt + 2, c->stream, s));
        HIPCHK(c, hipEventRecord(e1, c->stream));
        HIPCHK(c, hipEventSynchronize(e1));
        HIPCHK(c, hipEventElapsedTime(&t[s], e0, e1));
    }
    int best = 0;
    for (int s = 1; s < ns; s++)
        if (t[s] < t[best]) best = s;
    c->ord_shape = best;
    tune_cache_put(key, best + 1, t.data(), ns);
    return SL_OK;
}

// Lay the uploaded balls out as a one-system fast engine.
int configure_fast(sl_ctx* c)
{
    const FastLayout L = fast_layout(1, &c->n);
    const Phys P = phys_of(c->p);
    return fast_configure(c->fb, c->st, L, &P, c->p.precision, c->stream, &c->err);
}

}  // namespace

extern "C" {

int sl_params_preset(sl_params* out, int preset)
{
    if (!out) return SL_E_ARG;
    sl_params p{};
    // dust_const.hpp:21-39
    p.G = 6.67e-8;
    p.Ha = 4.7e-12;
    p.h_min = 1e-5 * .1;
    p.u_s = 0.1;
    p.u_r = 1e-5;
    p.kin = -1;
    p.kout = -1;
    p.dt = -1;
    p.far_gravity = 0;
    p.precision = SL_PREC_FP64;
    p.order = SL_ORDER_ORDERED;
    if (preset == 1) {
        p.u_s = 0.3;
        p.far_gravity = 1;
    } else if (preset != 0) {
        return SL_E_ARG;
    }
    *out = p;
    return SL_OK;
}

int sl_create(sl_ctx** out, int n_max, const sl_params* p)
{
    if (!out) return SL_E_ARG;
    *out = nullptr;
    if (n_max < 1 || n_max > (1 << 24)) return SL_E_ARG;
    sl_ctx* c = new (std::nothrow) sl_ctx();
    if (!c) return SL_E_ARG;
    int rc = check_params(c, p);
    if (rc) {
        delete c;
        return rc;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete c;
        return SL_E_NOGPU;
    }
    hipDeviceProp_t prop;
    if (hipGetDevice(&c->dev) != hipSuccess || hipGetDeviceProperties(&prop, c->dev) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete c;
        return SL_E_NOGPU;
    }
    c->p = *p;
    c->fb.n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->n_max = n_max;
    c->n_pad = (n_max + 255) / 256 * 256;
    auto fail = [&](int code) {
        free_ctx_buffers(c);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(SL_E_HIP);
    if (hipMalloc(&c->base, sizeof(double) * (size_t)NFIELD * c->n_pad) != hipSuccess) return fail(SL_E_HIP);
    const int max_blocks = n_max + 8;   // per-ball pair energies
    const int e_blocks = kick_blocks(n_max) + 8;
    if (hipMalloc(&c->pe_part, sizeof(double) * max_blocks) != hipSuccess) return fail(SL_E_HIP);
    if (hipMalloc(&c->e_part, sizeof(double) * 7 * e_blocks) != hipSuccess) return fail(SL_E_HIP);
    if (hipMalloc(&c->cnt, sizeof(unsigned long long) * 4) != hipSuccess) return fail(SL_E_HIP);
    if (hipMalloc(&c->nonfinite, sizeof(int)) != hipSuccess) return fail(SL_E_HIP);
    if (hipHostMalloc(&c->h_pe, sizeof(double) * max_blocks, 0) != hipSuccess) return fail(SL_E_HIP);
    if (hipHostMalloc(&c->h_e, sizeof(double) * 7 * e_blocks, 0) != hipSuccess) return fail(SL_E_HIP);
    if (hipHostMalloc(&c->h_cnt, sizeof(unsigned long long) * 2, 0) != hipSuccess) return fail(SL_E_HIP);
    if (hipHostMalloc(&c->h_nonfinite, sizeof(int), 0) != hipSuccess) return fail(SL_E_HIP);
    if (fast_alloc(c->fb, c->n_pad, 1) != 0) return fail(SL_E_HIP);
    for (int f = 0; f < NFIELD; f++) c->st.f[f] = c->base + (size_t)f * c->n_pad;
    if (hipMemsetAsync(c->base, 0, sizeof(double) * (size_t)NFIELD * c->n_pad, c->stream) != hipSuccess ||
        hipMemsetAsync(c->cnt, 0, sizeof(unsigned long long) * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->nonfinite, 0, sizeof(int), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(SL_E_HIP);
    *out = c;
    return SL_OK;
}

int sl_set_params(sl_ctx* c, const sl_params* p)
{
    if (!c) return SL_E_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (c->uploaded) {
        // The fast path keeps the current state in either of two buffer sets and
        // derives per-ball constants from the parameters: bring the canonical set up
        // to date, then re-derive for the new parameters.
        rc = fast_sync_to(c->fb, c->st, c->stream, &c->err);
        if (rc) return rc;
        const Phys P = phys_of(*p);
        rc = fast_set_phys(c->fb, c->st, &P, p->precision, c->stream, &c->err);
        if (rc) return rc;
    }
    if (p->far_gravity != c->p.far_gravity) c->ord_shape = -1;
    clear_ord_graphs(c);
    c->p = *p;
    return SL_OK;
}

int sl_set_option(sl_ctx* c, int option, int value)
{
    if (!c) return SL_E_ARG;
    switch (option) {
    case SL_OPT_GRAPHS:
        c->fb.use_graphs = value != 0;
        fast_clear_graphs(c->fb);
        clear_ord_graphs(c);
        return SL_OK;
    case SL_OPT_TILE_SPLIT:
        if (value != 0 && value != 1 && value != 2 && value != 4)
            return set_err(c, SL_E_ARG, "tile split must be 0 (autotune), 1, 2 or 4 (got %d)", value);
        c->fb.forced_split = value;
        c->fb.split = 0;
        fast_clear_graphs(c->fb);
        return SL_OK;
    case SL_OPT_TILE_QUEUE:
        if (value < 0 || value > 2)
            return set_err(c, SL_E_ARG, "tile queue must be 0 (autotune), 1 (per workgroup) or 2 (work queue) (got %d)", value);
        c->fb.forced_queue = value;
        c->fb.split = 0;
        fast_clear_graphs(c->fb);
        return SL_OK;
    case SL_OPT_BALL_WAVES:
        if (value != 0 && value != 4 && value != 8)
            return set_err(c, SL_E_ARG, "ball waves must be 0 (by size), 4 or 8 (got %d)", value);
        c->fb.forced_ball_waves = value;
        fast_clear_graphs(c->fb);
        return SL_OK;
    case SL_OPT_ORDERED_SHAPE:
        if (value < 0 || value > ordered_shapes())
            return set_err(c, SL_E_ARG, "ordered shape must be 0 (tuned) or 1..%d (got %d)", ordered_shapes(), value);
        c->forced_ord_shape = value;
        c->ord_shape = -1;
        clear_ord_graphs(c);
        return SL_OK;
    default:
        return set_err(c, SL_E_ARG, "unknown option %d", option);
    }
}

int sl_get_option(sl_ctx* c, int option, int* value)
{
    if (!c || !value) return SL_E_ARG;
    switch (option) {
    case SL_OPT_GRAPHS:
        *value = c->fb.use_graphs ? 1 : 0;
        return SL_OK;
    case SL_OPT_TILE_SPLIT:
        *value = c->fb.split;
        return SL_OK;
    case SL_OPT_TILE_QUEUE:
        *value = c->fb.split == 0 ? 0 : (c->fb.queue_launch ? 2 : 1);
        return SL_OK;
    case SL_OPT_BALL_WAVES:
        *value = fast_ball_waves(c->fb);
        return SL_OK;
    case SL_OPT_ORDERED_SHAPE:
        *value = c->ord_shape + 1;
        return SL_OK;
    default:
        return set_err(c, SL_E_ARG, "unknown option %d", option);
    }
}

int sl_upload(sl_ctx* c, int n, const double* pos3, const double* vel3, const double* w3,
              const double* acc3, const double* aacc3, const double* R, const double* m,
              const double* moi, const double* pos_prev3)
{
    if (!c) return SL_E_ARG;
    if (n < 1 || n > c->n_max) return set_err(c, SL_E_ARG, "n=%d outside [1, %d]", n, c->n_max);
    if (!pos3 || !vel3 || !w3 || !R || !m || !moi) return set_err(c, SL_E_ARG, "required array is NULL");
    for (int i = 0; i < n; i++) {
        if (!(R[i] > 0) || !std::isfinite(R[i])) return set_err(c, SL_E_ARG, "R[%d] = %g must be > 0", i, R[i]);
        if (!(m[i] > 0) || !std::isfinite(m[i])) return set_err(c, SL_E_ARG, "m[%d] = %g must be > 0", i, m[i]);
        if (!(moi[i] > 0) || !std::isfinite(moi[i])) return set_err(c, SL_E_ARG, "moi[%d] = %g must be > 0", i, moi[i]);
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(pos3[3 * i + k]) || !std::isfinite(vel3[3 * i + k]) || !std::isfinite(w3[3 * i + k]))
                return set_err(c, SL_E_ARG, "non-finite position/velocity/spin at ball %d", i);
    }
    const size_t np = (size_t)c->n_pad;
    std::vector<double> h((size_t)NFIELD * np, 0.0);
    auto put3 = [&](int f0, const double* a) {
        if (!a) return;
        for (int i = 0; i < n; i++) {
            h[(size_t)(f0 + 0) * np + i] = a[3 * i + 0];
            h[(size_t)(f0 + 1) * np + i] = a[3 * i + 1];
            h[(size_t)(f0 + 2) * np + i] = a[3 * i + 2];
        }
    };
    put3(PX, pos3);
    put3(VX, vel3);
    put3(WX, w3);
    put3(AX, acc3);
    put3(BX, aacc3);
    put3(HX, vel3);
    put3(GX, w3);
    put3(QX, pos_prev3 ? pos_prev3 : pos3);
    for (int i = 0; i < n; i++) {
        h[(size_t)RAD * np + i] = R[i];
        h[(size_t)MASS * np + i] = m[i];
        h[(size_t)MOI * np + i] = moi[i];
    }
    // padding balls: unit radius and mass far away are never read (loops stop at n)
    // the context counts as uploaded only once the fast engine is configured for the new
    // state: a failed configuration leaves it unusable rather than half set up
    c->uploaded = false;
    c->forces_valid = false;
    HIPCHK(c, hipMemcpyAsync(c->base, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n != c->n) c->ord_shape = -1;
    clear_ord_graphs(c);
    c->n = n;
    c->step = 0;
    const int rc = configure_fast(c);
    if (rc) return rc;
    c->uploaded = true;
    c->forces_valid = (acc3 != nullptr);
    return SL_OK;
}

int sl_init_forces(sl_ctx* c, sl_energy* e)
{
    if (!c) return SL_E_ARG;
    if (!c->uploaded) return set_err(c, SL_E_STATE, "init before upload");
    if (c->p.order == SL_ORDER_FAST) {
        const int rc = fast_init(c->fb, c->st, e, c->stream, &c->err);
        if (rc) return rc;
    } else {
        fast_canon_changed(c->fb);
        HIPCHK(c, launch_copy_prev(c->st, c->n, c->stream));
        HIPCHK(c, launch_zero_i32(c->nonfinite, 1, c->stream));
        int npe = 0;
        int rc = pair_pass(c, true, false, &npe);
        if (rc) return rc;
        rc = read_energy(c, npe, kick_blocks(c->n), e, true);
        if (rc) return rc;
    }
    c->forces_valid = true;
    return flush_profile(c);
}

int sl_step(sl_ctx* c, int n_steps, int write_every, sl_energy* e_out, int e_cap)
{
    if (!c) return SL_E_ARG;
    if (!c->uploaded) return set_err(c, SL_E_STATE, "step before upload");
    if (n_steps < 0) return set_err(c, SL_E_ARG, "n_steps < 0");
    if (!c->forces_valid) return set_err(c, SL_E_STATE, "no forces: call sl_init_forces or upload acc");
    if (c->p.order == SL_ORDER_FAST) {
        // the engine launches directly while profiling (events around each kernel),
        // else replays captured step graphs
        c->fb.hook = c->prof ? LaunchHook{hook_before, hook_after, c} : LaunchHook{};
        const int nw = fast_steps(c->fb, c->st, n_steps, &c->step, write_every, e_out, e_cap, c->stream, &c->err);
        const int rc = flush_profile(c);
        if (nw < 0) return nw;
        if (rc) return rc;
        return nw;
    }
    if (n_steps > 0) fast_canon_changed(c->fb);
    // Runs of steps without energies replay captured graphs (not while profiling, where
    // each launch is timed, nor under a profiling tool); write steps, and the step that
    // chooses the launch shape, are launched directly.
    const bool graphs = c->fb.use_graphs && !c->prof && !tool_attached();
    int nw = 0;
    for (int k = 0; k < n_steps;) {
        if (graphs && c->ord_shape >= 0) {
            int len = 0;
            while (k + len < n_steps && len < ORD_GRAPH_MAX_STEPS) {
                const long long s = c->step + 1 + len;
                if (write_every > 0 && s % write_every == 0) break;
                len++;
            }
            if (len > 0) {
                const int rc = ordered_run(c, len);
                if (rc) return rc;
                c->step += len;
                k += len;
                continue;
            }
        }
        k++;
        c->step++;
        const bool write = write_every > 0 && (c->step % write_every) == 0;
        HIPCHK(c, launch_drift(c->st, c->n, c->p.dt, c->stream));
        if (c->ord_shape < 0) {
            const int rt = tune_ordered(c);
            if (rt) return rt;
        }
        int npe = 0;
        int rc = pair_pass(c, write, true, &npe);
        if (rc) return rc;
        if (write) {
            sl_energy e;
            rc = read_energy(c, npe, kick_blocks(c->n), &e, true);
            if (rc) return rc;
            if (e_out && nw < e_cap) e_out[nw] = e;
            nw++;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int rc = flush_profile(c);
    if (rc) return rc;
    return nw;
}

int sl_download(sl_ctx* c, double* pos3, double* vel3, double* w3, double* acc3, double* aacc3)
{
    if (!c) return SL_E_ARG;
    if (!c->uploaded) return set_err(c, SL_E_STATE, "download before upload");
    const int rc = fast_sync_to(c->fb, c->st, c->stream, &c->err);
    if (rc) return rc;
    const size_t np = (size_t)c->n_pad;
    std::vector<double> h((size_t)NFIELD * np);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->base, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto get3 = [&](int f0, double* a) {
        if (!a) return;
        for (int i = 0; i < c->n; i++) {
            a[3 * i + 0] = h[(size_t)(f0 + 0) * np + i];
            a[3 * i + 1] = h[(size_t)(f0 + 1) * np + i];
            a[3 * i + 2] = h[(size_t)(f0 + 2) * np + i];
        }
    };
    get3(PX, pos3);
    get3(VX, vel3);
    get3(WX, w3);
    get3(AX, acc3);
    get3(BX, aacc3);
    return SL_OK;
}

int sl_download_half(sl_ctx* c, double* velh3, double* wh3)
{
    if (!c) return SL_E_ARG;
    if (!c->uploaded) return set_err(c, SL_E_STATE, "download before upload");
    const int rc = fast_sync_to(c->fb, c->st, c->stream, &c->err);
    if (rc) return rc;
    const size_t np = (size_t)c->n_pad;
    std::vector<double> h(6 * np);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->st.f[HX], sizeof(double) * 6 * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->n; i++)
        for (int k = 0; k < 3; k++) {
            if (velh3) velh3[3 * i + k] = h[(size_t)k * np + i];
            if (wh3) wh3[3 * i + k] = h[(size_t)(3 + k) * np + i];
        }
    return SL_OK;
}

int sl_synchronize(sl_ctx* c)
{
    if (!c) return SL_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SL_OK;
}

long long sl_step_count(const sl_ctx* c) { return c ? c->step : -1; }

const char* sl_last_error(const sl_ctx* c) { return c ? c->err.c_str() : "NULL context"; }

void sl_destroy(sl_ctx* c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    clear_ord_graphs(c);
    for (auto e : c->ev_pool) hipEventDestroy(e);
    free_ctx_buffers(c);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int sl_profile_enable(sl_ctx* c, int enable, int reset)
{
    if (!c) return SL_E_ARG;
    int rc = flush_profile(c);
    if (rc) return rc;
    c->prof = enable != 0;
    if (reset) {
        c->acc_prof = sl_profile{};
        HIPCHK(c, launch_zero_u64(c->cnt, 1, c->stream));
        HIPCHK(c, fast_reset_counters(c->fb, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return SL_OK;
}

int sl_profile_read(sl_ctx* c, sl_profile* out)
{
    if (!c || !out) return SL_E_ARG;
    int rc = flush_profile(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->cnt, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->acc_prof;
    out->contacts = (long long)c->h_cnt[0];
    long long fc = 0, fn = 0;
    if (fast_read_counters(c->fb, c->stream, &fc, &fn) != 0) return set_err(c, SL_E_HIP, "reading counters failed");
    out->contacts += fc;
    out->near_pairs = fn;
    return SL_OK;
}

}  // extern "C"
