// Surrogate model for collision outcomes: a 4-layer perceptron
//   x (n_in) -> 256 -> 256 -> 256 -> y (n_out), ReLU between layers,
// whose matrix products run on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16:
// bfloat16 operands, binary32 accumulation, 32 of K per instruction).  Training (forward, loss, backward, Adam) and
// inference are all device kernels; every reduction has a fixed order, so a run is
// reproducible bit for bit.
//
// Output layout: y[0..2] are the class logits (stick, bounce, fragment), y[3..] are
// regression targets.  Loss per sample: softmax cross-entropy of the logits plus the
// mean squared error of the regression outputs; the batch loss is their mean.
#include "spacelab/sl_api.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

using bf16 = unsigned short;
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int NL = 4;          // layers
constexpr int GM = 32;         // GEMM block tile rows
constexpr int GN = 32;         // GEMM block tile columns
constexpr int GK = 64;         // GEMM K step staged in LDS
constexpr int NCLS = 3;

__host__ __device__ inline bf16 to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (bf16)(u >> 16);
}

__host__ __device__ inline float from_bf16(bf16 h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// C[M][N] = op(A)[M][K] * op(B)[K][N] (+ bias[N]) with bfloat16 operands and binary32
// accumulation.  op(A)(m, k) = TA ? A[k * lda + m] : A[m * lda + k]; op(B)(k, n) =
// TB ? B[n * ldb + k] : B[k * ldb + n].  Optional ReLU; the result is stored in binary32
// (C) and/or bfloat16 (Cb).  A 256-thread block computes a 32 x 32 tile: wavefront w
// owns the 16 x 16 quarter (w / 2, w % 2), one v_mfma_f32_16x16x32_bf16 per 32 of K
// (lane l: row or column l % 16, K values 8 (l / 16) .. 8 (l / 16) + 7 of the step, the
// same K values for both operands).  Operand tiles of GK = 64 along K are staged through
// LDS (zero fill outside the matrices); global reads run along each matrix's contiguous
// index, and both tiles are stored K-contiguous so each lane reads its eight K values as
// one 16-byte word.  The small output tile keeps many workgroups in flight for the
// 256-wide layers.
//
// With csum, the blocks of the first row of tiles also form the column sums of the
// binary32 matrix csrc[K][ldcs] over its K rows for their GN columns (the bias gradient
// beside the weight gradient X^T dZ): 8 row phases per column, phase p adding rows p,
// p + 8, ... in order, then the phases added in order.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16* A, int lda, const bf16* B, int ldb,
                                              float* C, int ldc, bf16* Cb, int ldcb, const float* bias, int relu,
                                              const float* csrc, int ldcs, float* csum)
{
    __shared__ __attribute__((aligned(16))) bf16 As[GM][GK + 8];   // [m][k]
    __shared__ __attribute__((aligned(16))) bf16 Bt[GN][GK + 8];   // [n][k]
    const int tid = threadIdx.x, w = tid / 64, l = tid % 64;
    const int m0 = blockIdx.y * GM, n0 = blockIdx.x * GN;
    const int wm = (w / 2) * 16, wn = (w % 2) * 16;
    float4_t acc = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += GK) {
        for (int e = tid; e < GM * GK; e += 256) {
            int r, c;   // r: m within the tile, c: k within the step
            if (TA) {
                c = e / GM;
                r = e % GM;
            } else {
                r = e / GK;
                c = e % GK;
            }
            const int gm = m0 + r, gk = k0 + c;
            bf16 v = 0;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[r][c] = v;
        }
        for (int e = tid; e < GK * GN; e += 256) {
            int r, c;   // r: k within the step, c: n within the tile
            if (TB) {
                c = e / GK;
                r = e % GK;
            } else {
                r = e / GN;
                c = e % GN;
            }
            const int gk = k0 + r, gn = n0 + c;
            bf16 v = 0;
            if (gk < K && gn < N) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bt[c][r] = v;
        }
        __syncthreads();
        const int row = wm + (l % 16), col = wn + (l % 16);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 32) {
            const int kq = kk + 8 * (l / 16);
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&As[row][kq]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&Bt[col][kq]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    if (csum && blockIdx.y == 0) {
        __shared__ float part[8][GN];
        const int cl = tid % GN, ph = tid / GN;   // 8 phases of GN columns
        const int c = n0 + cl;
        float sum = 0.f;
        if (c < N) {
            // the loads run ahead of the (ordered) additions
#pragma unroll 8
            for (int r = ph; r < K; r += 8) sum += csrc[(size_t)r * ldcs + c];
        }
        part[ph][cl] = sum;
        __syncthreads();
        if (ph == 0 && c < N) {
            float v = part[0][cl];
            for (int q = 1; q < 8; q++) v += part[q][cl];
            csum[c] = v;
        }
    }
    for (int r = 0; r < 4; r++) {
        const int row = m0 + wm + 4 * (l / 16) + r;
        const int col = n0 + wn + (l % 16);
        if (row >= M || col >= N) continue;
        float v = acc[r];
        if (bias) v += bias[col];
        if (relu && v < 0.f) v = 0.f;
        if (C) C[(size_t)row * ldc + col] = v;
        if (Cb) Cb[(size_t)row * ldcb + col] = to_bf16(v);
    }
}

__global__ void k_f32_to_bf16(int n, const float* x, bf16* y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// Rows of x (n_in wide) into a zero-padded bfloat16 matrix ld wide.
__global__ void k_pack_input(int B, int n_in, int ld, const float* x, bf16* xb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * ld) return;
    const int r = i / ld, c = i % ld;
    xb[i] = c < n_in ? to_bf16(x[(size_t)r * n_in + c]) : (bf16)0;
}

// Loss and its gradient with respect to the outputs, one thread per sample:
// CE(softmax(y[0..2]), cls) + mean_k (y[3+k] - t_k)^2, gradient scaled by 1 / B.
__global__ void k_loss(int B, int n_out, int ldy, const float* y, const int* cls, const float* reg, float* dy,
                       bf16* dyb, float* loss_s)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* z = y + (size_t)b * ldy;
    float mx = z[0];
    for (int k = 1; k < NCLS; k++) mx = fmaxf(mx, z[k]);
    float se = 0.f, p[NCLS];
    for (int k = 0; k < NCLS; k++) {
        p[k] = expf(z[k] - mx);
        se += p[k];
    }
    const int c = cls[b];
    float loss = logf(se) + mx - z[c];
    const int nr = n_out - NCLS;
    float* g = dy + (size_t)b * ldy;
    const float invB = 1.0f / (float)B;
    for (int k = 0; k < NCLS; k++) g[k] = (p[k] / se - (k == c ? 1.f : 0.f)) * invB;
    for (int k = 0; k < nr; k++) {
        const float d = z[NCLS + k] - reg[(size_t)b * nr + k];
        loss += d * d / (float)nr;
        g[NCLS + k] = 2.f * d / (float)nr * invB;
    }
    for (int k = n_out; k < ldy; k++) g[k] = 0.f;
    for (int k = 0; k < ldy; k++) dyb[(size_t)b * ldy + k] = to_bf16(g[k]);
    loss_s[b] = loss;
}

// dZ = dH * (Z > 0): ReLU backward, in binary32 and bfloat16.
__global__ void k_relu_back(int n, const float* z, float* dz, bf16* dzb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = z[i] > 0.f ? dz[i] : 0.f;
    dz[i] = g;
    dzb[i] = to_bf16(g);
}

// Adam (Kingma & Ba 2015) on binary32 master weights, refreshing the bfloat16 copy.
__global__ void k_adam(int n, float* w, bf16* wb, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
}

// Softmax probabilities and predictive entropy (nats) of the class logits.
__global__ void k_predict(int B, int ldy, const float* y, float* prob, float* ent)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* z = y + (size_t)b * ldy;
    float mx = z[0];
    for (int k = 1; k < NCLS; k++) mx = fmaxf(mx, z[k]);
    float p[NCLS], se = 0.f;
    for (int k = 0; k < NCLS; k++) {
        p[k] = expf(z[k] - mx);
        se += p[k];
    }
    float h = 0.f;
    for (int k = 0; k < NCLS; k++) {
        p[k] /= se;
        prob[(size_t)b * NCLS + k] = p[k];
        if (p[k] > 0.f) h -= p[k] * logf(p[k]);
    }
    ent[b] = h;
}

struct Layer {
    int in, out, ldin, ldout;   // logical and padded widths
    float* W = nullptr;         // [in][out] binary32 master (padded rows/cols are zero)
    bf16* Wb = nullptr;         // bfloat16 copy
    float* b = nullptr;         // [out]
    float *gW = nullptr, *gb = nullptr;
    float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
};

}  // namespace

struct sl_mlp {
    int n_in = 0, width = 0, n_out = 0, max_batch = 0;
    hipStream_t stream = nullptr;
    Layer L[NL];
    // activations: Z (pre-activation, binary32) and H (post-activation, bfloat16) per layer
    bf16* xb = nullptr;              // packed input [max_batch][ld0]
    float* Z[NL] = {};
    bf16* Hb[NL] = {};
    float* dZ[NL] = {};
    bf16* dZb[NL] = {};
    float* dH = nullptr;             // scratch [max_batch][width]
    float *xin = nullptr, *reg = nullptr, *loss_s = nullptr, *prob = nullptr, *ent = nullptr;
    int* cls = nullptr;
    long long t = 0;                 // Adam step
    std::string err;
};

namespace {

int pad16(int n) { return (n + 15) / 16 * 16; }

int mlp_err(sl_mlp* m, int code, const char* what)
{
    if (m) m->err = what;
    return code;
}

template <bool TA, bool TB>
void gemm(hipStream_t st, int M, int N, int K, const bf16* A, int lda, const bf16* B, int ldb, float* C, int ldc,
          bf16* Cb, int ldcb, const float* bias, bool relu, const float* csrc = nullptr, int ldcs = 0,
          float* csum = nullptr)
{
    const dim3 grid((N + GN - 1) / GN, (M + GM - 1) / GM);
    k_gemm<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, Cb, ldcb, bias, relu ? 1 : 0, csrc, ldcs,
                                         csum);
}

void free_mlp(sl_mlp* m)
{
    for (auto& L : m->L) {
        void* p[] = {L.W, L.Wb, L.b, L.gW, L.gb, L.mW, L.vW, L.mb, L.vb};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    for (int k = 0; k < NL; k++) {
        void* p[] = {m->Z[k], m->Hb[k], m->dZ[k], m->dZb[k]};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    void* p[] = {m->xb, m->dH, m->xin, m->reg, m->loss_s, m->prob, m->ent, m->cls};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
}

// Forward pass of a batch already packed into m->xb: hidden layers keep their
// post-activation values (binary32 in Z, whose sign is also the ReLU mask, and
// bfloat16 in Hb for the next product); the last layer's outputs are in Z[NL-1].
void forward(sl_mlp* m, int B)
{
    const bf16* in = m->xb;
    for (int k = 0; k < NL; k++) {
        Layer& L = m->L[k];
        const bool last = k == NL - 1;
        gemm<false, false>(m->stream, B, L.ldout, L.ldin, in, L.ldin, L.Wb, L.ldout, m->Z[k], L.ldout,
                           last ? nullptr : m->Hb[k], L.ldout, L.b, !last);
        in = m->Hb[k];
    }
}

// Backward pass after forward() and the loss kernel (dZ[NL-1] set): parameter
// gradients into gW, gb.
void backward(sl_mlp* m, int B)
{
    for (int k = NL - 1; k >= 0; k--) {
        Layer& L = m->L[k];
        const bf16* Hprev = k ? m->Hb[k - 1] : m->xb;
        // weight gradient H^T dZ, and the bias gradient (column sums of dZ) in the same launch
        gemm<true, false>(m->stream, L.ldin, L.ldout, B, Hprev, L.ldin, m->dZb[k], L.ldout, L.gW, L.ldout, nullptr, 0,
                          nullptr, false, m->dZ[k], L.ldout, L.gb);
        if (k > 0) {
            gemm<false, true>(m->stream, B, L.ldin, L.ldout, m->dZb[k], L.ldout, L.Wb, L.ldout, m->dZ[k - 1], L.ldin,
                              nullptr, 0, nullptr, false);
            const int n = B * L.ldin;
            k_relu_back<<<(n + 255) / 256, 256, 0, m->stream>>>(n, m->Z[k - 1], m->dZ[k - 1], m->dZb[k - 1]);
        }
    }
}

int load_batch(sl_mlp* m, int B, const float* x)
{
    if (hipMemcpyAsync(m->xin, x, 4 * (size_t)B * m->n_in, hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return SL_E_HIP;
    const int ld0 = m->L[0].ldin, n = B * ld0;
    k_pack_input<<<(n + 255) / 256, 256, 0, m->stream>>>(B, m->n_in, ld0, m->xin, m->xb);
    return hipGetLastError() == hipSuccess ? SL_OK : SL_E_HIP;
}

// forward + loss for a labelled batch; mean loss to *loss
int loss_pass(sl_mlp* m, int B, const float* x, const int* cls, const float* reg, float* loss)
{
    for (int b = 0; b < B; b++)
        if (cls[b] < 0 || cls[b] >= NCLS) return mlp_err(m, SL_E_ARG, "class label outside 0..2");
    if (load_batch(m, B, x)) return mlp_err(m, SL_E_HIP, "input upload failed");
    const int nr = m->n_out - NCLS;
    if (hipMemcpyAsync(m->cls, cls, 4 * (size_t)B, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->reg, reg, 4 * (size_t)B * nr, hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "label upload failed");
    forward(m, B);
    const Layer& L = m->L[NL - 1];
    k_loss<<<(B + 255) / 256, 256, 0, m->stream>>>(B, m->n_out, L.ldout, m->Z[NL - 1], m->cls, m->reg, m->dZ[NL - 1],
                                                   m->dZb[NL - 1], m->loss_s);
    std::vector<float> ls(B);
    if (hipMemcpyAsync(ls.data(), m->loss_s, 4 * (size_t)B, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "loss read-back failed");
    double sum = 0;
    for (int b = 0; b < B; b++) sum += ls[b];
    if (loss) *loss = (float)(sum / B);
    return SL_OK;
}

int check_batch(sl_mlp* m, int B, const void* x)
{
    if (!m) return SL_E_ARG;
    if (B < 1 || B > m->max_batch || !x) return mlp_err(m, SL_E_ARG, "batch size outside [1, max_batch] or NULL input");
    return SL_OK;
}

}  // namespace

extern "C" {

int sl_mlp_create(sl_mlp** out, int n_in, int width, int n_out, int max_batch, unsigned long long seed)
{
    if (!out) return SL_E_ARG;
    *out = nullptr;
    if (n_in < 1 || width < 16 || n_out <= NCLS || max_batch < 1 || n_in > 4096 || width > 4096 || n_out > 4096)
        return SL_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SL_E_NOGPU;
    sl_mlp* m = new (std::nothrow) sl_mlp();
    if (!m) return SL_E_ARG;
    m->n_in = n_in;
    m->width = width;
    m->n_out = n_out;
    m->max_batch = max_batch;
    const int dims[NL + 1] = {n_in, width, width, width, n_out};
    auto fail = [&](int code) {
        free_mlp(m);
        delete m;
        return code;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail(SL_E_HIP);
    // He-uniform initialisation from a SplitMix64 stream; padded entries are zero
    uint64_t s = seed;
    auto rnd = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int k = 0; k < NL; k++) {
        Layer& L = m->L[k];
        L.in = dims[k];
        L.out = dims[k + 1];
        L.ldin = pad16(L.in);
        L.ldout = pad16(L.out);
        const size_t nw = (size_t)L.ldin * L.ldout;
        std::vector<float> w(nw, 0.f), b(L.ldout, 0.f);
        const double lim = std::sqrt(6.0 / L.in);
        for (int i = 0; i < L.in; i++)
            for (int j = 0; j < L.out; j++) w[(size_t)i * L.ldout + j] = (float)((2 * rnd() - 1) * lim);
        std::vector<bf16> wb(nw);
        for (size_t q = 0; q < nw; q++) wb[q] = to_bf16(w[q]);
        if (hipMalloc(&L.W, 4 * nw) || hipMalloc(&L.Wb, 2 * nw) || hipMalloc(&L.b, 4 * L.ldout) ||
            hipMalloc(&L.gW, 4 * nw) || hipMalloc(&L.gb, 4 * L.ldout) || hipMalloc(&L.mW, 4 * nw) ||
            hipMalloc(&L.vW, 4 * nw) || hipMalloc(&L.mb, 4 * L.ldout) || hipMalloc(&L.vb, 4 * L.ldout))
            return fail(SL_E_HIP);
        if (hipMemcpy(L.W, w.data(), 4 * nw, hipMemcpyHostToDevice) ||
            hipMemcpy(L.Wb, wb.data(), 2 * nw, hipMemcpyHostToDevice) ||
            hipMemcpy(L.b, b.data(), 4 * L.ldout, hipMemcpyHostToDevice) || hipMemset(L.mW, 0, 4 * nw) ||
            hipMemset(L.vW, 0, 4 * nw) || hipMemset(L.mb, 0, 4 * L.ldout) || hipMemset(L.vb, 0, 4 * L.ldout) ||
            hipMemset(L.gW, 0, 4 * nw) || hipMemset(L.gb, 0, 4 * L.ldout))
            return fail(SL_E_HIP);
        const size_t na = (size_t)max_batch * L.ldout;
        if (hipMalloc(&m->Z[k], 4 * na) || hipMalloc(&m->Hb[k], 2 * na) || hipMalloc(&m->dZ[k], 4 * na) ||
            hipMalloc(&m->dZb[k], 2 * na))
            return fail(SL_E_HIP);
    }
    const int ld0 = m->L[0].ldin;
    if (hipMalloc(&m->xb, 2 * (size_t)max_batch * ld0) || hipMalloc(&m->dH, 4 * (size_t)max_batch * pad16(width)) ||
        hipMalloc(&m->xin, 4 * (size_t)max_batch * n_in) || hipMalloc(&m->reg, 4 * (size_t)max_batch * (n_out - NCLS)) ||
        hipMalloc(&m->loss_s, 4 * (size_t)max_batch) || hipMalloc(&m->prob, 4 * (size_t)max_batch * NCLS) ||
        hipMalloc(&m->ent, 4 * (size_t)max_batch) || hipMalloc(&m->cls, 4 * (size_t)max_batch))
        return fail(SL_E_HIP);
    *out = m;
    return SL_OK;
}

int sl_mlp_n_params(const sl_mlp* m)
{
    if (!m) return SL_E_ARG;
    int n = 0;
    for (const Layer& L : m->L) n += L.in * L.out + L.out;
    return n;
}

// Flat parameter layout: for each layer W[in][out] (row-major) then b[out].
int sl_mlp_get_params(sl_mlp* m, float* w)
{
    if (!m || !w) return SL_E_ARG;
    for (Layer& L : m->L) {
        std::vector<float> W((size_t)L.ldin * L.ldout), b(L.ldout);
        if (hipMemcpy(W.data(), L.W, 4 * W.size(), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(b.data(), L.b, 4 * b.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return mlp_err(m, SL_E_HIP, "parameter read-back failed");
        for (int i = 0; i < L.in; i++)
            for (int j = 0; j < L.out; j++) *w++ = W[(size_t)i * L.ldout + j];
        for (int j = 0; j < L.out; j++) *w++ = b[j];
    }
    return SL_OK;
}

int sl_mlp_set_params(sl_mlp* m, const float* w)
{
    if (!m || !w) return SL_E_ARG;
    for (Layer& L : m->L) {
        const size_t nw = (size_t)L.ldin * L.ldout;
        std::vector<float> W(nw, 0.f), b(L.ldout, 0.f);
        for (int i = 0; i < L.in; i++)
            for (int j = 0; j < L.out; j++) W[(size_t)i * L.ldout + j] = *w++;
        for (int j = 0; j < L.out; j++) b[j] = *w++;
        std::vector<bf16> Wb(nw);
        for (size_t q = 0; q < nw; q++) Wb[q] = to_bf16(W[q]);
        if (hipMemcpy(L.W, W.data(), 4 * nw, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(L.Wb, Wb.data(), 2 * nw, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(L.b, b.data(), 4 * L.ldout, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(L.mW, 0, 4 * nw) != hipSuccess || hipMemset(L.vW, 0, 4 * nw) != hipSuccess ||
            hipMemset(L.mb, 0, 4 * L.ldout) != hipSuccess || hipMemset(L.vb, 0, 4 * L.ldout) != hipSuccess)
            return mlp_err(m, SL_E_HIP, "parameter upload failed");
    }
    m->t = 0;
    return SL_OK;
}

int sl_mlp_forward(sl_mlp* m, int B, const float* x, float* y, float* prob, float* ent)
{
    int rc = check_batch(m, B, x);
    if (rc) return rc;
    if (load_batch(m, B, x)) return mlp_err(m, SL_E_HIP, "input upload failed");
    forward(m, B);
    const Layer& L = m->L[NL - 1];
    k_predict<<<(B + 255) / 256, 256, 0, m->stream>>>(B, L.ldout, m->Z[NL - 1], m->prob, m->ent);
    std::vector<float> z((size_t)B * L.ldout);
    if (hipMemcpyAsync(z.data(), m->Z[NL - 1], 4 * z.size(), hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "output read-back failed");
    if (prob && hipMemcpyAsync(prob, m->prob, 4 * (size_t)B * NCLS, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "probability read-back failed");
    if (ent && hipMemcpyAsync(ent, m->ent, 4 * (size_t)B, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "entropy read-back failed");
    if (hipStreamSynchronize(m->stream) != hipSuccess) return mlp_err(m, SL_E_HIP, "forward failed");
    if (y)
        for (int b = 0; b < B; b++)
            for (int k = 0; k < m->n_out; k++) y[(size_t)b * m->n_out + k] = z[(size_t)b * L.ldout + k];
    return SL_OK;
}

int sl_mlp_gradients(sl_mlp* m, int B, const float* x, const int* cls, const float* reg, float* loss, float* grads)
{
    int rc = check_batch(m, B, x);
    if (rc) return rc;
    if (!cls || !reg) return mlp_err(m, SL_E_ARG, "labels are NULL");
    rc = loss_pass(m, B, x, cls, reg, loss);
    if (rc) return rc;
    backward(m, B);
    if (hipStreamSynchronize(m->stream) != hipSuccess) return mlp_err(m, SL_E_HIP, "backward failed");
    if (grads) {
        for (Layer& L : m->L) {
            std::vector<float> W((size_t)L.ldin * L.ldout), b(L.ldout);
            if (hipMemcpy(W.data(), L.gW, 4 * W.size(), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(b.data(), L.gb, 4 * b.size(), hipMemcpyDeviceToHost) != hipSuccess)
                return mlp_err(m, SL_E_HIP, "gradient read-back failed");
            for (int i = 0; i < L.in; i++)
                for (int j = 0; j < L.out; j++) *grads++ = W[(size_t)i * L.ldout + j];
            for (int j = 0; j < L.out; j++) *grads++ = b[j];
        }
    }
    return SL_OK;
}

int sl_mlp_train_step(sl_mlp* m, int B, const float* x, const int* cls, const float* reg, float lr, float* loss)
{
    int rc = check_batch(m, B, x);
    if (rc) return rc;
    if (!cls || !reg || !(lr > 0)) return mlp_err(m, SL_E_ARG, "labels are NULL or lr <= 0");
    rc = loss_pass(m, B, x, cls, reg, loss);
    if (rc) return rc;
    backward(m, B);
    m->t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)m->t), c2 = 1.f - std::pow(b2, (float)m->t);
    for (Layer& L : m->L) {
        const int nw = L.ldin * L.ldout;
        k_adam<<<(nw + 255) / 256, 256, 0, m->stream>>>(nw, L.W, L.Wb, L.gW, L.mW, L.vW, lr, b1, b2, eps, c1, c2);
        k_adam<<<(L.ldout + 255) / 256, 256, 0, m->stream>>>(L.ldout, L.b, nullptr, L.gb, L.mb, L.vb, lr, b1, b2, eps,
                                                             c1, c2);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess)
        return mlp_err(m, SL_E_HIP, "training step failed");
    return SL_OK;
}

const char* sl_mlp_last_error(const sl_mlp* m) { return m ? m->err.c_str() : "NULL model"; }

void sl_mlp_destroy(sl_mlp* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_mlp(m);
    delete m;
}

}  // extern "C"
