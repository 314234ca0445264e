// Small device utilities.  Counters are cleared by kernels rather than
// hipMemsetAsync so the same sequence is safe to capture in a HIP graph.
#include "spacelab/sl_api.h"
#include "sl_internal.hpp"
#include "sl_pair_ref.hpp"

namespace sl {

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_zero_i32(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_zero_u64<<<(n + 255) / 256, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_zero_i32(int* p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_zero_i32<<<(n + 255) / 256, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace sl

namespace sl {

// Arithmetic probe: the binary64 division, square root and logarithm exactly as
// the ordered kernels issue them (compiled with the same flags).
__global__ void k_probe(int n, const double* a, const double* b, double* q, double* s, double* lg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    q[i] = a[i] / b[i];
    s[i] = __builtin_sqrt(a[i]);
    lg[i] = ::log(a[i]);
}

}  // namespace sl

extern "C" int sl_probe_divsqrt(int n, const double* a, const double* b, double* q, double* s, double* lg)
{
    if (n < 1 || !a || !b || !q || !s || !lg) return SL_E_ARG;
    double* d = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n;
    if (hipMalloc(&d, 5 * bytes) != hipSuccess) return SL_E_HIP;
    int rc = SL_OK;
    if (hipMemcpy(d, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = SL_E_HIP;
    } else {
        sl::k_probe<<<(n + 255) / 256, 256>>>(n, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(q, d + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(s, d + 3 * (size_t)n, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(lg, d + 4 * (size_t)n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = SL_E_HIP;
    }
    hipFree(d);
    return rc;
}

namespace sl {

// The ordered kernels' division through a correctly rounded reciprocal.
__global__ void k_probe_div_rcp(int n, const double* a, const double* b, double* q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double y = 1.0 / b[i];
    q[i] = div_rcp(a[i], b[i], y);
}

}  // namespace sl

extern "C" int sl_probe_div_rcp(int n, const double* a, const double* b, double* q)
{
    if (n < 1 || !a || !b || !q) return SL_E_ARG;
    double* d = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n;
    if (hipMalloc(&d, 3 * bytes) != hipSuccess) return SL_E_HIP;
    int rc = SL_OK;
    if (hipMemcpy(d, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = SL_E_HIP;
    } else {
        sl::k_probe_div_rcp<<<(n + 255) / 256, 256>>>(n, d, d + n, d + 2 * (size_t)n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(q, d + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = SL_E_HIP;
    }
    hipFree(d);
    return rc;
}

namespace sl {

// Plain streaming copy, 16 bytes per thread: the reference point for scaling the
// FETCH_SIZE / WRITE_SIZE hardware counters to bytes.
__global__ void k_copy16(const double2* __restrict__ src, double2* __restrict__ dst, size_t n16)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

}  // namespace sl

extern "C" int sl_probe_copy(unsigned long long bytes, int reps, double* ms_per_copy)
{
    if (bytes < 16 || reps < 1) return SL_E_ARG;
    const size_t n16 = (size_t)(bytes / 16);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess) return SL_E_HIP;
    if (hipMalloc(&b, n16 * 16) != hipSuccess) {
        hipFree(a);
        return SL_E_HIP;
    }
    int rc = SL_OK;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    if (hipMemset(a, 1, n16 * 16) != hipSuccess) rc = SL_E_HIP;
    const unsigned blocks = (unsigned)((n16 + 255) / 256);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps && rc == SL_OK; r++) sl::k_copy16<<<blocks, 256>>>(a, b, n16);
    hipEventRecord(e1, 0);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = SL_E_HIP;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}
