// ls_calib.hip -- bandwidth calibration kernels for the measurements
// (liblscalib.so, not part of the solver ABI).
//
// STREAM-style triad c = a + s*b and copy c = a over large FP64 arrays with
// 16-byte accesses, one double2 per thread and no loop: the pattern that
// reached the highest HBM rate on gfx950 among the shapes tried (grid-stride
// loops with 8- or 16-byte accesses and 4 pairs per thread were 5-20%
// slower).  The benchmark times the triad with HIP events to report the
// achievable HBM rate next to the published peak, and rocprofv3 --pmc runs
// use it as the kernel of known traffic that scales FETCH_SIZE / WRITE_SIZE:
// the triad reads 2n and writes n doubles.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__global__ __launch_bounds__(256) void k_triad(const double2* __restrict__ a, const double2* __restrict__ b,
                                               double2* __restrict__ c, double s, int64_t n2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) {
        const double2 x = a[i], y = b[i];
        c[i] = make_double2(fma(s, y.x, x.x), fma(s, y.y, x.y));
    }
}

__global__ __launch_bounds__(256) void k_copy2(const double2* __restrict__ a, double2* __restrict__ c, int64_t n2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) c[i] = a[i];
}

}  // namespace

extern "C" {

// c[i] = a[i] + s*b[i] for i < n; n even, pointers 16-byte aligned.
// Returns 0, or -1 on bad arguments / a launch error.
int lsc_triad(const double* a, const double* b, double* c, double s, int64_t n, void* stream) {
    if (!a || !b || !c || n <= 0 || (n & 1) || ((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) return -1;
    const int64_t n2 = n / 2;
    const int64_t blocks = (n2 + 255) / 256;
    if (blocks > 0x7fffffffLL) return -1;
    hipLaunchKernelGGL(k_triad, dim3((unsigned)blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2*>(a), reinterpret_cast<const double2*>(b),
                       reinterpret_cast<double2*>(c), s, n2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// c[i] = a[i] for i < n; n even, pointers 16-byte aligned.
int lsc_copy(const double* a, double* c, int64_t n, void* stream) {
    if (!a || !c || n <= 0 || (n & 1) || ((uintptr_t)a | (uintptr_t)c) & 15) return -1;
    const int64_t n2 = n / 2;
    const int64_t blocks = (n2 + 255) / 256;
    if (blocks > 0x7fffffffLL) return -1;
    hipLaunchKernelGGL(k_copy2, dim3((unsigned)blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       reinterpret_cast<const double2*>(a), reinterpret_cast<double2*>(c), n2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
