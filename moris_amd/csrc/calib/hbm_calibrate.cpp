// hbm_calibrate -- streaming copy and write-only kernels over FP64 buffers of a known size.
//
// Used to scale the HBM counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports per dispatch:
// the factor between a counter and the bytes such a kernel demonstrably moves is applied to the
// same counter of the assembly kernels.  Prints the event-timed rate of each kernel.
//
//   mgpu_hbm_calibrate [n_values]      (default 51153914, i.e. ~410 MB per buffer)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_write(double* __restrict__ b, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = v;
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 1;                                                              \
        }                                                                          \
    } while (0)

int main(int argc, char** argv) {
    const size_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 51153914ull;
    if (n == 0 || n > (size_t)8 << 30) {
        std::fprintf(stderr, "n_values must be in 1 .. 8G\n");
        return 2;
    }
    double *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, n * sizeof(double)));
    CHECK(hipMalloc(&b, n * sizeof(double)));
    CHECK(hipMemset(a, 0, n * sizeof(double)));
    CHECK(hipMemset(b, 0, n * sizeof(double)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const unsigned grid = 16384, block = 256;
    const int iters = 20;
    float ms = 0.f;
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(block), 0, 0, a, b, n);
    CHECK(hipEventRecord(e0));
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(block), 0, 0, a, b, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= iters;
    std::printf("k_copy  %zu values: %.4f ms, %.1f GB/s (read + write), %.1f KiB read, %.1f KiB written\n", n, ms,
                2.0 * n * 8 / (ms * 1e-3) / 1e9, n * 8 / 1024.0, n * 8 / 1024.0);
    CHECK(hipEventRecord(e0));
    for (int it = 0; it < iters; ++it) hipLaunchKernelGGL(k_write, dim3(grid), dim3(block), 0, 0, b, n, 1.0);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= iters;
    std::printf("k_write %zu values: %.4f ms, %.1f GB/s, %.1f KiB written\n", n, ms, 1.0 * n * 8 / (ms * 1e-3) / 1e9,
                n * 8 / 1024.0);
    CHECK(hipGetLastError());
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
