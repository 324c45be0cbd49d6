// hbm_copy: plain device-to-device copy kernel of a known size, used to
// calibrate HBM traffic counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
// and to measure the achievable copy bandwidth.
//
//   hbm_copy [MiB per buffer, default 2048] [repetitions, default 10]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

#define CK(x)                                                                    \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 2048;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 10;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 1, bytes));
    CK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    // one 16-byte element per thread: the fastest of the shapes tried
    // (grid-stride loops over 2048 - 16384 workgroups with 1 - 8 loads in
    // flight per thread ran at 4.2 - 4.9 TB/s, this one at 5.4)
    const int block = 256;
    const unsigned grid = (unsigned)((n + block - 1) / block);
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(block), 0, 0, a, b, n);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(block), 0, 0, a, b, n);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double s = ms * 1e-3 / reps;
    std::printf("{\"bytes_read_per_launch\": %zu, \"bytes_written_per_launch\": %zu, \"ms_per_launch\": %.4f, "
                "\"GBps\": %.1f}\n", bytes, bytes, s * 1e3, 2.0 * bytes / s / 1e9);
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
