// Plain device copy used to calibrate the HBM traffic counters and to give a
// measured bandwidth next to the nominal 8 TB/s: copies a 2 GiB buffer with
// 16-byte vector loads and non-temporal stores, prints GB/s and the bytes per
// launch.  rocprofv3 FETCH_SIZE / WRITE_SIZE of this kernel against its known
// 2 GiB read and 2 GiB write give the counter scale factors.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void copy16(const f4* __restrict__ a, f4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(a[i], b + i);
}

#define CHK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20;
    const size_t bytes = (size_t)2 << 30;
    const size_t n = bytes / sizeof(f4);
    f4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 0, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(copy16, grid, block, 0, 0, a, b, n);
    CHK(hipEventRecord(e0));
    for (int k = 0; k < iters; ++k) hipLaunchKernelGGL(copy16, grid, block, 0, 0, a, b, n);
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("{\"kernel\": \"copy16\", \"bytes_read_per_launch\": %zu, \"bytes_written_per_launch\": %zu, "
                "\"ms_per_launch\": %.5f, \"gbs\": %.1f}\n",
                bytes, bytes, ms / iters, 2.0 * bytes * iters / (ms / 1e3) / 1e9);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
