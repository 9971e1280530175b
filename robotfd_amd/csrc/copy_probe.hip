// Plain device-to-device copy of a known size, the reference point that
// scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters into HBM bytes.
//     copy_probe [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy16(const uint4 *src, uint4 *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? (size_t)std::atol(argv[1]) : 256;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float total = 0;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy16, dim3(4096), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r) total += ms;
    }
    std::printf("copy_probe bytes_read=%zu bytes_written=%zu avg_ms=%.4f GB/s=%.1f\n", bytes, bytes,
                total / (reps - 1), 2.0 * bytes / (total / (reps - 1) * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
