// Plain device-to-device copy of a known number of bytes.  Used to
// calibrate the FETCH_SIZE / WRITE_SIZE hardware counters of rocprofv3
// (their reading for a copy of N bytes gives the counter-to-byte factor that
// converts the stage kernel's counters into HBM bytes per launch).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_f64x2(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 1024;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "copy_probe: hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    for (int rep = 0; rep < 3; ++rep) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(copy_f64x2, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
