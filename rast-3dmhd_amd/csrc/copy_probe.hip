// Streaming kernels of a known number of bytes: the measured HBM ceiling
// (copy, read-only, write-only) and the calibration of rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters (their reading for a copy of N bytes
// gives the counter-to-byte factor used for the stage kernel's traffic).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

template <int U>
__global__ __launch_bounds__(256) void copy_u(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) b[i + u * stride] = v[u];
    }
    for (; i < n; i += stride) b[i] = a[i];
}

template <int U>
__global__ __launch_bounds__(256) void copy_nt(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = a[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            __builtin_nontemporal_store(v[u].x, &b[i + u * stride].x);
            __builtin_nontemporal_store(v[u].y, &b[i + u * stride].y);
        }
    }
    for (; i < n; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(256) void read_only(const double2 *__restrict__ a, double *__restrict__ out, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double s = 0.0;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        double2 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        s += v0.x + v0.y + v1.x + v1.y + v2.x + v2.y + v3.x + v3.y;
    }
    for (; i < n; i += stride) s += a[i].x + a[i].y;
    if (s == 12345.678) out[0] = s;  // keeps the loads alive
}

__global__ __launch_bounds__(256) void write_only(double2 *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = make_double2(1.0, 2.0);
}

static float timed(void (*launch)(void *), void *arg) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int rep = 0; rep < 5; ++rep) {
        (void)hipEventRecord(e0);
        launch(arg);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

struct Args {
    double2 *a, *b;
    double *o;
    size_t n;
    int grid;
};

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 1024;
    const int mode = argc > 2 ? atoi(argv[2]) : -1;  // -1: all variants; 0: plain copy only
    const size_t bytes = mib << 20;
    Args g{};
    g.n = bytes / sizeof(double2);
    if (hipMalloc(&g.a, bytes) != hipSuccess || hipMalloc(&g.b, bytes) != hipSuccess ||
        hipMalloc(&g.o, 64) != hipSuccess) {
        fprintf(stderr, "copy_probe: hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(g.a, 1, bytes) != hipSuccess) return 1;
    int ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) ncu = prop.multiProcessorCount;
    const int grids[3] = {ncu * 8, ncu * 32, ncu * 128};
    for (int gi = 0; gi < 3; ++gi) {
        g.grid = grids[gi];
        float ms;
        ms = timed([](void *p) { Args *a = (Args *)p; hipLaunchKernelGGL(copy_u<1>, dim3(a->grid), dim3(256), 0, 0, a->a, a->b, a->n); }, &g);
        printf("copy U=1   grid=%6d %zu MiB: %.3f ms %.1f GB/s (read+write)\n", g.grid, mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
        if (mode == 0) continue;
        ms = timed([](void *p) { Args *a = (Args *)p; hipLaunchKernelGGL(copy_u<4>, dim3(a->grid), dim3(256), 0, 0, a->a, a->b, a->n); }, &g);
        printf("copy U=4   grid=%6d %zu MiB: %.3f ms %.1f GB/s (read+write)\n", g.grid, mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
        ms = timed([](void *p) { Args *a = (Args *)p; hipLaunchKernelGGL(copy_nt<4>, dim3(a->grid), dim3(256), 0, 0, a->a, a->b, a->n); }, &g);
        printf("copy NT U=4 grid=%6d %zu MiB: %.3f ms %.1f GB/s (read+write)\n", g.grid, mib, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
        ms = timed([](void *p) { Args *a = (Args *)p; hipLaunchKernelGGL(read_only, dim3(a->grid), dim3(256), 0, 0, a->a, a->o, a->n); }, &g);
        printf("read       grid=%6d %zu MiB: %.3f ms %.1f GB/s\n", g.grid, mib, ms, 1.0 * bytes / (ms * 1e-3) / 1e9);
        ms = timed([](void *p) { Args *a = (Args *)p; hipLaunchKernelGGL(write_only, dim3(a->grid), dim3(256), 0, 0, a->b, a->n); }, &g);
        printf("write      grid=%6d %zu MiB: %.3f ms %.1f GB/s\n", g.grid, mib, ms, 1.0 * bytes / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(g.a);
    (void)hipFree(g.b);
    (void)hipFree(g.o);
    return 0;
}
