// Horizontal means and mean flux profiles of the resident state (a
// diagnostic beside the step): the reference's HORIZONTAL_MEAN
// (3dmhdsub.f:3143-3223) of T and rho, and the mean convective flux FCONM
// (3dmhdsub.f:1650-1661), the horizontal mean of
//     -RW (2.5 (T - TTM) + 0.5 |u|^2),   u = RU (1/RO),
// on every interior plane.  One workgroup per plane: a first pass sums T and
// rho, a second pass the convective flux with the plane's mean T; each sum
// runs in a fixed order (per-thread strides, then a fixed reduction tree), so
// the result is the same on every run.  (The reference sums sequentially;
// the two orders agree to rounding.)

#include "mhd_device.h"

namespace rmhd {

constexpr int PROF_THREADS = 512;

// Deterministic workgroup sum: wavefront butterflies, then the wavefront
// partials in order by thread 0.  Every thread returns the total.
__device__ double block_sum(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();  // red may still be read from a previous call
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
        for (int w = 1; w < nw; ++w) s += red[w];
        red[16] = s;
    }
    __syncthreads();
    return red[16];
}

// Values of the state at offset o of plane k, FP64 form (register q).
struct StateFp64 {
    const double *q;
    long long fs;
    __device__ double get(int f, int, long long o) const { return q[f * fs + o]; }
};

// Mixed form: stored values h and resting residual R about the background.
struct StateMixed {
    const float *h;
    const unsigned *rr;
    Background bg;
    long long fs;
    __device__ double get(int f, int k, long long o) const {
        if (f == 4)
            return bg.tt[k] + (double)trep(bg.ro32[k], bg.tt32[k], h[3 * fs + o], h[4 * fs + o]) +
                   (double)r_get(rr, fs, 4, o);
        const double v = (double)h[f * fs + o] + (double)r_get(rr, fs, f, o);
        return f == 3 ? bg.ro[k] + v : v;
    }
};

// out: [3][npz] = TTM, ROM, FCONM of the planes k = 1..npz.
template <class S>
__global__ __launch_bounds__(PROF_THREADS) void k_flux_profiles(const Params p, const S st, double *__restrict__ out) {
    __shared__ double red[17];
    const int k = blockIdx.x + 1;
    const long long pl = p.plane, base = (long long)k * pl;
    const double cnt = (double)(float)(p.npx * p.npy);
    double st_ = 0.0, sr = 0.0;
    for (long long v = threadIdx.x; v < pl; v += blockDim.x) {
        st_ += st.get(4, k, base + v);
        sr += st.get(3, k, base + v);
    }
    const double ttm = block_sum(st_, red) / cnt;
    const double rom = block_sum(sr, red) / cnt;
    double sc = 0.0;
    for (long long v = threadIdx.x; v < pl; v += blockDim.x) {
        const long long o = base + v;
        const double ri = 1.0 / st.get(3, k, o);
        const double rw = st.get(2, k, o);
        const double uu = st.get(0, k, o) * ri, vv = st.get(1, k, o) * ri, ww = rw * ri;
        const double w1 = st.get(4, k, o) - ttm;
        sc += -(rw * (2.5 * w1 + 0.5 * (uu * uu + vv * vv + ww * ww)));
    }
    const double fconm = block_sum(sc, red) / cnt;
    if (threadIdx.x == 0) {
        out[k - 1] = ttm;
        out[p.npz + k - 1] = rom;
        out[2 * p.npz + k - 1] = fconm;
    }
}

hipError_t launch_flux_profiles(const Params &p, const double *q, double *out, hipStream_t s) {
    const StateFp64 st{q, p.fsize};
    hipLaunchKernelGGL(k_flux_profiles<StateFp64>, dim3(p.npz), dim3(PROF_THREADS), 0, s, p, st, out);
    return hipGetLastError();
}

hipError_t launch_flux_profiles_mixed(const Params &p, const float *h, const unsigned *rr, const Background &bg,
                                      double *out, hipStream_t s) {
    const StateMixed st{h, rr, bg, p.fsize};
    hipLaunchKernelGGL(k_flux_profiles<StateMixed>, dim3(p.npz), dim3(PROF_THREADS), 0, s, p, st, out);
    return hipGetLastError();
}

}  // namespace rmhd
