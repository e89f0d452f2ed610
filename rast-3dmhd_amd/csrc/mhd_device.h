// Device helpers shared by the FP64 and the mixed-precision kernels.
#pragma once

#include "mhd_kernels.h"

namespace rmhd {

__device__ __forceinline__ int wrapi(int x, int n) {
    int r = x % n;
    return r < 0 ? r + n : r;
}

// XCD-aware block order.  The hardware deals consecutive workgroup ids to
// the 8 XCDs in turn; remap so that each XCD receives a contiguous run of
// tiles (tiles that share halo rows then share an L2).  A bijection for any
// grid size.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int x = bid & 7, local = bid >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + local;
}


// CFL limits of one point (sub:750-856).  mach = OGAMMA*UU/TT before the
// square root; w1..w4 the four pointwise limits.
template <bool MAG>
__device__ __forceinline__ void cfl_point(const Params &p, const double *q, double ddx, double ddy, double ddz,
                                          double rkap, double &mach, double &w1, double &w2, double &w3,
                                          double &w4) {
    const double ru = q[F_RU], rv = q[F_RV], rw = q[F_RW], ro = q[F_RO], tt = q[F_TT];
    const double r = 1.0 / ro;
    double uu = (ru * ru + rv * rv + rw * rw) * (r * r);
    const double mm = p.ogamma * uu / tt;
    if (mm > mach || !(mm == mm)) mach = mm;  // NaN propagates into the max
    if (MAG) {
        const double bx = q[F_BX], by = q[F_BY], bz = q[F_BZ];
        const double b2 = bx * bx + by * by + bz * bz;
        uu = uu + p.gamma * tt + b2 * p.obeta / ro + 2.0 * sqrt(uu * (p.gamma * tt + b2 * p.obeta / ro));
    } else {
        uu = uu + p.gamma * tt + 2.0 * sqrt(uu * p.gamma * tt);
    }
    double dd = fmin(fmin(ddx, ddy), ddz);
    const double a1 = dd / sqrt(uu);
    const double a2 = 0.5 * dd * dd * p.repr * p.cv * ro / rkap;
    const double a3 = 0.375 * dd * dd * p.re * ro;
    w1 = (a1 < w1 || !(a1 == a1)) ? a1 : w1;
    w2 = (a2 < w2 || !(a2 == a2)) ? a2 : w2;
    w3 = (a3 < w3 || !(a3 == a3)) ? a3 : w3;
    if (MAG) w4 = fmin(w4, 0.5 * dd * dd * p.rm);
}

// Workgroup reduction of the CFL accumulators and one set of 64-bit atomics
// on order-preserving keys (min/max are exact, so the order is irrelevant).
// NaN keys sort above +Inf: they win the max and, through the key of -NaN
// (below -Inf), the min, so a non-finite state cannot go unnoticed.
static __device__ void cfl_reduce(double mach, double w1, double w2, double w3, double w4, int lmag,
                           DevScalars *__restrict__ sc) {
    __shared__ unsigned long long red[16][5];
    unsigned long long km = dkey(mach), k1 = dkey(w1), k2 = dkey(w2), k3 = dkey(w3), k4 = dkey(w4);
    // a NaN limit becomes -NaN for the min reduction
    if (!(w1 == w1)) k1 = 0;
    if (!(w2 == w2)) k2 = 0;
    if (!(w3 == w3)) k3 = 0;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o;
        o = __shfl_xor(km, off, 64); km = o > km ? o : km;
        o = __shfl_xor(k1, off, 64); k1 = o < k1 ? o : k1;
        o = __shfl_xor(k2, off, 64); k2 = o < k2 ? o : k2;
        o = __shfl_xor(k3, off, 64); k3 = o < k3 ? o : k3;
        o = __shfl_xor(k4, off, 64); k4 = o < k4 ? o : k4;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        red[wv][0] = km; red[wv][1] = k1; red[wv][2] = k2; red[wv][3] = k3; red[wv][4] = k4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < nw; ++v) {
            km = red[v][0] > km ? red[v][0] : km;
            k1 = red[v][1] < k1 ? red[v][1] : k1;
            k2 = red[v][2] < k2 ? red[v][2] : k2;
            k3 = red[v][3] < k3 ? red[v][3] : k3;
            k4 = red[v][4] < k4 ? red[v][4] : k4;
        }
        atomicMax(&sc->key_mach, km);
        atomicMin(&sc->key_min[0], k1);
        atomicMin(&sc->key_min[1], k2);
        atomicMin(&sc->key_min[2], k3);
        if (lmag) atomicMin(&sc->key_min[3], k4);
    }
}

}  // namespace rmhd
