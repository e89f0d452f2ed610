// Plane means and flux relaxation of the optional FLUXES branches (FP64 form).
//
// HORIZONTAL_MEAN (3dmhdsub.f:3143-3223, NGRID=0 branch) is reproduced in the
// reference's own summation order for an NPEY-slab decomposition: each slab
// (rows j = s*NRY .. s*NRY+NRY-1 of a plane) is summed sequentially, rows
// outer and points inner, then divided by FLOAT(NPX*NPY), and the slab
// quotients are added in slab order.  One workgroup per (plane, slab) forms
// a slab's sum; one workgroup combines them per plane.  The means feed
//   ROM   density diffusion (ID, sub:1301-1371), with the ghost planes of RO
//         holding the mirrored planes 2 and npz-1 (ID or LREM);
//   TTM   Rempel conduction (LREM, sub:1536-1581) and the relaxation;
//   HRAD  the radiative heating of the mean profile (LREM, sub:1567-1573);
//   radd  RLAX*TFAC*CORRECT(K) of the flux relaxation (sub:1615-1765), whose
//         saved factor TFAC and maximum TTM_MAX live in RelaxState and whose
//         vertical integral runs rank by rank over an NPEZ-rank decomposition
//         as the reference's does.
// These run on the stage input before each stage kernel; the optional
// branches are not a throughput path and favour the reference's arithmetic
// order over parallel reductions.

#include "mhd_kernels.h"

namespace rmhd {

namespace {

// The reference's sums are one sequential chain per slab (rows outer, points
// inner; the slab is a contiguous run of NRY*NPX points of the plane).  A
// workgroup per (plane, slab) stages the run through LDS in chunks of
// SUM_CHUNK values (coalesced, double-buffered) and its thread 0 adds them
// in order, so the chain runs at the latency of an FP64 add instead of that
// of a dependent global load.
constexpr int SUM_T = 256, SUM_CHUNK = 1024;

// sum + v[0] + v[1] + ... + v[m-1], added one by one in this order; the LDS
// reads of a group of 8 are issued before its adds.
__device__ __forceinline__ double chain_add(double sum, const double *v, int m) {
    int i = 0;
    for (; i + 8 <= m; i += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum = sum + x[u];
    }
    for (; i < m; ++i) sum = sum + v[i];
    return sum;
}

// Slab sums of plane k (0..npz+1) of RO (WHAT & 1) and TT (WHAT & 2).
template <int WHAT>
__global__ __launch_bounds__(SUM_T) void k_slab_sums(const Params p, const double *__restrict__ q,
                                                     double *__restrict__ part) {
    __shared__ double buf[2][2][SUM_CHUNK];
    const int npl = p.npz + 2, nsl = p.npey;
    const int b = blockIdx.x;  // b = k * nsl + sl
    const int k = b / nsl, sl = b - k * nsl;
    const int nry = p.npy / nsl;
    const long long fs = p.fsize, pl = p.plane, n = (long long)nry * p.npx;
    int kr = k;
    if (p.mirror && k == 0) kr = 2;
    if (p.mirror && k == p.npz + 1) kr = p.npz - 1;
    const double *ro = q + 3 * fs + (long long)kr * pl + (long long)sl * nry * p.npx;
    const double *tt = q + 4 * fs + (long long)k * pl + (long long)sl * nry * p.npx;
    auto stage = [&](long long c0, int slot) {
        for (int i = threadIdx.x; i < SUM_CHUNK; i += SUM_T) {
            if (c0 + i < n) {
                if (WHAT & 1) buf[slot][0][i] = ro[c0 + i];
                if (WHAT & 2) buf[slot][1][i] = tt[c0 + i];
            }
        }
    };
    // RO's chain on thread 0, TT's on thread 64 (another SIMD)
    const int f = threadIdx.x == 0 ? 0 : threadIdx.x == 64 ? 1 : -1;
    const bool chain = f >= 0 && (WHAT & (1 << f));
    double sum = 0.0;
    const long long nch = (n + SUM_CHUNK - 1) / SUM_CHUNK;
    stage(0, 0);
    __syncthreads();
    for (long long c = 0; c < nch; ++c) {
        const int slot = (int)(c & 1);
        if (c + 1 < nch) stage((c + 1) * SUM_CHUNK, slot ^ 1);
        if (chain) sum = chain_add(sum, buf[slot][f], (int)min((long long)SUM_CHUNK, n - c * SUM_CHUNK));
        __syncthreads();
    }
    if (chain) part[(long long)f * npl * nsl + b] = sum;
    else if (f >= 0) part[(long long)f * npl * nsl + b] = 0.0;
}

// The slab quotients of plane k added in slab order.
__device__ __forceinline__ double slab_mean(const Params &p, const double *part, int k) {
    const double cnt = (double)(float)(p.npx * p.npy);
    double v = 0.0;
    for (int sl = 0; sl < p.npey; ++sl) {
        const double w = part[k * p.npey + sl] / cnt;
        v = sl == 0 ? w : v + w;
    }
    return v;
}

// One workgroup: ROM, WWZ (ID=1), TTM, then HRAD (LREM).
__global__ void k_plane_means(const Params p, const Metrics m, const double *__restrict__ part, double *rom,
                              double *wwz, double *ttm, double *hrad, int what) {
    const int npl = p.npz + 2;
    for (int k = threadIdx.x; k < npl; k += blockDim.x) {
        if (what & 1) {
            const double v = slab_mean(p, part, k);
            rom[k] = v;
            if (p.id == 1) wwz[k] = 1.0 / v;  // sub:1323-1325
        }
        if (what & 2) ttm[k] = slab_mean(p, part + (size_t)npl * p.npey, k);
    }
    __syncthreads();
    if (p.lrem) {
        for (int k = 1 + threadIdx.x; k <= p.npz; k += blockDim.x) {
            const double tp1 = ttm[k + 1], t0 = ttm[k], tm1 = ttm[k - 1];
            hrad[k] = ((tp1 - 2.0 * t0 + tm1) * p.h2z * m.dzzdz[k] * m.dzzdz[k] + (tp1 - tm1) * p.hz * m.d2zzdz2[k])
                          * m.rkapa[k]
                      + (tp1 - tm1) * p.hz * m.dzzdz[k] * m.dkapa[k];
        }
    }
}

// Slab sums of the convective-flux summand -RW*(2.5*(TT-TTM) + 0.5*|u|^2) of
// planes k = 1..npz+1 (sub:1653-1662); u = RU*(1/RO) as FLUXES forms it.  The
// summands are formed by the staging threads, the chain by thread 0, as in
// k_slab_sums.
__global__ __launch_bounds__(SUM_T) void k_fconm_sums(const Params p, const double *__restrict__ q,
                                                      const double *__restrict__ ttm, double *__restrict__ part) {
    __shared__ double buf[2][SUM_CHUNK];
    const int nsl = p.npey;
    const int b = blockIdx.x;  // b = (k - 1) * nsl + sl
    const int k = 1 + b / nsl, sl = b % nsl;
    const int nry = p.npy / nsl;
    const long long fs = p.fsize, pl = p.plane, n = (long long)nry * p.npx;
    const long long so = (long long)sl * nry * p.npx;
    const int kr = (p.mirror && k == p.npz + 1) ? p.npz - 1 : k;
    const double *ru = q + (long long)k * pl + so, *rv = q + fs + (long long)k * pl + so;
    const double *rw = q + 2 * fs + (long long)k * pl + so;
    const double *ro = q + 3 * fs + (long long)kr * pl + so, *tt = q + 4 * fs + (long long)k * pl + so;
    const double tm = ttm[k];
    auto stage = [&](long long c0, int slot) {
        for (int i = threadIdx.x; i < SUM_CHUNK; i += SUM_T) {
            const long long o = c0 + i;
            if (o < n) {
                const double ri = 1.0 / ro[o];
                const double uu = ru[o] * ri, vv = rv[o] * ri, ww = rw[o] * ri;
                buf[slot][i] = -(rw[o] * (2.5 * (tt[o] - tm) + 0.5 * (uu * uu + vv * vv + ww * ww)));
            }
        }
    };
    double sum = 0.0;
    const long long nch = (n + SUM_CHUNK - 1) / SUM_CHUNK;
    stage(0, 0);
    __syncthreads();
    for (long long c = 0; c < nch; ++c) {
        const int slot = (int)(c & 1);
        if (c + 1 < nch) stage((c + 1) * SUM_CHUNK, slot ^ 1);
        if (threadIdx.x == 0) sum = chain_add(sum, buf[slot], (int)min((long long)SUM_CHUNK, n - c * SUM_CHUNK));
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k * nsl + sl] = sum;
}

// One thread: TFAC, FCONM, FRADM, ALPHA and the per-rank integral CORRECT,
// then radd[k] = (RLAX*TFAC)*CORRECT(K) (fixed-T top, fixed-flux bottom).
__global__ void k_relax(const Params p, const Metrics m, const double *__restrict__ zee,
                        const double *__restrict__ part, const double *__restrict__ ttm, double *__restrict__ work,
                        DevScalars *__restrict__ sc, RelaxState *__restrict__ rs, double *__restrict__ radd) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int npz = p.npz, npl = npz + 2;
    double *fconm = work, *fradm = work + npl, *corr = work + 2 * npl, *endval = work + 3 * npl;
    double tfac = rs->tfac;
    if (sc->timt > p.tstart && tfac > 1e-4) {
        // the bottom boundary plane's mean temperature (sub:1619-1630)
        const double dt = sc->dt;
        if (ttm[npz] <= rs->ttm_max) {
            tfac = tfac - dt / (3.0 * p.toff);
        } else {
            rs->ttm_max = ttm[npz];
            tfac = tfac + dt / (3.0 * p.toff);
            tfac = fmin(tfac, 2.0);
        }
        for (int k = 1; k <= npz + 1; ++k) fconm[k] = slab_mean(p, part, k);
        // radiative flux (sub:1665-1677): resident plane k is padded K = k+1
        for (int k = 2; k <= npz - 1; ++k)
            fradm[k] = (ttm[k + 1] - ttm[k - 1]) * p.hz * m.dzzdz[k] * m.rkapa[k] / p.repr;
        fradm[1] = (-3.0 * ttm[1] + 4.0 * ttm[2] - ttm[3]) * p.hz * m.dzzdz[1] * m.rkapa[1] / p.repr;
        fradm[npz] = (3.0 * ttm[npz] - 4.0 * ttm[npz - 1] + ttm[npz - 2]) * p.hz * m.dzzdz[npz] * m.rkapa[npz] / p.repr;
        const double ftot = p.lrem ? (double)(8.07f * 0.4f) * p.theta : p.theta / p.repr;
        // per rank, local planes 2..NZ (padded K = pz*NRZ + kl); the bottom
        // rank's last one is its ghost plane and is never used (sub:1693-1705)
        const int nrz = npz / p.npez, nzl = nrz + 2;
        for (int pz = 0; pz < p.npez; ++pz) {
            const int kl_end = pz == p.npez - 1 ? nzl - 1 : nzl;
            double c = 0.0, aprev = 0.0;
            for (int kl = 2; kl <= kl_end; ++kl) {
                const int kg = pz * nrz + kl, k = kg - 1;
                double a = 1.0 - (fconm[k] + fradm[k]) / ftot;
                a = a / (1.0 + 4.0 * fabs(a));
                if (kl == 2) c = 0.0;
                else c = c + 0.5 * (a + aprev) * (zee[kg - 1] - zee[kg - 2]);
                if (kl <= nzl - 1) corr[k] = c;
                aprev = a;
            }
            endval[pz] = c;
        }
        // offsets: the end values of the ranks above, in rank order (sub:1733-1752)
        for (int pz = 0; pz < p.npez; ++pz) {
            double sum = 0.0;
            for (int ipe = 1; ipe <= pz; ++ipe) sum = sum + endval[ipe - 1];
            for (int kl = 2; kl <= nzl - 1; ++kl) corr[pz * nrz + kl - 1] = corr[pz * nrz + kl - 1] + sum;
        }
    } else {
        for (int k = 1; k <= npz; ++k) corr[k] = 0.0;
    }
    rs->tfac = tfac;
    rs->icall = rs->icall + 1;
    const double rt = p.rlax * tfac;
    for (int k = 1; k <= npz; ++k) radd[k] = rt * corr[k];
}

}  // namespace

hipError_t launch_opt_means(const Params &p, const Metrics &m, const double *q, const OptBuffers &o,
                            DevScalars *sc, hipStream_t s) {
    const int npl = p.npz + 2;
    const int what = (p.id != 0 ? 1 : 0) | ((p.lrem || p.rlax > 0.0) ? 2 : 0);
    if (what) {
        const int n = npl * p.npey;
        if (what == 1) hipLaunchKernelGGL(k_slab_sums<1>, dim3(n), dim3(SUM_T), 0, s, p, q, o.part);
        else if (what == 2) hipLaunchKernelGGL(k_slab_sums<2>, dim3(n), dim3(SUM_T), 0, s, p, q, o.part);
        else hipLaunchKernelGGL(k_slab_sums<3>, dim3(n), dim3(SUM_T), 0, s, p, q, o.part);
        hipLaunchKernelGGL(k_plane_means, dim3(1), dim3(256), 0, s, p, m, o.part, o.rom, o.wwz, o.ttm, o.hrad, what);
    }
    if (p.rlax > 0.0) {
        const int n = (p.npz + 1) * p.npey;
        hipLaunchKernelGGL(k_fconm_sums, dim3(n), dim3(SUM_T), 0, s, p, q, o.ttm, o.part);
        hipLaunchKernelGGL(k_relax, dim3(1), dim3(64), 0, s, p, m, o.zee, o.part, o.ttm, o.work, sc, o.rs, o.radd);
    }
    return hipGetLastError();
}

}  // namespace rmhd
