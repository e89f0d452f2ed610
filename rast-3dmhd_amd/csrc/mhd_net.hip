// Subgrid-closure network: three 3x3x3 convolutions cin -> width -> width
// -> cout with GELU after the first two, on BF16 MFMA with FP32
// accumulation (include/rast3dmhd_closure.h; no counterpart in the
// reference).  Training (forward, backward, Adam) and inference run on the
// device.
//
// Layout: activations channel-last, a[z][y][x][c] (BF16, c contiguous),
// periodic in x and y, zero padding in z.  Master parameters FP32 in the
// order W1[co][tap][ci], b1, W2, b2, W3, b3 with tap = 9(dz+1)+3(dy+1)+(dx+1).
//
// k_conv: implicit GEMM D[voxel][co] = sum_k A[voxel][k] B[k][co] with
//   k = tap*CIN + ci, on v_mfma_f32_16x16x32_bf16 (lane l holds
//   A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16], D[4(l/16)..+3][l%16]).  A
//   workgroup (4 wavefronts) owns a 64 x 4 column tile and marches down a z
//   chunk with a ring of three input planes (halo 1) in LDS; each wavefront
//   computes one row of 64 voxels (4 M-tiles) for all output channels.  Each
//   lane holds its B fragments of the whole layer in registers (most in
//   accumulation registers, which the MFMA reads directly) for the whole
//   march, so the fully unrolled K loop reads only A from LDS, one K block
//   ahead of the MFMAs that use it.  The voxel stride in LDS is an odd
//   number of 16-byte units so the 16 lanes of a fragment read hit distinct
//   banks.  The epilogue adds the bias and applies GELU (hidden layers), or
//   multiplies by GELU'(z) (backward).
//   The backward pass for the input reuses it with the flipped, transposed
//   kernel.
// k_dw: weight gradient dW[co][tap][ci] = sum_v G[v][co] X[v+off][ci] as
//   D[co][ci] over K = voxels; per-workgroup partials, reduced in a fixed
//   order (deterministic).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rast3dmhd_closure.h"
#include "mhd_kernels.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CTX = 64, CTY = 4;          // conv output tile per plane
constexpr int LW = CTX + 2, LH = CTY + 2; // LDS plane with halo
constexpr int DSEG = 64;                  // dW: voxels per segment (two K blocks)
constexpr int GS = DSEG + 8;              // dW: LDS row stride of G (odd number of 16-B units)
constexpr int XS = LW + 7;                // dW: LDS row stride of X (73, odd)

template <int CIN>
constexpr int cstride() { return ((CIN / 8) % 2 == 0) ? CIN + 8 : CIN; }

__device__ __forceinline__ int wrapn(int x, int n) {
    int r = x % n;
    return r < 0 ? r + n : r;
}

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26
// (absolute error below 1.5e-7, far under the BF16 rounding of the
// activations): branch-free, one hardware reciprocal and one exponential.
__device__ __forceinline__ float erf_as(float u) {
    const float s = fabsf(u);
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.3275911f, s, 1.0f));
    float q = __builtin_fmaf(1.061405429f, t, -1.453152027f);
    q = __builtin_fmaf(q, t, 1.421413741f);
    q = __builtin_fmaf(q, t, -0.284496736f);
    q = __builtin_fmaf(q, t, 0.254829592f);
    const float e = 1.0f - q * t * __expf(-s * s);
    return copysignf(e, u);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erf_as(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erf_as(x * 0.70710678118654752f)) + x * __expf(-0.5f * x * x) * 0.39894228040143268f;
}

// EPI 0: out32 = acc + bias.  EPI 1: zout = acc + bias, out16 = gelu(zout).
// EPI 2: out16 = acc * gelu'(zprev).  EPI 3: out32 = acc.
// EPI 4: out16 = gelu(acc + bias) (inference: the pre-activation, which only
// the backward pass reads, is not stored).
template <int CIN, int NT, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_conv(const bf16 *__restrict__ x, const bf16 *__restrict__ wt, int kp,
                                              const float *__restrict__ bias, const float *__restrict__ zprev,
                                              bf16 *__restrict__ out16, float *__restrict__ out32,
                                              float *__restrict__ zout, int cout, int ostride, int nx, int ny,
                                              int nz, int kchunk) {
    constexpr int CS = cstride<CIN>();
    constexpr int PLANE = LW * LH * CS;
    constexpr int NCH = CIN / 8;
    constexpr int NLOAD = (LW * LH * NCH + 255) / 256;
    constexpr int KTOT = 27 * CIN, KB = (KTOT + 31) / 32;
    extern __shared__ uint4 lds_raw[];
    bf16 *const lds = reinterpret_cast<bf16 *>(lds_raw);

    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lr = lane & 15, lg = lane >> 4;
    const int ntx = (nx + CTX - 1) / CTX, nty = (ny + CTY - 1) / CTY;
    const int bid = blockIdx.x;
    const int tz = bid / (ntx * nty);
    const int rem = bid - tz * (ntx * nty);
    const int x0 = (rem % ntx) * CTX, y0 = (rem / ntx) * CTY;
    const int zb = tz * kchunk, ze = min(nz, zb + kchunk);
    if (zb >= nz) return;

    // this thread's 16-byte pieces of an input plane with halo: offsets in a
    // plane of x (-1: no piece) and in an LDS plane, computed once
    int goff[NLOAD], loff[NLOAD];
#pragma unroll
    for (int i = 0; i < NLOAD; ++i) {
        const int c = t + 256 * i;
        goff[i] = -1;
        loff[i] = 0;
        if (c < LW * LH * NCH) {
            const int vox = c / NCH, part = c - vox * NCH;
            const int row = vox / LW, col = vox - row * LW;
            const int gy = wrapn(y0 + row - 1, ny), gx = wrapn(x0 + col - 1, nx);
            goff[i] = (gy * nx + gx) * CIN + part * 8;
            loff[i] = vox * CS + part * 8;
        }
    }
    const long long pstride = (long long)nx * ny * CIN;
    uint4 pre[NLOAD];
    auto fetch = [&](int zz) {
        const bool in = zz >= 0 && zz < nz;
        const bf16 *xz = x + (in ? zz : 0) * pstride;
#pragma unroll
        for (int i = 0; i < NLOAD; ++i) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (in && goff[i] >= 0) v = *reinterpret_cast<const uint4 *>(xz + goff[i]);
            pre[i] = v;
        }
    };
    auto slot = [](int zz) { return ((zz % 3) + 3) % 3; };
    auto store = [&](int zz) {
        bf16 *pl = lds + slot(zz) * PLANE;
#pragma unroll
        for (int i = 0; i < NLOAD; ++i)
            if (goff[i] >= 0) *reinterpret_cast<uint4 *>(pl + loff[i]) = pre[i];
    };
    // the layer's B fragments (weights) of this lane, held in registers
    // for the whole march: every plane uses all of them
    bf16x8 wb[KB][NT];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            wb[kb][nt] = *reinterpret_cast<const bf16x8 *>(wt + (long long)(16 * nt + lr) * kp + kb * 32 + lg * 8);
    for (int zz = zb - 1; zz <= zb + 1; ++zz) {
        fetch(zz);
        store(zz);
    }
    __syncthreads();

    for (int z = zb; z < ze; ++z) {
        const bool more = z + 1 < ze;
        if (more) fetch(z + 2);
        f32x4 acc[4][NT];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
        // the A fragments of K block kb: 4 M tiles of this wavefront's row
        auto frag = [&](int kb, bf16x8 *a) __attribute__((always_inline)) {
            const int k0 = kb * 32 + lg * 8;
            if (k0 < KTOT) {
                const int tap = k0 / CIN, ci0 = k0 - tap * CIN;
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                const bf16 *pl = lds + slot(z + dz) * PLANE + ((wv + 1 + dy) * LW + 1 + dx) * CS + ci0;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) a[mt] = *reinterpret_cast<const bf16x8 *>(pl + (16 * mt + lr) * CS);
            } else {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) a[mt] = bf16x8{};
            }
        };
        // K loop, software-pipelined: the LDS reads of block kb + 1 are
        // issued before the MFMAs of block kb, which cover their latency
        // (one wavefront per SIMD: nothing else would)
        bf16x8 fa[2][4];
        frag(0, fa[0]);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            const int c = kb & 1;
            if (kb + 1 < KB) frag(kb + 1, fa[c ^ 1]);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[c][mt], wb[kb][nt], acc[mt][nt], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        // epilogue: lane (lr, lg) holds voxels x0 + 16 mt + 4 lg + r of this
        // wavefront's row for channel 16 nt + lr.  The row and channel
        // checks are per lane; the column check only in a ragged tile.
        const int gy = y0 + wv;
        const long long rowo = (((long long)z * ny + gy) * nx + x0) * ostride;
        auto emit = [&](float v, int co, float bco, long long o) __attribute__((always_inline)) {
            if (EPI == 0) out32[o] = v + bco;
            if (EPI == 1 || EPI == 4) {
                const float zz = v + bco;
                if (EPI == 1) zout[o] = zz;
                out16[o] = (bf16)gelu(zz);
            }
            if (EPI == 2) out16[o] = (bf16)(v * gelu_grad(zprev[o]));
            if (EPI == 3) out32[o] = v;
        };
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = 16 * nt + lr;
            if (gy < ny && co < cout) {
                const float bco = (EPI == 0 || EPI == 1 || EPI == 4) ? bias[co] : 0.f;
                if (x0 + CTX <= nx) {
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            emit(acc[mt][nt][r], co, bco, rowo + (long long)(16 * mt + 4 * lg + r) * ostride + co);
                } else {
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int xo = 16 * mt + 4 * lg + r;
                            if (x0 + xo < nx) emit(acc[mt][nt][r], co, bco, rowo + (long long)xo * ostride + co);
                        }
                }
            }
        }
        __syncthreads();
        if (more) store(z + 2);
        __syncthreads();
    }
}

// Weight gradient partials: wave w accumulates taps w, w+4, ...; the
// workgroup walks segments of DSEG voxels of one (z, y) row.
template <int CI, int CO>  // padded to multiples of 16
__global__ __launch_bounds__(256) void k_dw(const bf16 *__restrict__ x, int cis, int ci_real,
                                            const bf16 *__restrict__ g, int cos_, int co_real, int nx, int ny,
                                            int nz, float *__restrict__ partial) {
    constexpr int MT = CO / 16, NT = CI / 16, NTAP = 7;
    __shared__ __align__(16) bf16 sg[CO * GS];
    __shared__ __align__(16) bf16 sx[9 * CI * XS];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, lr = lane & 15, lg = lane >> 4;
    const int nsegx = (nx + DSEG - 1) / DSEG;
    const long long nseg = (long long)nz * ny * nsegx;
    f32x4 acc[NTAP][MT][NT];
#pragma unroll
    for (int s = 0; s < NTAP; ++s)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[s][mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (long long seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const int sx0 = (int)(seg % nsegx) * DSEG;
        const long long zy = seg / nsegx;
        const int y = (int)(zy % ny), z = (int)(zy / ny);
        __syncthreads();
        // G of the segment, transposed: sg[co][v] (16-byte loads of 8 channels)
        for (int e = t; e < DSEG * (CO / 8); e += 256) {
            const int v = e / (CO / 8), c8 = e - v * (CO / 8);
            const int gx = sx0 + v;
            bf16x8 w = bf16x8{};
            if (c8 * 8 < co_real && gx < nx)
                w = *reinterpret_cast<const bf16x8 *>(g + (((long long)z * ny + y) * nx + gx) * cos_ + c8 * 8);
#pragma unroll
            for (int q = 0; q < 8; ++q) sg[(c8 * 8 + q) * GS + v] = (c8 * 8 + q < co_real) ? w[q] : (bf16)0.f;
        }
        // X rows (z+dz, y+dy), x = sx0-1 .. sx0+DSEG, transposed: sx[row][ci][x]
        for (int e = t; e < 9 * LW * (CI / 8); e += 256) {
            const int row = e / (LW * (CI / 8)), r2 = e - row * (LW * (CI / 8));
            const int col = r2 / (CI / 8), c8 = r2 - col * (CI / 8);
            const int zz = z + row / 3 - 1, yy = wrapn(y + row % 3 - 1, ny), xx = wrapn(sx0 + col - 1, nx);
            bf16x8 w = bf16x8{};
            if (c8 * 8 < ci_real && zz >= 0 && zz < nz)
                w = *reinterpret_cast<const bf16x8 *>(x + (((long long)zz * ny + yy) * nx + xx) * cis + c8 * 8);
#pragma unroll
            for (int q = 0; q < 8; ++q) sx[(row * CI + c8 * 8 + q) * XS + col] = (c8 * 8 + q < ci_real) ? w[q] : (bf16)0.f;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NTAP; ++s) {
            const int tap = wv + 4 * s;
            if (tap >= 27) break;
            const int row = tap / 3, dx = tap % 3 - 1;  // row = 3(dz+1) + (dy+1)
#pragma unroll
            for (int kb = 0; kb < DSEG / 32; ++kb) {
                const int k0 = kb * 32 + lg * 8;
                bf16x8 a[MT], b[NT];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) a[mt] = *reinterpret_cast<const bf16x8 *>(&sg[(16 * mt + lr) * GS + k0]);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const bf16 *px = &sx[(row * CI + 16 * nt + lr) * XS + 1 + dx + k0];
#pragma unroll
                    for (int e = 0; e < 8; ++e) b[nt][e] = px[e];
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[s][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[s][mt][nt], 0, 0, 0);
            }
        }
    }
    // partial[blk][tap][co][ci]
#pragma unroll
    for (int s = 0; s < NTAP; ++s) {
        const int tap = wv + 4 * s;
        if (tap >= 27) break;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = 16 * mt + 4 * lg + r, ci = 16 * nt + lr;
                    partial[(((long long)blockIdx.x * 27 + tap) * CO + co) * CI + ci] = acc[s][mt][nt][r];
                }
    }
}

// dW[co][tap][ci] (params order) = sum over the block partials: a workgroup
// takes 32 consecutive outputs, its 8 thread groups sum the blocks c, c+8,
// c+16, ... and the 8 group sums are added in order (deterministic).
constexpr int DWR_OUT = 32, DWR_GRP = 8;
__global__ __launch_bounds__(DWR_OUT *DWR_GRP) void k_dw_reduce(const float *__restrict__ partial, int nblk, int co_pad,
                                                               int ci_pad, int co_real, int ci_real,
                                                               float *__restrict__ dw) {
    __shared__ float red[DWR_GRP][DWR_OUT];
    const int j = threadIdx.x % DWR_OUT, c = threadIdx.x / DWR_OUT;
    const int e = blockIdx.x * DWR_OUT + j;
    const int ne = co_real * 27 * ci_real;
    float s = 0.f;
    if (e < ne) {
        const int co = e / (27 * ci_real), r = e - co * 27 * ci_real, tap = r / ci_real, ci = r - tap * ci_real;
        for (int b = c; b < nblk; b += DWR_GRP) s += partial[(((long long)b * 27 + tap) * co_pad + co) * ci_pad + ci];
    }
    red[c][j] = s;
    __syncthreads();
    if (c == 0 && e < ne) {
        float t = red[0][j];
        for (int q = 1; q < DWR_GRP; ++q) t += red[q][j];
        dw[e] = t;
    }
}

// db partials: block b sums G over its voxel range.  A thread loads 8
// channels (16 bytes) at a time; the per-channel sums of the threads are
// then added in thread order (deterministic).
__global__ __launch_bounds__(256) void k_db(const bf16 *__restrict__ g, int cos_, int co_real, long long nvox,
                                            float *__restrict__ partial) {
    __shared__ float red[256 * 8];
    const int nch = cos_ / 8, vpb = 256 / nch;  // 16-byte chunks per voxel, voxels per pass
    const int t = threadIdx.x, ch = t % nch, vt = t / nch;
    const long long per = (nvox + gridDim.x - 1) / gridDim.x;
    const long long v0 = blockIdx.x * per, v1 = min(nvox, v0 + per);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (long long v = v0 + vt; v < v1; v += vpb) {
        const bf16x8 w = *reinterpret_cast<const bf16x8 *>(g + v * cos_ + ch * 8);
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] += (float)w[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) red[t * 8 + q] = acc[q];
    __syncthreads();
    if (t < co_real) {
        const int cc = t / 8, q = t % 8;
        float s = 0.f;
        for (int u = 0; u < vpb; ++u) s += red[(u * nch + cc) * 8 + q];
        partial[blockIdx.x * co_real + t] = s;
    }
}

// db = sum over the block partials: 8 thread groups sum the blocks c, c+8,
// ..., then the group sums are added in order (deterministic).
__global__ __launch_bounds__(DWR_OUT *DWR_GRP) void k_db_reduce(const float *__restrict__ partial, int nblk,
                                                               int co_real, float *__restrict__ db) {
    __shared__ float red[DWR_GRP][DWR_OUT];
    const int co = threadIdx.x % DWR_OUT, c = threadIdx.x / DWR_OUT;
    float s = 0.f;
    if (co < co_real)
        for (int b = c; b < nblk; b += DWR_GRP) s += partial[b * co_real + co];
    red[c][co] = s;
    __syncthreads();
    if (c == 0 && co < co_real) {
        float t = red[0][co];
        for (int q = 1; q < DWR_GRP; ++q) t += red[q][co];
        db[co] = t;
    }
}

// BF16 weight panels of one layer from the FP32 master W[co][tap][ci]:
// forward wf[co_pad][kpf] at k = tap*ci_str + ci; backward (input
// gradient) wb[ci_pad][kpb] at k = tap'*co_str + co with tap' = 26 - tap.
__global__ void k_pack_w(const float *__restrict__ w, int co_real, int ci_real, int ci_str, int co_str,
                         bf16 *__restrict__ wf, int co_pad, int kpf, bf16 *__restrict__ wb, int ci_pad, int kpb) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nf = (long long)co_pad * kpf, nb = (long long)ci_pad * kpb;
    if (e < nf) {
        const int co = (int)(e / kpf), k = (int)(e - (long long)co * kpf);
        const int tap = k / ci_str, ci = k - tap * ci_str;
        float v = 0.f;
        if (co < co_real && tap < 27 && ci < ci_real) v = w[((long long)co * 27 + tap) * ci_real + ci];
        wf[e] = (bf16)v;
    }
    if (e < nb) {
        const int ci = (int)(e / kpb), k = (int)(e - (long long)ci * kpb);
        const int tp = k / co_str, co = k - tp * co_str;
        float v = 0.f;
        if (ci < ci_real && tp < 27 && co < co_real) v = w[((long long)co * 27 + (26 - tp)) * ci_real + ci];
        wb[e] = (bf16)v;
    }
}

// MSE loss of y against t (first cout channels of stride 8) and its
// gradient g = 2 (y - t) / (n cout) as BF16 with stride 8 (channel >= cout
// zero); per-block partial sums of the squared error.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y, const float *__restrict__ tgt, int cout,
                                              long long nvox, float gscale, bf16 *__restrict__ g,
                                              double *__restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nvox * 8;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e & 7);
        float d = 0.f;
        if (c < cout) {
            d = y[e] - tgt[e];
            s += (double)d * (double)d;
        }
        if (g) g[e] = (bf16)(gscale * d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gr = grad[e];
    const float mm = b1 * m[e] + (1.f - b1) * gr;
    const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_f32_to_bf16(const float *__restrict__ a, bf16 *__restrict__ b, long long n, int cin, int cstr) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * cstr) return;
    const long long v = e / cstr;
    const int c = (int)(e - v * cstr);
    b[e] = (bf16)(c < cin ? a[v * cin + c] : 0.f);
}

struct Sample {
    int nx, ny, nz;
    bf16 *x;   // [v][8]
    float *t;  // [v][8]
};

struct Layer {
    int ci, co;          // real channel counts
    int cis, cos_;       // storage strides of input / output activations (multiples of 8)
    int co_pad, ci_pad;  // MFMA paddings (multiples of 16)
    int kpf, kpb;
    long long woff, boff;  // offsets into the parameter vector
    bf16 *wf = nullptr, *wb = nullptr;
};

}  // namespace

struct mhd_net {
    int dev = 0;
    hipStream_t stream = nullptr;
    int cin = 0, width = 0, cout = 0;
    long long nparam = 0;
    Layer L[3];
    float *p = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    int adam_t = 0;
    // workspace for volumes up to cap voxels
    long long cap = 0;
    bf16 *a0 = nullptr, *a1 = nullptr, *a2 = nullptr, *g1 = nullptr, *g2 = nullptr, *g3 = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *tmp32 = nullptr, *dx32 = nullptr;
    float *partial = nullptr;
    double *lpart = nullptr;
    int nblk_dw = 512;
    std::vector<Sample> samples;
    char err[512] = {0};
};

static char g_net_err[512];

static int nfail(mhd_net_t *n, int code, const char *fmt, ...) {
    char *buf = n ? n->err : g_net_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define NHIP(n, x)                                                                         \
    do {                                                                                   \
        hipError_t r_ = (x);                                                               \
        if (r_ != hipSuccess)                                                              \
            return nfail((n), MHD_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(r_),    \
                         __FILE__, __LINE__);                                              \
    } while (0)

// z chunk per workgroup so that the grid has about `target` workgroups.
static int kchunk_for(int nx, int ny, int nz, int target, int *nblocks) {
    const int tiles = ((nx + CTX - 1) / CTX) * ((ny + CTY - 1) / CTY);
    int nc = (target + tiles - 1) / tiles;
    if (nc > nz) nc = nz;
    if (nc < 1) nc = 1;
    const int kc = (nz + nc - 1) / nc;
    *nblocks = tiles * ((nz + kc - 1) / kc);
    return kc;
}

template <int CIN, int NT, int EPI>
static hipError_t conv(hipStream_t s, const bf16 *x, const bf16 *w, int kp, const float *bias, const float *zprev,
                       bf16 *out16, float *out32, float *zout, int cout, int ostride, int nx, int ny, int nz) {
    const int lds = 3 * LW * LH * cstride<CIN>() * (int)sizeof(bf16);
    if (kp != 32 * ((27 * CIN + 31) / 32)) return hipErrorInvalidValue;  // the kernel's K blocks
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if ((long long)nx * ny * CIN > 0x7fffffffLL) return hipErrorInvalidValue;  // in-plane offsets are 32-bit
    static std::atomic<unsigned long long> attr{0};
    hipError_t e = rmhd::lds_limit_once(attr, reinterpret_cast<const void *>(&k_conv<CIN, NT, EPI>), lds);
    if (e != hipSuccess) return e;
    // as many workgroups as fit the CUs at once (LDS and registers): each
    // streams a longer z chunk, which amortises its prologue (three planes +
    // the weights); per device
    static std::atomic<int> resident[64];
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    int res = resident[dev].load(std::memory_order_relaxed);
    if (!res) {
        int per_cu = 0, ncu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_conv<CIN, NT, EPI>),
                                                         256, lds);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        res = std::max(1, per_cu) * std::max(1, ncu);
        resident[dev].store(res, std::memory_order_relaxed);
    }
    int nb;
    const int kc = kchunk_for(nx, ny, nz, res, &nb);
    hipLaunchKernelGGL((k_conv<CIN, NT, EPI>), dim3(nb), dim3(256), lds, s, x, w, kp, bias, zprev, out16, out32, zout,
                       cout, ostride, nx, ny, nz, kc);
    return hipGetLastError();
}

static int ensure_workspace(mhd_net_t *n, long long nvox) {
    if (nvox <= n->cap) return MHD_OK;
    void **bufs[11] = {(void **)&n->a0, (void **)&n->a1, (void **)&n->a2, (void **)&n->g1, (void **)&n->g2,
                       (void **)&n->g3, (void **)&n->z1, (void **)&n->z2, (void **)&n->y, (void **)&n->tmp32,
                       (void **)&n->dx32};
    for (auto b : bufs)
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    const size_t w = (size_t)n->width;
    NHIP(n, hipMalloc(&n->a0, nvox * 8 * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->a1, nvox * w * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->a2, nvox * w * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->g1, nvox * w * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->g2, nvox * w * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->g3, nvox * 8 * sizeof(bf16)));
    NHIP(n, hipMalloc(&n->z1, nvox * w * sizeof(float)));
    NHIP(n, hipMalloc(&n->z2, nvox * w * sizeof(float)));
    NHIP(n, hipMalloc(&n->y, nvox * 8 * sizeof(float)));
    NHIP(n, hipMalloc(&n->tmp32, nvox * 8 * sizeof(float)));
    NHIP(n, hipMalloc(&n->dx32, nvox * 8 * sizeof(float)));
    n->cap = nvox;
    return MHD_OK;
}

static hipError_t pack_weights(mhd_net_t *n) {
    for (int l = 0; l < 3; ++l) {
        const Layer &L = n->L[l];
        const long long tot = std::max((long long)L.co_pad * L.kpf, (long long)L.ci_pad * L.kpb);
        hipLaunchKernelGGL(k_pack_w, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, n->stream, n->p + L.woff, L.co,
                           L.ci, L.cis, L.cos_, L.wf, L.co_pad, L.kpf, L.wb, L.ci_pad, L.kpb);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// forward of a BF16 input x (stride 8) into n->y (FP32, stride 8) on stream s;
// keep: also store the pre-activations z1, z2 the backward pass reads
static hipError_t forward_on(mhd_net_t *n, hipStream_t s, const bf16 *x, int nx, int ny, int nz, bool keep) {
    const Layer *L = n->L;
    hipError_t e = keep ? conv<8, 2, 1>(s, x, L[0].wf, L[0].kpf, n->p + L[0].boff, nullptr, n->a1, nullptr, n->z1,
                                        L[0].co, L[0].cos_, nx, ny, nz)
                        : conv<8, 2, 4>(s, x, L[0].wf, L[0].kpf, n->p + L[0].boff, nullptr, n->a1, nullptr, nullptr,
                                        L[0].co, L[0].cos_, nx, ny, nz);
    if (e == hipSuccess)
        e = keep ? conv<32, 2, 1>(s, n->a1, L[1].wf, L[1].kpf, n->p + L[1].boff, nullptr, n->a2, nullptr, n->z2,
                                  L[1].co, L[1].cos_, nx, ny, nz)
                 : conv<32, 2, 4>(s, n->a1, L[1].wf, L[1].kpf, n->p + L[1].boff, nullptr, n->a2, nullptr, nullptr,
                                  L[1].co, L[1].cos_, nx, ny, nz);
    if (e == hipSuccess)
        e = conv<32, 1, 0>(s, n->a2, L[2].wf, L[2].kpf, n->p + L[2].boff, nullptr, nullptr, n->y, nullptr, L[2].co, 8,
                           nx, ny, nz);
    return e;
}

static hipError_t forward(mhd_net_t *n, int nx, int ny, int nz, bool keep) {
    return forward_on(n, n->stream, n->a0, nx, ny, nz, keep);
}

template <int CI, int CO>
static hipError_t dw(mhd_net_t *n, const bf16 *x, const Layer &L, const bf16 *g, int nx, int ny, int nz) {
    const long long nseg = (long long)nz * ny * ((nx + DSEG - 1) / DSEG);
    const int nblk = (int)std::min<long long>(n->nblk_dw, nseg);
    hipLaunchKernelGGL((k_dw<CI, CO>), dim3(nblk), dim3(256), 0, n->stream, x, L.cis, L.ci, g, L.cos_, L.co, nx, ny, nz,
                       n->partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int ne = L.co * 27 * L.ci;
    hipLaunchKernelGGL(k_dw_reduce, dim3((ne + DWR_OUT - 1) / DWR_OUT), dim3(DWR_OUT * DWR_GRP), 0, n->stream,
                       n->partial, nblk, CO, CI, L.co, L.ci, n->grad + L.woff);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long nvox = (long long)nx * ny * nz;
    const int nb = (int)std::min<long long>(256, (nvox + 255) / 256);
    hipLaunchKernelGGL(k_db, dim3(nb), dim3(256), 0, n->stream, g, L.cos_, L.co, nvox, n->partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_db_reduce, dim3(1), dim3(DWR_OUT * DWR_GRP), 0, n->stream, n->partial, nb, L.co,
                       n->grad + L.boff);
    return hipGetLastError();
}

// backward from g3 (BF16, stride 8): parameter gradients into grad; with
// want_dx the input gradient (FP32, stride 8) into dx32.
static hipError_t backward(mhd_net_t *n, int nx, int ny, int nz, bool want_dx) {
    const Layer *L = n->L;
    hipError_t e = dw<32, 16>(n, n->a2, L[2], n->g3, nx, ny, nz);
    if (e == hipSuccess)
        e = conv<8, 2, 2>(n->stream, n->g3, L[2].wb, L[2].kpb, nullptr, n->z2, n->g2, nullptr, nullptr, L[2].ci,
                          L[2].cis, nx, ny, nz);
    if (e == hipSuccess) e = dw<32, 32>(n, n->a1, L[1], n->g2, nx, ny, nz);
    if (e == hipSuccess)
        e = conv<32, 2, 2>(n->stream, n->g2, L[1].wb, L[1].kpb, nullptr, n->z1, n->g1, nullptr, nullptr, L[1].ci,
                           L[1].cis, nx, ny, nz);
    if (e == hipSuccess) e = dw<16, 32>(n, n->a0, L[0], n->g1, nx, ny, nz);
    if (e == hipSuccess && want_dx)
        e = conv<32, 1, 3>(n->stream, n->g1, L[0].wb, L[0].kpb, nullptr, nullptr, nullptr, n->dx32, nullptr, L[0].ci,
                           8, nx, ny, nz);
    return e;
}

static int upload_input(mhd_net_t *n, const float *x, long long nvox, bf16 *dst) {
    NHIP(n, hipMemcpyAsync(n->tmp32, x, nvox * n->cin * sizeof(float), hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((nvox * 8 + 255) / 256)), dim3(256), 0, n->stream, n->tmp32, dst,
                       nvox, n->cin, 8);
    NHIP(n, hipGetLastError());
    return MHD_OK;
}

static int check_dims(mhd_net_t *n, int nx, int ny, int nz, const char *who) {
    if (!n) return nfail(nullptr, MHD_ERR_ARG, "%s: null handle", who);
    if (nx < 1 || ny < 1 || nz < 1) return nfail(n, MHD_ERR_ARG, "%s: volume %dx%dx%d", who, nx, ny, nz);
    return MHD_OK;
}

extern "C" int mhd_net_create(int cin, int width, int cout, mhd_net_t **out) {
    if (!out) return nfail(nullptr, MHD_ERR_ARG, "mhd_net_create: null argument");
    *out = nullptr;
    if (cin != 8 || width != 32 || cout < 1 || cout > 8)
        return nfail(nullptr, MHD_ERR_SCOPE, "mhd_net_create: built for cin=8, width=32, cout<=8 (got %d, %d, %d)",
                     cin, width, cout);
    mhd_net_t *n = new (std::nothrow) mhd_net();
    if (!n) return nfail(nullptr, MHD_ERR_ARG, "mhd_net_create: out of host memory");
    n->cin = cin;
    n->width = width;
    n->cout = cout;
    const int ci[3] = {cin, width, width}, co[3] = {width, width, cout};
    const int cis[3] = {8, width, width}, cos_[3] = {width, width, 8};
    long long off = 0;
    for (int l = 0; l < 3; ++l) {
        Layer &L = n->L[l];
        L.ci = ci[l];
        L.co = co[l];
        L.cis = cis[l];
        L.cos_ = cos_[l];
        L.co_pad = (co[l] + 15) / 16 * 16;
        L.ci_pad = (ci[l] + 15) / 16 * 16;
        L.kpf = (27 * L.cis + 31) / 32 * 32;
        L.kpb = (27 * L.cos_ + 31) / 32 * 32;
        L.woff = off;
        off += (long long)L.co * 27 * L.ci;
        L.boff = off;
        off += L.co;
    }
    n->nparam = off;
    hipError_t e = hipGetDevice(&n->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    for (float **b : {&n->p, &n->grad, &n->am, &n->av})
        if (e == hipSuccess) e = hipMalloc(b, off * sizeof(float));
    for (int l = 0; l < 3 && e == hipSuccess; ++l) {
        Layer &L = n->L[l];
        e = hipMalloc(&L.wf, (size_t)L.co_pad * L.kpf * sizeof(bf16));
        if (e == hipSuccess) e = hipMalloc(&L.wb, (size_t)L.ci_pad * L.kpb * sizeof(bf16));
    }
    if (e == hipSuccess) e = hipMalloc(&n->partial, (size_t)n->nblk_dw * 27 * 32 * 32 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&n->lpart, 1024 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(n->p, 0, off * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->am, 0, off * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->av, 0, off * sizeof(float));
    if (e != hipSuccess) {
        const int rc = nfail(nullptr, MHD_ERR_HIP, "mhd_net_create: %s", hipGetErrorString(e));
        mhd_net_destroy(n);
        return rc;
    }
    *out = n;
    return MHD_OK;
}

extern "C" long long mhd_net_num_params(const mhd_net_t *n) { return n ? n->nparam : 0; }

extern "C" int mhd_net_set_params(mhd_net_t *n, const float *params) {
    if (!n || !params) return nfail(n, MHD_ERR_ARG, "mhd_net_set_params: null argument");
    NHIP(n, hipSetDevice(n->dev));
    NHIP(n, hipMemcpyAsync(n->p, params, n->nparam * sizeof(float), hipMemcpyHostToDevice, n->stream));
    NHIP(n, hipMemsetAsync(n->am, 0, n->nparam * sizeof(float), n->stream));
    NHIP(n, hipMemsetAsync(n->av, 0, n->nparam * sizeof(float), n->stream));
    n->adam_t = 0;
    NHIP(n, pack_weights(n));
    NHIP(n, hipStreamSynchronize(n->stream));
    return MHD_OK;
}

extern "C" int mhd_net_get_params(mhd_net_t *n, float *params) {
    if (!n || !params) return nfail(n, MHD_ERR_ARG, "mhd_net_get_params: null argument");
    NHIP(n, hipSetDevice(n->dev));
    NHIP(n, hipMemcpyAsync(params, n->p, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    return MHD_OK;
}

extern "C" int mhd_net_forward(mhd_net_t *n, int nx, int ny, int nz, const float *x, float *y) {
    int rc = check_dims(n, nx, ny, nz, "mhd_net_forward");
    if (rc) return rc;
    if (!x || !y) return nfail(n, MHD_ERR_ARG, "mhd_net_forward: null array");
    const long long nvox = (long long)nx * ny * nz;
    NHIP(n, hipSetDevice(n->dev));
    rc = ensure_workspace(n, nvox);
    if (rc) return rc;
    rc = upload_input(n, x, nvox, n->a0);
    if (rc) return rc;
    NHIP(n, forward(n, nx, ny, nz, false));
    // y (stride 8) -> host (stride cout)
    std::vector<float> tmp((size_t)nvox * 8);
    NHIP(n, hipMemcpyAsync(tmp.data(), n->y, nvox * 8 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    for (long long v = 0; v < nvox; ++v)
        for (int c = 0; c < n->cout; ++c) y[v * n->cout + c] = tmp[(size_t)v * 8 + c];
    return MHD_OK;
}

extern "C" int mhd_net_gradient(mhd_net_t *n, int nx, int ny, int nz, const float *x, const float *dy,
                                float *dparams, float *dx) {
    int rc = check_dims(n, nx, ny, nz, "mhd_net_gradient");
    if (rc) return rc;
    if (!x || !dy || !dparams) return nfail(n, MHD_ERR_ARG, "mhd_net_gradient: null array");
    const long long nvox = (long long)nx * ny * nz;
    NHIP(n, hipSetDevice(n->dev));
    rc = ensure_workspace(n, nvox);
    if (rc) return rc;
    rc = upload_input(n, x, nvox, n->a0);
    if (rc) return rc;
    NHIP(n, forward(n, nx, ny, nz, true));
    std::vector<float> g8((size_t)nvox * 8, 0.f);
    for (long long v = 0; v < nvox; ++v)
        for (int c = 0; c < n->cout; ++c) g8[(size_t)v * 8 + c] = dy[v * n->cout + c];
    NHIP(n, hipMemcpyAsync(n->tmp32, g8.data(), g8.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((nvox * 8 + 255) / 256)), dim3(256), 0, n->stream, n->tmp32,
                       n->g3, nvox, 8, 8);
    NHIP(n, hipGetLastError());
    NHIP(n, backward(n, nx, ny, nz, dx != nullptr));
    NHIP(n, hipMemcpyAsync(dparams, n->grad, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    if (dx) {
        NHIP(n, hipMemcpyAsync(g8.data(), n->dx32, nvox * 8 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    }
    NHIP(n, hipStreamSynchronize(n->stream));
    if (dx)
        for (long long v = 0; v < nvox; ++v)
            for (int c = 0; c < n->cin; ++c) dx[v * n->cin + c] = g8[(size_t)v * 8 + c];
    return MHD_OK;
}

extern "C" int mhd_net_add_sample(mhd_net_t *n, int nx, int ny, int nz, const float *x, const float *t) {
    int rc = check_dims(n, nx, ny, nz, "mhd_net_add_sample");
    if (rc) return rc;
    if (!x || !t) return nfail(n, MHD_ERR_ARG, "mhd_net_add_sample: null array");
    const long long nvox = (long long)nx * ny * nz;
    NHIP(n, hipSetDevice(n->dev));
    rc = ensure_workspace(n, nvox);
    if (rc) return rc;
    Sample s{nx, ny, nz, nullptr, nullptr};
    NHIP(n, hipMalloc(&s.x, nvox * 8 * sizeof(bf16)));
    hipError_t e = hipMalloc(&s.t, nvox * 8 * sizeof(float));
    if (e != hipSuccess) {
        (void)hipFree(s.x);
        return nfail(n, MHD_ERR_HIP, "mhd_net_add_sample: %s", hipGetErrorString(e));
    }
    n->samples.push_back(s);
    rc = upload_input(n, x, nvox, s.x);
    if (rc) return rc;
    std::vector<float> t8((size_t)nvox * 8, 0.f);
    for (long long v = 0; v < nvox; ++v)
        for (int c = 0; c < n->cout; ++c) t8[(size_t)v * 8 + c] = t[v * n->cout + c];
    NHIP(n, hipMemcpyAsync(s.t, t8.data(), t8.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    NHIP(n, hipStreamSynchronize(n->stream));
    return MHD_OK;
}

extern "C" int mhd_net_num_samples(const mhd_net_t *n) { return n ? (int)n->samples.size() : 0; }

extern "C" int mhd_net_clear_samples(mhd_net_t *n) {
    if (!n) return nfail(nullptr, MHD_ERR_ARG, "mhd_net_clear_samples: null handle");
    NHIP(n, hipSetDevice(n->dev));
    NHIP(n, hipStreamSynchronize(n->stream));
    for (auto &s : n->samples) {
        (void)hipFree(s.x);
        (void)hipFree(s.t);
    }
    n->samples.clear();
    return MHD_OK;
}

// loss of the current parameters on sample s (forward only); with train,
// one Adam step on it.
static int sample_step(mhd_net_t *n, int s, bool train, float lr, double *loss) {
    const Sample &S = n->samples[s];
    const long long nvox = (long long)S.nx * S.ny * S.nz;
    NHIP(n, hipMemcpyAsync(n->a0, S.x, nvox * 8 * sizeof(bf16), hipMemcpyDeviceToDevice, n->stream));
    NHIP(n, forward(n, S.nx, S.ny, S.nz, train));
    const int nb = (int)std::min<long long>(1024, (nvox * 8 + 255) / 256);
    const float gscale = (float)(2.0 / ((double)nvox * n->cout));
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, n->stream, n->y, S.t, n->cout, nvox, gscale,
                       train ? n->g3 : nullptr, n->lpart);
    NHIP(n, hipGetLastError());
    if (train) {
        NHIP(n, backward(n, S.nx, S.ny, S.nz, false));
        ++n->adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->p, n->grad,
                           n->am, n->av, n->nparam, lr, b1, b2, 1e-8f, c1, c2);
        NHIP(n, hipGetLastError());
        NHIP(n, pack_weights(n));
    }
    if (loss) {
        std::vector<double> lp(nb);
        NHIP(n, hipMemcpyAsync(lp.data(), n->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, n->stream));
        NHIP(n, hipStreamSynchronize(n->stream));
        double s2 = 0;
        for (double v : lp) s2 += v;
        *loss = s2 / ((double)nvox * n->cout);
    }
    return MHD_OK;
}

extern "C" int mhd_net_train(mhd_net_t *n, int iters, float lr, double *loss_first, double *loss_last) {
    if (!n) return nfail(nullptr, MHD_ERR_ARG, "mhd_net_train: null handle");
    if (n->samples.empty()) return nfail(n, MHD_ERR_STATE, "mhd_net_train: no samples");
    if (iters < 1 || !(lr > 0.f)) return nfail(n, MHD_ERR_ARG, "mhd_net_train: iters %d, lr %g", iters, lr);
    NHIP(n, hipSetDevice(n->dev));
    for (int it = 0; it < iters; ++it) {
        double *lp = (it == 0) ? loss_first : (it == iters - 1 ? loss_last : nullptr);
        const int rc = sample_step(n, it % (int)n->samples.size(), true, lr, lp);
        if (rc) return rc;
    }
    if (iters == 1 && loss_last && loss_first) *loss_last = *loss_first;
    NHIP(n, hipStreamSynchronize(n->stream));
    return MHD_OK;
}

extern "C" int mhd_net_loss(mhd_net_t *n, int s, double *loss) {
    if (!n) return nfail(nullptr, MHD_ERR_ARG, "mhd_net_loss: null handle");
    if (s < 0 || s >= (int)n->samples.size()) return nfail(n, MHD_ERR_ARG, "mhd_net_loss: sample %d", s);
    NHIP(n, hipSetDevice(n->dev));
    return sample_step(n, s, false, 0.f, loss);
}

extern "C" int mhd_net_time_forward(mhd_net_t *n, int nx, int ny, int nz, int reps, double *ms) {
    int rc = check_dims(n, nx, ny, nz, "mhd_net_time_forward");
    if (rc) return rc;
    if (reps < 1) return nfail(n, MHD_ERR_ARG, "mhd_net_time_forward: reps %d", reps);
    NHIP(n, hipSetDevice(n->dev));
    rc = ensure_workspace(n, (long long)nx * ny * nz);
    if (rc) return rc;
    NHIP(n, forward(n, nx, ny, nz, false));  // warm-up (also sets kernel attributes)
    hipEvent_t a, b;
    NHIP(n, hipEventCreate(&a));
    NHIP(n, hipEventCreate(&b));
    NHIP(n, hipEventRecord(a, n->stream));
    for (int r = 0; r < reps; ++r) NHIP(n, forward(n, nx, ny, nz, false));
    NHIP(n, hipEventRecord(b, n->stream));
    NHIP(n, hipEventSynchronize(b));
    float t = 0;
    NHIP(n, hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = t / reps;
    return MHD_OK;
}

extern "C" const char *mhd_net_last_error(const mhd_net_t *n) { return n ? n->err : g_net_err; }

extern "C" void mhd_net_destroy(mhd_net_t *n) {
    if (!n) return;
    (void)hipSetDevice(n->dev);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (auto &s : n->samples) {
        (void)hipFree(s.x);
        (void)hipFree(s.t);
    }
    void *bufs[] = {n->p, n->grad, n->am, n->av, n->a0, n->a1, n->a2, n->g1, n->g2, n->g3, n->z1, n->z2,
                    n->y, n->tmp32, n->dx32, n->partial, n->lpart, n->L[0].wf, n->L[0].wb, n->L[1].wf,
                    n->L[1].wb, n->L[2].wf, n->L[2].wb};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

// Device-side entry points for the closure-corrected step (mhd_capi.hip).
namespace rmhd {

int net_shape(const mhd_net_t *n, int *cin, int *cout) {
    if (!n) return MHD_ERR_ARG;
    *cin = n->cin;
    *cout = n->cout;
    return MHD_OK;
}

// Workspace for volumes of nvox points and the device the network lives on.
int net_prepare(mhd_net_t *n, long long nvox, int *dev) {
    *dev = n->dev;
    const int rc = ensure_workspace(n, nvox);
    if (rc) return rc;
    return hipStreamSynchronize(n->stream) == hipSuccess ? MHD_OK : MHD_ERR_HIP;
}

// y = net(x) for a BF16 input (stride 8) resident on the device, queued on
// stream s; *y points to the output (FP32, stride 8) in the network's
// workspace, valid until the next use of the network.
hipError_t net_forward_device(mhd_net_t *n, hipStream_t s, int nx, int ny, int nz, const __bf16 *x,
                              const float **y) {
    *y = n->y;
    return forward_on(n, s, x, nx, ny, nz, false);
}

}  // namespace rmhd
