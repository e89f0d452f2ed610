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

// Entry i of a read-only per-plane table (grid metrics, background
// profiles) at a workgroup-uniform index: read through the constant
// address space, so the compiler issues a scalar load.  A vector load here
// would share the in-order vector-memory counter with the plane prefetch
// and the wait for it would also wait for every load issued before it.
// The tables are written only between kernels (host copies, or earlier
// kernels on the stream, such as the plane means of the optional branches).
template <class T>
__device__ __forceinline__ T uload(const T *p, int i) {
    return ((const __attribute__((address_space(4))) T *)p)[i];
}


// x, held back until y has been computed: an empty asm that reads y and
// may change x, so that nothing derived from x is scheduled before y.  Used
// to keep the first use of a loaded value (and so the wait for the load,
// which also waits for every load issued before it) late in a plane.
template <class X, class Y>
__device__ __forceinline__ X after(X x, Y y) {
    asm volatile("" : "+v"(x) : "v"(y));
    return x;
}

// x is needed here: an empty asm that reads x, so the wait for the load of
// x happens at this point (once, before a loop) rather than at its first use
// inside the loop, where it would be repeated on every iteration.
template <class X>
__device__ __forceinline__ void settle(X x) {
    asm volatile("" ::"v"(x));
}

// Stored values of the mixed form (mhd_mixed.hip) as other kernels read them.

// 1/x of the FP32 perturbation arithmetic by the hardware reciprocal
// (v_rcp_f32, within 1 ulp) instead of the correctly rounded division
// sequence (about ten instructions).  The mixed form is judged by its error
// against the FP64 form, where one ulp of 1/rho is far below the FP32
// rounding of the stored values; every use (stencil, packing, unpacking,
// boundary planes) goes through this function, so they stay consistent.
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// T' represented by a stored (rho', P') pair: (P' - T0 rho') / (rho0 + rho')
// in FP32.  Every kernel that needs T' of a stored pair calls this function
// (the stencil's ring included), so the value the residual compensates
// against is bit-identical to the value the right-hand side sees.
__device__ __forceinline__ float trep(float r0, float t0, float rh, float pp) {
    const float ri = rcp32(r0 + rh);
    return (pp - t0 * rh) * ri;
}

// BF16 bit patterns of the resting residual R: round to nearest even
// (v_cvt_pk_bf16_f32; NaN stays NaN), and back (exact).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned short to_bf16(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
// Two residuals in one 32-bit word: lo in the low half, hi in the high half.
__device__ __forceinline__ unsigned pack_bf16x2(float lo, float hi) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t));
}
__host__ __device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// R packs the BF16 residuals of two fields per 32-bit word: word
// (f/2)*fs + o holds field f in its low (f even) or high (f odd) half.
__device__ __forceinline__ float r_half(unsigned w, int f) {
    return from_bf16((unsigned short)((f & 1) ? (w >> 16) : (w & 0xffffu)));
}
__device__ __forceinline__ float r_get(const unsigned *__restrict__ rr, long long fs, int f, long long o) {
    return r_half(rr[(f >> 1) * fs + o], f);
}
__device__ __forceinline__ void r_put(unsigned *__restrict__ rr, long long fs, int f, long long o, float r) {
    unsigned *const w = &rr[(f >> 1) * fs + o];
    const unsigned b = to_bf16(r);
    *w = (f & 1) ? ((*w & 0x0000ffffu) | (b << 16)) : ((*w & 0xffff0000u) | b);
}

}  // namespace rmhd
