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


}  // namespace rmhd
