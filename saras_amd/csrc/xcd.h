// Workgroup order for gfx950's 8 XCDs.
//
// The dispatcher deals workgroups round-robin to the XCDs, each with its own L2: workgroup b runs on
// XCD b % 8.  A stencil kernel whose consecutive workgroups touch neighbouring rows and planes then
// fetches every neighbour plane into several L2s.  xcd_block() gives the logical block that
// workgroup blockIdx.x should process so that each XCD gets one contiguous range of logical blocks
// (the last gridDim.x % 8 workgroups keep their own index).  Indexing every output and every
// per-block partial by the logical block leaves a kernel's results unchanged.
#ifndef SR_XCD_H
#define SR_XCD_H

#include <hip/hip_runtime.h>

namespace sr {

__device__ inline unsigned xcd_block() {
    unsigned nb = gridDim.x, b = blockIdx.x;
    unsigned per = nb >> 3;
    if (b < per * 8u) return (b & 7u) * per + (b >> 3);
    return b;
}

}  // namespace sr

#endif
