// pf_launch.h — launch-shape helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

// MI355X has 8 XCDs, each with its own L2; consecutive workgroups of a
// dispatch are placed round-robin over the XCDs.  xcd_block() renumbers
// workgroups so that each XCD receives one contiguous range of work items
// (neighbouring image tiles touch the same cells, so their cell data stays
// in one XCD's L2).  Workgroups past the last full round keep their id.
constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int per = nb / kNumXcd;
    if (b >= per * kNumXcd) return b;
    return (b % kNumXcd) * per + (b / kNumXcd);
}

}  // namespace pf
