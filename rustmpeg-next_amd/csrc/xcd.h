// xcd.h - workgroup order for the 8 XCDs of an MI355X (device code only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rmh {

// Workgroup index -> work index such that the workgroups one XCD runs
// (hardware ids k, k + 8, k + 16, ... of a 1-D grid go to XCD k) take a
// contiguous run of work indices: neighbouring tiles then share that XCD's L2.
// A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_local_block(uint32_t bid, uint32_t nb) {
    const uint32_t x = bid & 7, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

}  // namespace rmh
