// Workgroup placement over the MI355X's 8 XCDs (device code).
#pragma once

#include <hip/hip_runtime.h>

namespace bem {

// Workgroup id -> logical id, so that consecutive logical ids land on the same
// XCD (each XCD has its own L2): dispatch places workgroup g on XCD g % 8, and
// the logical ids of XCD x are one contiguous range, taken in dispatch order.
__device__ __forceinline__ int xcd_remap(int g, int n) {
    const int per = n / 8, rem = n % 8;
    const int xcd = g % 8, idx = g / 8;
    return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

}  // namespace bem
