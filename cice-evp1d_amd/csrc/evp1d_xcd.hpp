// evp1d_xcd.hpp - workgroup-to-tile mapping for the 8 XCDs of the chip.
//
// Workgroups of one launch are dispatched round-robin over the 8 XCDs (workgroup
// b runs on XCD b % 8), and each XCD has its own L2.  With `xcd` set, the
// workgroup with hardware index b processes logical tile L(b), a bijection that
// hands each XCD one contiguous eighth of the tiles, so the neighbour rows a
// tile gathers from were streamed through the same L2.
#pragma once

#include <hip/hip_runtime.h>

constexpr int kXcds = 8;
__device__ __forceinline__ int tile_of(int b, int n, bool xcd) {
  if (!xcd) return b;
  const int q = n / kXcds, r = n % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
