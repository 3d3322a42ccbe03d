// scan.hpp -- one-block exclusive scan used by the counting sorts (cells,
// mesh column bins): start[i] = base + sum_{j<i} count[j], start[n] = base +
// total; count and fill are cleared for the next use.  Call with all threads
// of a block (<= 1024 threads, multiple of 64).
#pragma once
#include <hip/hip_runtime.h>

namespace mmamd {

__device__ inline void block_scan_clear(int *count, int *start, int *fill, int n, int base) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = base;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int b = 0; b < n; b += blockDim.x) {
    const int c = b + threadIdx.x;
    const int v = c < n ? count[c] : 0;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
      int w = lane < nw ? wsum[lane] : 0;
      for (int o = 1; o < 16; o <<= 1) {
        const int y = __shfl_up(w, o, 64);
        if (lane >= o) w += y;
      }
      if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    const int excl = carry + (wid > 0 ? wsum[wid - 1] : 0) + x - v;
    if (c < n) {
      start[c] = excl;
      count[c] = 0;
      fill[c] = 0;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) start[n] = carry;
}

}  // namespace mmamd
