// reduce.hpp -- deterministic cross-block reductions.
//
// Each block writes NV partial sums; the last block to finish (atomic
// counter per system) sums all partials with every thread taking a strided
// share and a fixed-shape tree, so the result does not depend on block
// scheduling.  Slots listed as "top-2" (largest and second-largest value)
// are merged instead of summed.
#pragma once
#include <hip/hip_runtime.h>

namespace mmamd {

// merge two (largest, second largest) pairs
__device__ __forceinline__ void top2_merge(double &a1, double &a2, double b1, double b2) {
  if (b1 > a1) { a2 = fmax(a1, b2); a1 = b1; }
  else a2 = fmax(a2, b1);
}

__device__ __forceinline__ double load_coherent(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide reduction of NV values per thread (sum for c < NSUM, top-2 on
// the pair (NSUM, NSUM+1) when NV == NSUM + 2).  Result valid in thread 0.
template <int NV, int NSUM, int TB>
__device__ __forceinline__ void block_reduce(double *v, double (*red)[NV]) {
  for (int o = 32; o >= 1; o >>= 1) {
    for (int c = 0; c < NSUM; c++) v[c] += __shfl_xor(v[c], o);
    if (NV > NSUM) {
      const double b1 = __shfl_xor(v[NSUM], o), b2 = __shfl_xor(v[NSUM + 1], o);
      top2_merge(v[NSUM], v[NSUM + 1], b1, b2);
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < NV; c++) red[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < NV; c++) v[c] = red[0][c];
    for (int ww = 1; ww < TB / 64; ww++) {
      for (int c = 0; c < NSUM; c++) v[c] += red[ww][c];
      if (NV > NSUM) top2_merge(v[NSUM], v[NSUM + 1], red[ww][NSUM], red[ww][NSUM + 1]);
    }
  }
  __syncthreads();
}

// Store this block's partials only (thread 0 keeps the block's values in v).
template <int NV, int NSUM, int TB>
__device__ void store_partials(double *v, double *part, int blk, double (*red)[NV]) {
  block_reduce<NV, NSUM, TB>(v, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < NV; c++) part[(long)blk * NV + c] = v[c];
}

// Store this block's partials and, in the last block, reduce all nblk
// partials into out[0..NV).  Returns true in the thread that holds the
// final values (thread 0 of the last block).
template <int NV, int NSUM, int TB>
__device__ bool finish_partials(double *v, double *part, int nblk, int blk, int *counter,
                                double (*red)[NV], int *flag) {
  block_reduce<NV, NSUM, TB>(v, red);
  if (threadIdx.x == 0) {
    for (int c = 0; c < NV; c++) part[(long)blk * NV + c] = v[c];
    __threadfence();
    *flag = (atomicAdd(counter, 1) == nblk - 1);
  }
  __syncthreads();
  if (!*flag) return false;
  __threadfence();
  double t[NV];
  for (int c = 0; c < NV; c++) t[c] = 0;
  for (int b = threadIdx.x; b < nblk; b += TB) {
    const double *p = part + (long)b * NV;
    for (int c = 0; c < NSUM; c++) t[c] += load_coherent(p + c);
    if (NV > NSUM) top2_merge(t[NSUM], t[NSUM + 1], load_coherent(p + NSUM), load_coherent(p + NSUM + 1));
  }
  block_reduce<NV, NSUM, TB>(t, red);
  if (threadIdx.x == 0) {
    for (int c = 0; c < NV; c++) v[c] = t[c];
    *counter = 0;
    return true;
  }
  return false;
}

}  // namespace mmamd
