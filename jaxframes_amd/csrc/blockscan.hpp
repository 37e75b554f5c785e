// Workgroup-wide scans over LDS arrays (wave64).
#pragma once

#include "common.hpp"

namespace jfx {

// Inclusive scan of v across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += y;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint64_t y = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += y;
  }
  return v;
}

// In-place exclusive scan of s[0..N) (LDS) by a BLOCK-thread workgroup.
// Thread t owns the PER consecutive entries starting at t*PER. s_tmp needs
// BLOCK/64 + 1 words. Returns the total. Ends with a barrier.
template <int N, int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan_lds(uint32_t* s, uint32_t* s_tmp) {
  constexpr int PER = (N + BLOCK - 1) / BLOCK;
  constexpr int W = BLOCK / 64;
  const int tid = threadIdx.x;
  const int wid = tid >> 6;
  const uint32_t lane = lane_id();
  uint32_t v[PER];
  uint32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    int idx = tid * PER + k;
    v[k] = idx < N ? s[idx] : 0u;
    tsum += v[k];
  }
  uint32_t incl = wave_incl_scan(tsum);
  if (lane == 63) s_tmp[wid] = incl;
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint32_t t = s_tmp[w];
      s_tmp[w] = acc;
      acc += t;
    }
    s_tmp[W] = acc;
  }
  __syncthreads();
  uint32_t excl = incl - tsum + s_tmp[wid];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    int idx = tid * PER + k;
    if (idx < N) s[idx] = excl;
    excl += v[k];
  }
  uint32_t total = s_tmp[W];
  __syncthreads();
  return total;
}

}  // namespace jfx
