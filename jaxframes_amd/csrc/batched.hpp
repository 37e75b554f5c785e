// Batched groupby: many independent groupby instances of uneven row counts,
// concatenated into one table with frame offsets, run as ONE device pipeline.
//
// Every row gets the composite key (frame << key_bits) | (ordered(key) -
// ordered(min)); one stable sort of the composites (over exactly
// frame_bits + key_bits bits, no host read) puts the rows of each frame
// together in key order, and one head-flag segmented reduction over the
// whole table yields every frame's groups, frame-major. The composite is
// built here by k_batch_compose.
#pragma once

#include "common.hpp"

namespace jfx {

constexpr int BC_BLOCK = 256;
constexpr int BC_ROWS = 4096;        // rows per workgroup
constexpr int BC_LOCAL_FRAMES = 64;  // frame offsets staged in LDS per workgroup

// First index f in [0, m] with a[f] > x (a nondecreasing, m+1 entries).
__device__ __forceinline__ uint32_t bc_upper(const int64_t* a, uint32_t m, int64_t x) {
  uint32_t lo = 0, hi = m + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// offsets: nframes + 1 nondecreasing row offsets (offsets[0] = 0,
// offsets[nframes] = n). Keys outside [min, min + 2^key_bits) raise
// DEVERR_BOUNDS (the range was fixed when the pipeline was set up).
template <typename K, typename CK>
__global__ void __launch_bounds__(BC_BLOCK)
k_batch_compose(const K* __restrict__ keys, int oc, uint64_t omin, uint32_t key_bits,
                const int64_t* __restrict__ offsets, uint32_t nframes, uint64_t n,
                CK* __restrict__ out, uint32_t* err) {
  __shared__ int64_t s_off[BC_LOCAL_FRAMES + 1];
  __shared__ uint32_t s_f0, s_nf;
  const uint64_t r0 = (uint64_t)blockIdx.x * BC_ROWS;
  if (r0 >= n) return;
  const uint64_t r1 = (r0 + BC_ROWS) < n ? (r0 + BC_ROWS) : n;
  if (threadIdx.x == 0) {
    const uint32_t f0 = bc_upper(offsets, nframes, (int64_t)r0) - 1u;
    const uint32_t f1 = bc_upper(offsets, nframes, (int64_t)(r1 - 1)) - 1u;
    s_f0 = f0;
    s_nf = f1 - f0 + 1u;
  }
  __syncthreads();
  const uint32_t f0 = s_f0, nf = s_nf;
  const bool local = nf <= (uint32_t)BC_LOCAL_FRAMES;
  if (local)
    for (uint32_t i = threadIdx.x; i <= nf; i += BC_BLOCK) s_off[i] = offsets[f0 + i];
  __syncthreads();
  const uint64_t kmask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
  uint32_t bad = 0;
  uint32_t f = 0;  // frame index relative to f0; rows only increase per thread
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += BC_BLOCK) {
    uint32_t fr;
    if (local) {
      while (f + 1 < nf && s_off[f + 1] <= (int64_t)r) ++f;
      fr = f0 + f;
    } else {
      fr = bc_upper(offsets, nframes, (int64_t)r) - 1u;
    }
    const uint64_t d = (uint64_t)(ordered(keys[r], oc) - (K)omin);
    bad |= (d & ~kmask) != 0 ? 1u : 0u;
    out[r] = (CK)(((uint64_t)fr << key_bits) | (d & kmask));
  }
  if (ballot64(bad != 0u) != 0ull && lane_id() == 0) atomicOr(err, DEVERR_BOUNDS);
}

static __global__ void k_copy_u32(const uint32_t* src, uint32_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

static __global__ void k_zero_u32(uint32_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = 0u;
}

}  // namespace jfx
