// Chunked single-pass digit scatter ("reduce-then-scan" without lookback).
//
// The input is cut into C contiguous chunks, one per workgroup, each a whole
// number of tiles. A histogram pass stores every chunk's digit counts; a
// column scan turns them into each chunk's starting offset per digit (digit
// major, chunk minor: the stable order); the scatter workgroup then walks its
// chunk tile by tile, ranking keys with wave64 ballots, staging the tile in
// LDS in digit order and appending each digit run at that digit's running
// offset. Consecutive tiles of one chunk append to adjacent addresses, so
// partial cache lines are completed in the same L2 before they reach HBM;
// no workgroup waits on another.
#pragma once

#include <type_traits>

#include "blockscan.hpp"
#include "common.hpp"
#include "onesweep.hpp"

namespace jfx {

struct ChunkGeom {
  uint64_t n;
  uint64_t chunk;    // keys per chunk (multiple of the tile)
  uint32_t nchunks;
};

// Where per-chunk counts come from when the column scan reads them.
enum CountMode : int {
  CM_DIRECT = 0,   // counts[b][d], stride R
  CM_ROTATE = 1,   // low-11-bit counts rotated by `sub`: counts[b][(d + sub) & 2047]
};

// Column scan, one wave per digit: lane l sums chunks [l*per, (l+1)*per),
// a wave exclusive scan orders the lane sums, and each lane writes the
// exclusive prefix of digit d over chunks < b to colex[b*R + d]; lane 63
// writes the digit total.
__device__ __forceinline__ uint32_t chunk_count(const uint32_t* counts, uint32_t cstride, int mode,
                                                uint32_t sub, uint32_t b, uint32_t d) {
  const uint32_t* row = counts + (size_t)b * cstride;
  if (mode == CM_DIRECT) return row[d];
  return row[(d + sub) & 2047u];
}

static __global__ void __launch_bounds__(256)
k_chunk_colscan(const uint32_t* counts, uint32_t cstride, int mode, uint32_t sub, uint32_t R,
                uint32_t nchunks, uint32_t* colex, uint32_t* total) {
  const uint32_t d = blockIdx.x * 4u + (threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  if (d >= R) return;
  const uint32_t per = (nchunks + 63u) / 64u;
  const uint32_t b0 = lane * per;
  const uint32_t b1 = (b0 + per) < nchunks ? (b0 + per) : nchunks;
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += chunk_count(counts, cstride, mode, sub, b, d);
  const uint32_t incl = wave_incl_scan(sum);
  uint32_t run = incl - sum;
  for (uint32_t b = b0; b < b1; ++b) {
    colex[(size_t)b * R + d] = run;
    run += chunk_count(counts, cstride, mode, sub, b, d);
  }
  if (lane == 63) total[d] = incl;
}

// Chunk scatter. Workgroup b owns keys [b*chunk, min((b+1)*chunk, n)).
// PF: software-pipeline the next tile's loads into registers. WPE: minimum
// waves per SIMD the register allocation must allow (4 = two 8-wave
// workgroups per CU, i.e. at most 128 VGPRs).
template <typename K, typename PT, int D, int W, int KPT, bool PF, int WPE, typename DigitF>
__global__ void __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_chunk_scatter(ScatterArgs<K, PT> a, ChunkGeom g, const uint32_t* colex, const uint32_t* total,
                DigitF dig) {
  constexpr int RADIX = 1 << D;
  constexpr int BLOCK = W * 64;
  constexpr int TILE = BLOCK * KPT;
  constexpr bool HAS_PAY = !std::is_same<PT, NoPay>::value;
  constexpr int XB = (HAS_PAY && sizeof(PT) > sizeof(K)) ? (int)sizeof(PT) : (int)sizeof(K);
  constexpr uint32_t MASK = RADIX - 1;
  using HT = typename std::conditional<(D > 8), uint16_t, uint32_t>::type;

  __shared__ HT s_whist[W * RADIX];
  __shared__ uint32_t s_run[RADIX];     // next output position of each digit
  __shared__ uint32_t s_tstart[RADIX];  // tile-local exclusive digit start
  __shared__ __align__(16) unsigned char s_x[TILE * XB];
  K* s_keys = reinterpret_cast<K*>(s_x);
  // scan scratch lives in the tile buffer: both scans run while it is free
  uint32_t* s_tmp = reinterpret_cast<uint32_t*>(s_x);

  const int tid = threadIdx.x;
  const int wid = tid >> 6;
  const uint32_t lane = lane_id();
  const uint32_t R = a.radix_eff;
  // per-digit loops only visit digits that can occur (rounded to whole waves)
  const uint32_t RL = ((R + 63u) & ~63u) < (uint32_t)RADIX ? ((R + 63u) & ~63u) : (uint32_t)RADIX;
  const uint32_t b = blockIdx.x;
  const uint64_t cbeg = (uint64_t)b * g.chunk;
  const uint64_t cend = (cbeg + g.chunk) < g.n ? (cbeg + g.chunk) : g.n;

  // running offsets: digit base (exclusive scan of totals) + this chunk's column prefix
  for (uint32_t d = tid; d < RADIX; d += BLOCK) s_run[d] = d < R ? total[d] : 0u;
  __syncthreads();
  block_excl_scan_lds<RADIX, BLOCK>(s_run, s_tmp);
  for (uint32_t d = tid; d < R; d += BLOCK) s_run[d] += colex[(size_t)b * R + d];
  __syncthreads();

  HT* wh = s_whist + wid * RADIX;
  const uint32_t wb = (uint32_t)wid * 64u * KPT;
  // Software pipeline (PF): the next tile's keys and payload are loaded into
  // registers while the current tile is ranked and written.
  K nkey[KPT];
  PT npay[KPT];
  auto fetch = [&](uint64_t tb, K* kd, PT* pd) {
    const uint32_t n2 = tb < cend ? (uint32_t)((cend - tb) < (uint64_t)TILE ? (cend - tb) : (uint64_t)TILE) : 0u;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      kd[i] = li < n2 ? a.keys_in[tb + li] : K(0);
    }
    if constexpr (HAS_PAY) {
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        const uint32_t li = wb + (uint32_t)i * 64u + lane;
        if (a.iota) pd[i] = (PT)(tb + li);
        else pd[i] = li < n2 ? a.pay_in[tb + li] : PT(0);
      }
    }
  };
  if constexpr (PF) fetch(cbeg, nkey, npay);
  for (uint64_t tbase = cbeg; tbase < cend; tbase += TILE) {
    const uint32_t tn = (uint32_t)((cend - tbase) < (uint64_t)TILE ? (cend - tbase) : (uint64_t)TILE);
    for (int w = 0; w < W; ++w)
      for (uint32_t d = tid; d < RL; d += BLOCK) s_whist[w * RADIX + d] = HT(0);
    K key[KPT];
    uint32_t dg[KPT];
    uint32_t slot[KPT];
    PT pay[KPT];
    if constexpr (PF) {
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        key[i] = nkey[i];
        if constexpr (HAS_PAY) pay[i] = npay[i];
      }
      fetch(tbase + TILE, nkey, npay);
    } else {
      fetch(tbase, key, pay);
    }
#pragma unroll
    for (int i = 0; i < KPT; ++i) dg[i] = dig(key[i]) & MASK;
    __syncthreads();

#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      const bool valid = li < tn;
      const uint64_t peers = match_digit<D>(dg[i], ballot64(valid));
      if (valid) {
        const uint32_t below = mbcnt(peers);
        const uint32_t cnt = wh[dg[i]];
        slot[i] = cnt + below;
        lds_order();
        if (below == 0u) wh[dg[i]] = (HT)(cnt + (uint32_t)__popcll(peers));
        lds_order();
      }
      // keep each item's ballots next to their use: hoisting the ballots of
      // all items ahead of the LDS chain exhausts the SGPRs
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    for (uint32_t d = tid; d < RADIX; d += BLOCK) {
      uint32_t sum = 0;
      if (d < RL) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const uint32_t c = s_whist[w * RADIX + d];
          s_whist[w * RADIX + d] = (HT)sum;
          sum += c;
        }
      }
      s_tstart[d] = sum;
    }
    __syncthreads();
    block_excl_scan_lds<RADIX, BLOCK>(s_tstart, s_tmp);
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      if (li < tn) {
        slot[i] += s_tstart[dg[i]] + wh[dg[i]];
        s_keys[slot[i]] = key[i];
      }
    }
    __syncthreads();
    uint32_t gpos[KPT];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const uint32_t j = (uint32_t)k * BLOCK + tid;
      if (j < tn) {
        const K kk = s_keys[j];
        const uint32_t d = dig(kk) & MASK;
        uint32_t pos = s_run[d] + (j - s_tstart[d]);
        if ((uint64_t)pos >= a.n) {
          atomicOr(a.err, DEVERR_BOUNDS);
          pos = 0xFFFFFFFFu;
        } else if (a.reverse) {
          pos = (uint32_t)(a.n - 1) - pos;
        }
        gpos[k] = pos;
        if (a.write_keys && pos != 0xFFFFFFFFu) a.keys_out[pos] = kk;
      }
    }
    if constexpr (HAS_PAY) {
      __syncthreads();
      PT* s_pay = reinterpret_cast<PT*>(s_x);
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        const uint32_t li = wb + (uint32_t)i * 64u + lane;
        if (li < tn) s_pay[slot[i]] = pay[i];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < KPT; ++k) {
        const uint32_t j = (uint32_t)k * BLOCK + tid;
        if (j < tn && gpos[k] != 0xFFFFFFFFu) a.pay_out[gpos[k]] = s_pay[j];
      }
    }
    __syncthreads();
    // advance the running offsets by this tile's digit counts
    for (uint32_t d = tid; d < R; d += BLOCK) {
      const uint32_t next = (d + 1 < (uint32_t)RADIX) ? s_tstart[d + 1] : tn;
      s_run[d] += next - s_tstart[d];
    }
    __syncthreads();
  }
}

}  // namespace jfx
