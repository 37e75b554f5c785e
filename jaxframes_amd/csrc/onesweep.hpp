// One stable digit-scatter pass ("onesweep"): per-tile wave64 ballot ranking,
// decoupled-lookback digit offsets across tiles, LDS-staged coalesced writes.
// Shared by the LSD radix sort and the hash partitioner.
#pragma once

#include <type_traits>

#include "blockscan.hpp"
#include "common.hpp"

namespace jfx {

struct NoPay {};

template <typename K, typename PT>
struct ScatterArgs {
  const K* keys_in;
  K* keys_out;
  const PT* pay_in;      // ignored when iota != 0
  PT* pay_out;
  uint64_t n;
  const uint32_t* base;  // [radix] global exclusive start of each digit
  uint32_t* desc;        // [num_tiles][radix] lookback descriptors (zeroed)
  uint32_t* desc_next;   // descriptors of the next pass to zero, or null
  uint32_t* tile_counter;
  uint32_t* err;
  uint32_t radix_eff;    // digits are < radix_eff (<= 1 << D)
  int write_keys;
  int reverse;           // write to n-1-pos (descending final pass)
  int iota;              // payload = row index (uint32)
};

// Wave64 multisplit match: the mask of lanes in `valid` whose D-bit digit
// equals this lane's, using one ballot per digit bit. Per bit: a
// sign-extended bit (0 or ~0), one ballot, and an XNOR-AND on each 32-bit
// half of the mask (one v_bitop3 each on gfx950).
template <int D>
__device__ __forceinline__ uint64_t match_digit(uint32_t dg, uint64_t valid) {
  uint32_t lo = (uint32_t)valid, hi = (uint32_t)(valid >> 32);
#pragma unroll
  for (int b = 0; b < D; ++b) {
    const uint32_t x = (uint32_t)((int32_t)(dg << (31 - b)) >> 31);
    const uint64_t bb = __ballot(x != 0u);
    lo &= ~((uint32_t)bb ^ x);
    hi &= ~((uint32_t)(bb >> 32) ^ x);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Bits needed to represent digits below r (at least 1).
__host__ __device__ inline int digit_bits_for(uint32_t r) {
  int b = 1;
  while ((1u << b) < r) ++b;
  return b;
}

// Radix digit: ((ordered(key) - sub) >> shift) masked by the caller.
template <typename K>
struct RadixDigit {
  int oc;
  uint32_t shift;
  K sub;
  __device__ __forceinline__ uint32_t operator()(K k) const {
    return (uint32_t)((ordered(k, oc) - sub) >> shift);
  }
};

// Persistent-tile onesweep scatter. Tile t (taken from an atomic counter so
// every earlier tile is already running) holds TILE = W*64*KPT keys, loaded
// wave-striped: item i of lane l in wave w is key w*64*KPT + i*64 + l. That
// order is the input order, so ranking within the wave by (item, lane) and
// across waves by wave index is stable.
template <typename K, typename PT, int D, int W, int KPT, typename DigitF>
__global__ void __launch_bounds__(W * 64) k_onesweep(ScatterArgs<K, PT> a, DigitF dig) {
  constexpr int RADIX = 1 << D;
  constexpr int BLOCK = W * 64;
  constexpr int TILE = BLOCK * KPT;
  constexpr bool HAS_PAY = !std::is_same<PT, NoPay>::value;
  constexpr int XB = (HAS_PAY && sizeof(PT) > sizeof(K)) ? (int)sizeof(PT) : (int)sizeof(K);
  constexpr uint32_t MASK = RADIX - 1;

  // Per-wave digit counters; 16 bits suffice (a wave ranks at most 64*KPT
  // keys) and halve the LDS footprint of the 11-bit radix.
  using HT = typename std::conditional<(D > 8), uint16_t, uint32_t>::type;
  __shared__ HT s_whist[W * RADIX];
  __shared__ uint32_t s_tstart[RADIX];
  __shared__ uint32_t s_delta[RADIX];
  __shared__ uint32_t s_tmp[W + 1];
  __shared__ uint32_t s_tile;
  __shared__ __align__(16) unsigned char s_x[TILE * XB];
  K* s_keys = reinterpret_cast<K*>(s_x);

  const int tid = threadIdx.x;
  const int wid = tid >> 6;
  const uint32_t lane = lane_id();
  const uint32_t R = a.radix_eff;

  if (tid == 0) s_tile = atomicAdd(a.tile_counter, 1u);
  for (int i = tid; i < W * RADIX; i += BLOCK) s_whist[i] = HT(0);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tbase = (uint64_t)tile * TILE;
  const uint32_t tn = (uint32_t)((a.n - tbase) < (uint64_t)TILE ? (a.n - tbase) : (uint64_t)TILE);
  if (a.desc_next)
    for (uint32_t d = tid; d < R; d += BLOCK) a.desc_next[(size_t)tile * R + d] = 0u;

  K key[KPT];
  uint32_t dg[KPT];
  uint32_t slot[KPT];
  PT pay[KPT];
  const uint32_t wb = (uint32_t)wid * 64u * KPT;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const uint32_t li = wb + (uint32_t)i * 64u + lane;
    key[i] = li < tn ? a.keys_in[tbase + li] : K(0);
  }
  if constexpr (HAS_PAY) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      if (a.iota) pay[i] = (PT)(tbase + li);
      else pay[i] = li < tn ? a.pay_in[tbase + li] : PT(0);
    }
  }
#pragma unroll
  for (int i = 0; i < KPT; ++i) dg[i] = dig(key[i]) & MASK;

  // Wave-level multisplit ranking: peers = lanes with the same digit.
  HT* wh = s_whist + wid * RADIX;
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

  // Per digit: exclusive offsets across waves and the tile count.
  for (uint32_t d = tid; d < RADIX; d += BLOCK) {
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t c = s_whist[w * RADIX + d];
      s_whist[w * RADIX + d] = (HT)sum;
      sum += c;
    }
    s_tstart[d] = sum;
    s_delta[d] = sum;
    if (d < R) lookback_publish(a.desc, tile, R, d, sum);
  }
  __syncthreads();
  block_excl_scan_lds<RADIX, BLOCK>(s_tstart, s_tmp);

  // Keys to LDS in tile-sorted order.
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const uint32_t li = wb + (uint32_t)i * 64u + lane;
    if (li < tn) {
      slot[i] += s_tstart[dg[i]] + wh[dg[i]];
      s_keys[slot[i]] = key[i];
    }
  }
  // Global offsets of this tile's digit runs.
  for (uint32_t d = tid; d < R; d += BLOCK) {
    const uint32_t excl = lookback_u32(a.desc, tile, R, d, s_delta[d], a.err);
    s_delta[d] = a.base[d] + excl - s_tstart[d];
  }
  __syncthreads();

  uint32_t gpos[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const uint32_t j = (uint32_t)k * BLOCK + tid;
    if (j < tn) {
      const K kk = s_keys[j];
      uint32_t pos = s_delta[dig(kk) & MASK] + j;
      if ((uint64_t)pos >= a.n) {  // never taken when the offsets are consistent
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
}

}  // namespace jfx
