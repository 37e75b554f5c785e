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
// equals this lane's, using one ballot per digit bit. Per bit: the bit
// sign-extended (v_bfe_i32: 0 or ~0), one ballot of it, and on each 32-bit
// half of the mask m := m & ~(ballot ^ bit) as one v_bitop3 (truth table
// 0x90 over (m, ballot, bit)): four VALU instructions per bit. The asm
// barrier keeps the ballot on the extracted bit; otherwise the compiler
// re-derives it with a shift and a signed compare (five per bit).
template <int D>
__device__ __forceinline__ uint64_t match_digit(uint32_t dg, uint64_t valid) {
  uint32_t lo = (uint32_t)valid, hi = (uint32_t)(valid >> 32);
#pragma unroll
  for (int b = 0; b < D; ++b) {
    uint32_t x = (uint32_t)__builtin_amdgcn_sbfe((int)dg, b, 1);
    asm volatile("" : "+v"(x));
    const uint64_t bb = ballot64(x != 0u);
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bb, x, 0x90);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bb >> 32), x, 0x90);
  }
  return ((uint64_t)hi << 32) | lo;
}

// The same for two digits at once, their instructions interleaved (no
// dependent pair of VALU instructions back to back).
template <int D>
__device__ __forceinline__ void match_digit2(uint32_t d0, uint32_t d1, uint64_t v0, uint64_t v1,
                                             uint64_t& m0, uint64_t& m1) {
  uint32_t lo0 = (uint32_t)v0, hi0 = (uint32_t)(v0 >> 32);
  uint32_t lo1 = (uint32_t)v1, hi1 = (uint32_t)(v1 >> 32);
#pragma unroll
  for (int b = 0; b < D; ++b) {
    uint32_t x0 = (uint32_t)__builtin_amdgcn_sbfe((int)d0, b, 1);
    uint32_t x1 = (uint32_t)__builtin_amdgcn_sbfe((int)d1, b, 1);
    asm volatile("" : "+v"(x0), "+v"(x1));
    const uint64_t b0 = ballot64(x0 != 0u);
    const uint64_t b1 = ballot64(x1 != 0u);
    lo0 = __builtin_amdgcn_bitop3_b32(lo0, (uint32_t)b0, x0, 0x90);
    hi0 = __builtin_amdgcn_bitop3_b32(hi0, (uint32_t)(b0 >> 32), x0, 0x90);
    lo1 = __builtin_amdgcn_bitop3_b32(lo1, (uint32_t)b1, x1, 0x90);
    hi1 = __builtin_amdgcn_bitop3_b32(hi1, (uint32_t)(b1 >> 32), x1, 0x90);
  }
  m0 = ((uint64_t)hi0 << 32) | lo0;
  m1 = ((uint64_t)hi1 << 32) | lo1;
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
  // keys) and halve the LDS footprint of the 11-bit radix. Each wave's row
  // has one spare entry (index RADIX) that lanes past the tile's end count
  // into when keys are ranked in pairs.
  using HT = typename std::conditional<(D > 8), uint16_t, uint32_t>::type;
  constexpr int WS = RADIX + 1;
  // Keys ranked in pairs: workgroups of 16 waves whose LDS holds a CU alone
  // (a second workgroup would not fit in 160 KB), with payloads of at most 4
  // bytes (the 8-byte ones leave no VGPRs for it at 4 waves per SIMD).
  constexpr size_t LDS_BYTES = sizeof(HT) * W * WS + 8u * RADIX + 4u * (W + 2) + (size_t)TILE * XB;
  constexpr bool RANK2 = W == 16 && 2 * LDS_BYTES > 160u * 1024u && sizeof(PT) <= 4;
  static_assert(!RANK2 || KPT % 2 == 0, "keys per lane are ranked in pairs");
  __shared__ HT s_whist[W * WS];
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
  for (int i = tid; i < W * WS; i += BLOCK) s_whist[i] = HT(0);
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

  // Wave-level multisplit ranking: peers = lanes with the same digit.
  HT* wh = s_whist + wid * WS;
  if constexpr (RANK2) {
    // Two keys per lane at a time. Each pair's digits are taken inside the
    // loop, so the first pairs are ranked while the later keys are still
    // loading (the loads complete in order). Every lane of a digit group
    // writes the same new count; lanes past the tile's end count into the
    // spare entry (no branch). At 100M uint64 keys + uint32 rows, 16384-key
    // tiles: 544-574 us per pass against 584-630 us for one key at a time.
    // It takes 14-30 more VGPRs, which would halve the occupancy of the
    // smaller tiles (two workgroups per CU), so only tiles that hold a CU
    // alone by their LDS use it.
#pragma unroll
    for (int i = 0; i < KPT; i += 2) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      const bool v0 = li < tn, v1 = li + 64u < tn;
      dg[i] = dig(key[i]) & MASK;
      dg[i + 1] = dig(key[i + 1]) & MASK;
      uint64_t p0, p1;
      match_digit2<D>(dg[i], dg[i + 1], ballot64(v0), ballot64(v1), p0, p1);
      const uint32_t e0 = v0 ? dg[i] : (uint32_t)RADIX;
      const uint32_t e1 = v1 ? dg[i + 1] : (uint32_t)RADIX;
      const uint32_t c0 = wh[e0];
      slot[i] = c0 + mbcnt(p0);
      lds_order();
      wh[e0] = (HT)(c0 + (uint32_t)__popcll(p0));
      lds_order();
      const uint32_t c1 = wh[e1];
      slot[i + 1] = c1 + mbcnt(p1);
      lds_order();
      wh[e1] = (HT)(c1 + (uint32_t)__popcll(p1));
      lds_order();
      // keep each pair's ballots next to their use: hoisting the ballots of
      // all keys ahead of the LDS chain exhausts the SGPRs
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < KPT; ++i) dg[i] = dig(key[i]) & MASK;
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
      // keep each key's ballots next to their use
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();

  // Per digit: exclusive offsets across waves and the tile count.
  for (uint32_t d = tid; d < RADIX; d += BLOCK) {
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t c = s_whist[w * WS + d];
      s_whist[w * WS + d] = (HT)sum;
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
