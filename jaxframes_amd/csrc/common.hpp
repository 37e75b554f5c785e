// Shared device helpers for the sort-centric kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace jfx {

// Key type codes shared with include/jfx.h (JFX_KEY_*).
enum KeyType : int { KT_U32 = 0, KT_I32 = 1, KT_F32 = 2, KT_U64 = 3, KT_I64 = 4, KT_F64 = 5 };

// Ordered-key transform code: how raw key bits become unsigned bits whose
// unsigned order is the key order.
enum OrderCode : int { ORD_UNSIGNED = 0, ORD_SIGNED = 1, ORD_FLOAT = 2 };

__host__ __device__ inline int key_bytes(int kt) { return kt >= KT_U64 ? 8 : 4; }
__host__ __device__ inline int order_code(int kt) {
  return (kt == KT_U32 || kt == KT_U64) ? ORD_UNSIGNED
       : (kt == KT_I32 || kt == KT_I64) ? ORD_SIGNED : ORD_FLOAT;
}

// Order-preserving unsigned image of a key. Signed: flip the sign bit. Float:
// -0.0 -> +0.0 and every NaN -> +qNaN first (so -0.0 == +0.0 tie and NaN sorts
// last, the order of a stable float argsort), then invert negatives and set
// the sign bit of non-negatives.
__device__ __forceinline__ uint32_t ordered(uint32_t b, int oc) {
  if (oc == ORD_SIGNED) return b ^ 0x80000000u;
  if (oc == ORD_FLOAT) {
    uint32_t a = b & 0x7FFFFFFFu;
    b = (a == 0u) ? 0u : (a > 0x7F800000u ? 0x7FC00000u : b);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  return b;
}
__device__ __forceinline__ uint64_t ordered(uint64_t b, int oc) {
  if (oc == ORD_SIGNED) return b ^ 0x8000000000000000ull;
  if (oc == ORD_FLOAT) {
    uint64_t a = b & 0x7FFFFFFFFFFFFFFFull;
    b = (a == 0ull) ? 0ull : (a > 0x7FF0000000000000ull ? 0x7FF8000000000000ull : b);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  }
  return b;
}

// Equality as the reference's `!=` head test sees it: float NaN is never
// equal to anything, -0.0 equals +0.0.
__device__ __forceinline__ bool key_equal(uint32_t a, uint32_t b, int oc) {
  if (oc == ORD_FLOAT) {
    if ((a & 0x7FFFFFFFu) > 0x7F800000u || (b & 0x7FFFFFFFu) > 0x7F800000u) return false;
    if (((a | b) & 0x7FFFFFFFu) == 0u) return true;
  }
  return a == b;
}
__device__ __forceinline__ bool key_equal(uint64_t a, uint64_t b, int oc) {
  if (oc == ORD_FLOAT) {
    if ((a & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull ||
        (b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return false;
    if (((a | b) & 0x7FFFFFFFFFFFFFFFull) == 0ull) return true;
  }
  return a == b;
}

// ---- wave64 helpers -------------------------------------------------------
// The i1 builtin: the predicate goes straight into a scalar lane mask.
__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
// popcount(mask & lanes below me)
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// ---- cross-workgroup coherent word access ---------------------------------
// Relaxed atomics at agent scope are coherent across the XCDs' L2 caches.
// The lookback descriptors pack flag and value in one word, so no ordering
// with other memory is needed.
__device__ __forceinline__ uint32_t ld_acq_word(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_acq_word(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lookback descriptor (32-bit): 2 flag bits + 30-bit count.
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_PFX = 2u << 30;
constexpr uint32_t LB_VAL = (1u << 30) - 1u;
// Bound on polls of one descriptor before a workgroup gives up and raises the
// device error flag; a correct schedule never comes near it.
constexpr uint32_t LB_SPIN_LIMIT = 1u << 21;

// Device error codes written to the error word.
constexpr uint32_t DEVERR_LOOKBACK = 1u;  // a bounded wait ran out
constexpr uint32_t DEVERR_BOUNDS = 2u;    // a computed position fell outside its buffer

// Publish a tile's own count for one counter column (tile 0 publishes its
// count as an inclusive prefix straight away).
__device__ __forceinline__ void lookback_publish(uint32_t* desc, uint32_t tile, uint32_t stride,
                                                 uint32_t col, uint32_t count) {
  st_word(desc + (size_t)tile * stride + col, (tile == 0 ? LB_PFX : LB_AGG) | count);
}

// Decoupled lookback for one counter column: after lookback_publish, walks
// back over earlier tiles, returns the exclusive prefix over tiles < tile and
// publishes this tile's inclusive prefix. desc is laid out [tile][stride]
// with this column at offset `col` (tiles * stride < 2^32). Descriptors are
// fetched four tiles at a time and consumed in order. The four loads must
// be independent so that a round costs one round trip: their addresses are
// 32-bit offsets clamped to tile 0 (always valid) and the loads are
// unconditional, the value of a tile below 0 being replaced afterwards. (With
// 64-bit index products under per-load branches the compiler gave the
// products a high half that aliased the previous load's destination
// register and waited for that load before issuing the next: three round
// trips per round.)
__device__ __forceinline__ uint32_t lookback_u32(uint32_t* desc, uint32_t tile, uint32_t stride,
                                                 uint32_t col, uint32_t count, uint32_t* err) {
  if (tile == 0) return 0u;
  uint32_t excl = 0;
  int32_t t = (int32_t)tile - 1;
  uint32_t spins = 0;
  bool done = false;
  while (!done) {
    uint32_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t tq = t - q;
      v[q] = ld_acq_word(desc + ((uint32_t)(tq > 0 ? tq : 0) * stride + col));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (t - q < 0) v[q] = LB_PFX;
    int used = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (done || used < q) continue;
      const uint32_t f = v[q] & ~LB_VAL;
      if (f == 0u) continue;  // not published yet: refetch from here
      excl += v[q] & LB_VAL;
      used = q + 1;
      if (f == LB_PFX) done = true;
    }
    if (done) break;
    t -= used;
    if (used == 0) {
      if (++spins > LB_SPIN_LIMIT) {
        atomicOr(err, DEVERR_LOOKBACK);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  st_word(desc + (size_t)tile * stride + col, LB_PFX | (excl + count));
  return excl;
}

// Wave-parallel lookback for a single counter (desc[tile]): each round, lane
// l reads tile-1-l; the nearest inclusive prefix ends the walk. Must be called
// by all 64 lanes of one wave; every lane returns the exclusive prefix and
// lane 0 publishes this tile's inclusive prefix.
__device__ __forceinline__ uint32_t wave_lookback_u32(uint32_t* desc, uint32_t tile, uint32_t count,
                                                      uint32_t* err) {
  if (tile == 0) return 0u;
  const uint32_t lane = __lane_id();
  uint32_t excl = 0;
  int64_t base = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t t = base - (int64_t)lane;
    const uint32_t v = t >= 0 ? ld_acq_word(desc + t) : LB_PFX;
    const uint32_t f = v & ~LB_VAL;
    const uint64_t pfx = __ballot(f == LB_PFX);
    const uint64_t zero = __ballot(f == 0u);
    const int first = pfx ? __builtin_ctzll(pfx) : 64;
    const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
    if (zero & need) {
      if (++spins > LB_SPIN_LIMIT) {
        if (lane == 0) atomicOr(err, DEVERR_LOOKBACK);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t x = ((int)lane <= first) ? (v & LB_VAL) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    excl += x;
    if (first < 64) break;
    base -= 64;
  }
  if (lane == 0) st_word(desc + tile, LB_PFX | (excl + count));
  return excl;
}

template <typename T> __device__ __forceinline__ T ldg(const T* p) { return __builtin_nontemporal_load(p); }

}  // namespace jfx
