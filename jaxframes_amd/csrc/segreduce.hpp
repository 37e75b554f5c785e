// Segmented-reduction state for the groupby aggregations.
//
// One state per (partial) group: compensated sum (s, c), row count, min, max.
// Floating sums use TwoSum (Knuth) compensation: every add and every combine
// keeps the rounding error in c, so s + c carries about twice the working
// precision. Integer sums wrap in the value type, as XLA's int segment_sum.
#pragma once

#include <math.h>

#include "common.hpp"

namespace jfx {

template <typename V> struct ValTraits;
template <> struct ValTraits<float> {
  using U = uint32_t;
  static constexpr bool is_float = true;
  __device__ static float lowest() { return -INFINITY; }
  __device__ static float highest() { return INFINITY; }
};
template <> struct ValTraits<double> {
  using U = uint64_t;
  static constexpr bool is_float = true;
  __device__ static double lowest() { return -INFINITY; }
  __device__ static double highest() { return INFINITY; }
};
template <> struct ValTraits<int32_t> {
  using U = uint32_t;
  static constexpr bool is_float = false;
  __device__ static int32_t lowest() { return INT32_MIN; }
  __device__ static int32_t highest() { return INT32_MAX; }
};
template <> struct ValTraits<int64_t> {
  using U = uint64_t;
  static constexpr bool is_float = false;
  __device__ static int64_t lowest() { return INT64_MIN; }
  __device__ static int64_t highest() { return INT64_MAX; }
};
template <> struct ValTraits<uint32_t> {
  using U = uint32_t;
  static constexpr bool is_float = false;
  __device__ static uint32_t lowest() { return 0u; }
  __device__ static uint32_t highest() { return 0xFFFFFFFFu; }
};
template <> struct ValTraits<uint64_t> {
  using U = uint64_t;
  static constexpr bool is_float = false;
  __device__ static uint64_t lowest() { return 0ull; }
  __device__ static uint64_t highest() { return ~0ull; }
};

// Exact rounding error of a + b (TwoSum, no branches, no FMA involved). When
// the sum is not finite the error is 0, so that a sum that reached +-inf
// stays +-inf (the formula would give inf - inf = NaN).
template <typename V>
__device__ __forceinline__ V two_sum(V a, V b, V& err) {
  const V s = a + b;
  const V bp = s - a;
  const V ap = s - bp;
  const V e = (a - ap) + (b - bp);
  err = __builtin_isfinite(s) ? e : V(0);
  return s;
}

// NaN-propagating min / max (IEEE 754-2019 minimum/maximum: one
// v_minimum3/v_maximum3 on gfx950 for float32). A NaN anywhere in a group
// makes its min and max NaN, as numpy's minimum/maximum reductions do.
template <typename V> __device__ __forceinline__ V vmin(V a, V b) {
  if constexpr (ValTraits<V>::is_float) return __builtin_elementwise_minimum(a, b);
  else return b < a ? b : a;
}
template <typename V> __device__ __forceinline__ V vmax(V a, V b) {
  if constexpr (ValTraits<V>::is_float) return __builtin_elementwise_maximum(a, b);
  else return b > a ? b : a;
}

// MM = false drops the min / max fields from every update and exchange when
// no min or max output is requested.
template <typename V, bool MM = true>
struct AggState {
  V s;
  V c;
  uint32_t n;
  V mn;
  V mx;

  __device__ __forceinline__ static AggState identity() {
    AggState r;
    r.s = V(0);
    r.c = V(0);
    r.n = 0u;
    r.mn = ValTraits<V>::highest();
    r.mx = ValTraits<V>::lowest();
    return r;
  }
  __device__ __forceinline__ void add(V x) {
    if constexpr (ValTraits<V>::is_float) {
      V e;
      s = two_sum(s, x, e);
      c += e;
    } else {
      using U = typename ValTraits<V>::U;
      s = (V)((U)s + (U)x);
    }
    n += 1u;
    if constexpr (MM) {
      mn = vmin(mn, x);
      mx = vmax(mx, x);
    }
  }
  // this := this (earlier rows) then b (later rows)
  __device__ __forceinline__ void merge(const AggState& b) {
    if constexpr (ValTraits<V>::is_float) {
      V e;
      s = two_sum(s, b.s, e);
      c = c + b.c + e;
    } else {
      using U = typename ValTraits<V>::U;
      s = (V)((U)s + (U)b.s);
    }
    n += b.n;
    if constexpr (MM) {
      mn = vmin(mn, b.mn);
      mx = vmax(mx, b.mx);
    }
  }
  __device__ __forceinline__ V sum() const {
    if constexpr (ValTraits<V>::is_float) return s + c;
    else return s;
  }
};

template <typename V, bool MM>
__device__ __forceinline__ AggState<V, MM> shfl_up_state(const AggState<V, MM>& x, int off) {
  AggState<V, MM> r;
  r.s = __shfl_up(x.s, off, 64);
  if constexpr (ValTraits<V>::is_float) r.c = __shfl_up(x.c, off, 64);
  else r.c = V(0);
  r.n = __shfl_up(x.n, off, 64);
  if constexpr (MM) {
    r.mn = __shfl_up(x.mn, off, 64);
    r.mx = __shfl_up(x.mx, off, 64);
  } else {
    r.mn = x.mn;
    r.mx = x.mx;
  }
  return r;
}

}  // namespace jfx
