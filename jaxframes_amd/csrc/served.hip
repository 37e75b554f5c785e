// Indexed filter -> groupby for resident frames.
//
// A resident frame is indexed once: rows ordered by (key, value) -- groups
// contiguous, values ascending inside each group in the order-preserving
// unsigned image (NaN last) -- and the float64 prefix sums of the ordered
// values (NaN and infinities counted as 0, infinities counted per group). A query `groupby(key).agg(sum, count, mean) of
// the rows with value <cmp> t` then needs, per group, one or two binary
// searches for t inside the group's values and a difference of two prefix
// sums: the rows that pass form at most two runs of the group. The result of
// a group is the exact sum of its passing float32 values up to the float64
// prefix rounding (<= ~2^-52 * sum |x| over the group), rounded to the value
// type; groups where nothing passes are dropped by the caller (mask output).
// Semantics are those of filtering the rows and running
// SortBasedGroupBy.groupby_aggregate (parallel_algorithms.py:733-803) on them:
// IEEE comparisons (a NaN value passes only !=, and then makes its group's sum
// NaN), -0.0 == +0.0.
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "runtime.hpp"

namespace jfx {

template <typename V>
struct ServedArgs {
  const V* vals;              // values in (key, value) order
  const double* prefix;       // n + 1 prefix sums of vals (NaN and +-inf as 0)
  const uint32_t* starts;     // G + 1 group starts
  const uint32_t* nan_start;  // G: first NaN row of each group (its end if none)
  const uint32_t* ninf_end;   // G: end of the group's leading -inf rows
  const uint32_t* pinf_start; // G: first +inf row of the group (nan_start if none)
  uint32_t G;
  int cmp;
  V thr;
  int64_t* out_count;
  V* out_sum;
  V* out_mean;
  uint8_t* out_mask;          // 1 where the group has a passing row
};

// unsigned integer of the value's width
template <typename V>
struct SvU {
  using U = typename std::conditional<sizeof(V) == 4, uint32_t, uint64_t>::type;
};

template <typename V>
__device__ __forceinline__ typename SvU<V>::U ord_val(V v) {
  using U = typename SvU<V>::U;
  return ordered(__builtin_bit_cast(U, v), ORD_FLOAT);
}

// first i in [lo, hi) with image(vals[i]) > x (upper) / >= x (lower)
template <typename V>
__device__ __forceinline__ uint32_t sv_bound(const V* v, uint32_t lo, uint32_t hi,
                                             typename SvU<V>::U x, bool upper) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    const auto m = ord_val(v[mid]);
    if (upper ? (m <= x) : (m < x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename V>
__global__ void __launch_bounds__(256) k_served_groups(ServedArgs<V> a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G) return;
  const uint32_t lo = a.starts[g], hi = a.starts[g + 1], hn = a.nan_start[g];
  const bool tnan = a.thr != a.thr;
  const auto x = ord_val(a.thr);
  // up to two passing runs [r0, r1) and [r2, r3) among the non-NaN rows
  uint32_t r0 = lo, r1 = lo, r2 = hn, r3 = hn;
  bool nan_pass = false;
  if (!tnan) {
    switch (a.cmp) {
      case JFX_CMP_GT: r0 = sv_bound(a.vals, lo, hn, x, true); r1 = hn; break;
      case JFX_CMP_GE: r0 = sv_bound(a.vals, lo, hn, x, false); r1 = hn; break;
      case JFX_CMP_LT: r0 = lo; r1 = sv_bound(a.vals, lo, hn, x, false); break;
      case JFX_CMP_LE: r0 = lo; r1 = sv_bound(a.vals, lo, hn, x, true); break;
      case JFX_CMP_EQ:
        r0 = sv_bound(a.vals, lo, hn, x, false);
        r1 = sv_bound(a.vals, r0, hn, x, true);
        break;
      default:  // NE: everything but the equal run, NaN rows included
        r0 = lo;
        r1 = sv_bound(a.vals, lo, hn, x, false);
        r2 = sv_bound(a.vals, r1, hn, x, true);
        r3 = hn;
        nan_pass = true;
        break;
    }
  } else if (a.cmp == JFX_CMP_NE) {  // t is NaN: every row passes !=
    r0 = lo;
    r1 = hn;
    nan_pass = true;
  }
  const uint32_t nn = nan_pass ? hi - hn : 0u;
  const uint64_t cnt = (uint64_t)(r1 - r0) + (uint64_t)(r3 - r2) + nn;
  // finite rows through the prefix sums; -inf rows lead and +inf rows end the
  // group's non-NaN run, so the passing runs' overlaps with them count them
  double s = (a.prefix[r1] - a.prefix[r0]) + (a.prefix[r3] - a.prefix[r2]);
  const uint32_t ne = a.ninf_end[g], ps = a.pinf_start[g];
  auto overlap = [](uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1) {
    const uint32_t b = x0 > y0 ? x0 : y0, e = x1 < y1 ? x1 : y1;
    return e > b ? e - b : 0u;
  };
  const uint32_t nneg = overlap(r0, r1, lo, ne) + overlap(r2, r3, lo, ne);
  const uint32_t npos = overlap(r0, r1, ps, hn) + overlap(r2, r3, ps, hn);
  if (npos) s += __builtin_inf();
  if (nneg) s -= __builtin_inf();
  if (nn) s = __builtin_nan("");
  a.out_count[g] = (int64_t)cnt;
  if (a.out_sum) a.out_sum[g] = (V)s;
  if (a.out_mean) a.out_mean[g] = cnt ? (V)(s / (double)cnt) : (V)0;
  a.out_mask[g] = cnt ? 1 : 0;
}

}  // namespace jfx

extern "C" {

int jfx_served_groups(int val_type, const void* vals, const double* prefix, const uint32_t* starts,
                      const uint32_t* nan_start, const uint32_t* ninf_end,
                      const uint32_t* pinf_start, uint32_t ngroups, int cmp, const void* value,
                      int64_t* out_count, void* out_sum, void* out_mean, uint8_t* out_mask,
                      void* stream) {
  JFX_CHECK(val_type == JFX_F32 || val_type == JFX_F64, JFX_ERR_UNSUPPORTED,
            "indexed filters take float32 / float64 values");
  JFX_CHECK(cmp >= JFX_CMP_GT && cmp <= JFX_CMP_NE, JFX_ERR_INVALID, "bad comparison code");
  if (ngroups == 0) return JFX_OK;
  JFX_CHECK(vals && prefix && starts && nan_start && ninf_end && pinf_start && value && out_count &&
                out_mask,
            JFX_ERR_INVALID, "null argument");
  hipStream_t s = jfx::as_stream(stream);
  const uint32_t blocks = (ngroups + 255) / 256;
  if (val_type == JFX_F32) {
    jfx::ServedArgs<float> a{static_cast<const float*>(vals), prefix, starts, nan_start, ninf_end,
                             pinf_start, ngroups, cmp,
                             *static_cast<const float*>(value), out_count,
                             static_cast<float*>(out_sum), static_cast<float*>(out_mean), out_mask};
    JFX_LAUNCH("served_groups", s, (jfx::k_served_groups<float><<<blocks, 256, 0, s>>>(a)));
  } else {
    jfx::ServedArgs<double> a{static_cast<const double*>(vals), prefix, starts, nan_start, ninf_end,
                              pinf_start, ngroups, cmp,
                              *static_cast<const double*>(value), out_count,
                              static_cast<double*>(out_sum), static_cast<double*>(out_mean), out_mask};
    JFX_LAUNCH("served_groups", s, (jfx::k_served_groups<double><<<blocks, 256, 0, s>>>(a)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // extern "C"
