// Composite keys for multi-column sort / groupby / join.
//
// The reference orders rows by several key columns with one stable argsort
// per column, last key first (multi_column_lexsort, parallel_algorithms.py:
// 25-66), and finds groups where any key changes (:697-711). When every key
// column is an integer column and the value ranges of the columns fit in 64
// bits together, the row order and the groups are those of ONE composite key:
//
//   composite = sum_i field_i << shift_i,  field_i = ordered(k_i) - ordered(min_i)
//                                          (descending: ordered(max_i) - ordered(k_i))
//
// with the first key in the most significant field. The map is strictly
// monotonic per column, so a stable sort of the composite key is the stable
// lexicographic sort, and equal composites are equal key tuples. One radix
// sort (over the varying composite bits only) replaces K sorts, and the
// composite goes through the single-key groupby, including its small-range
// paths when the product of the ranges is small.
#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "runtime.hpp"

namespace jfx {

constexpr int PACK_MAXCOLS = 8;

struct PackCols {
  const void* col[PACK_MAXCOLS];
  void* out[PACK_MAXCOLS];
  uint64_t sub[PACK_MAXCOLS];   // ordered min (ascending) or ordered max (descending)
  uint64_t mask[PACK_MAXCOLS];  // (1 << bits) - 1
  uint32_t shift[PACK_MAXCOLS];
  int kb[PACK_MAXCOLS];         // 4 or 8
  int oc[PACK_MAXCOLS];         // ORD_UNSIGNED or ORD_SIGNED
  int desc[PACK_MAXCOLS];
  int ncols;
};

__device__ __forceinline__ uint64_t unordered_int(uint64_t o, int kb, int oc) {
  if (oc == ORD_SIGNED) o ^= kb == 4 ? 0x80000000ull : 0x8000000000000000ull;
  return o;
}

__global__ void __launch_bounds__(256) k_pack_keys(PackCols c, uint64_t n, uint64_t* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t v = 0;
    for (int j = 0; j < c.ncols; ++j) {
      uint64_t o;
      if (c.kb[j] == 4) o = ordered(static_cast<const uint32_t*>(c.col[j])[i], c.oc[j]);
      else o = ordered(static_cast<const uint64_t*>(c.col[j])[i], c.oc[j]);
      const uint64_t f = c.desc[j] ? c.sub[j] - o : o - c.sub[j];
      v |= (f & c.mask[j]) << c.shift[j];
    }
    out[i] = v;
  }
}

__global__ void __launch_bounds__(256) k_unpack_keys(PackCols c, const uint64_t* packed, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t v = packed[i];
    for (int j = 0; j < c.ncols; ++j) {
      const uint64_t f = (v >> c.shift[j]) & c.mask[j];
      const uint64_t o = c.desc[j] ? c.sub[j] - f : c.sub[j] + f;
      const uint64_t raw = unordered_int(o, c.kb[j], c.oc[j]);
      if (c.kb[j] == 4) static_cast<uint32_t*>(c.out[j])[i] = (uint32_t)raw;
      else static_cast<uint64_t*>(c.out[j])[i] = raw;
    }
  }
}

namespace {

int fill_cols(PackCols& c, const int* key_types, int ncols, const uint64_t* sub,
              const uint32_t* bits, const int* descending) {
  JFX_CHECK(ncols >= 1 && ncols <= PACK_MAXCOLS, JFX_ERR_INVALID, "1..8 key columns");
  JFX_CHECK(key_types && sub && bits, JFX_ERR_INVALID, "null column description");
  c.ncols = ncols;
  uint32_t shift = 0;
  // the last column is the least significant field
  for (int j = ncols - 1; j >= 0; --j) {
    const int t = key_types[j];
    JFX_CHECK(t == JFX_U32 || t == JFX_I32 || t == JFX_U64 || t == JFX_I64, JFX_ERR_UNSUPPORTED,
              "composite keys take integer columns only");
    JFX_CHECK(bits[j] <= 64, JFX_ERR_INVALID, "field wider than 64 bits");
    c.kb[j] = key_bytes(t);
    c.oc[j] = order_code(t);
    c.desc[j] = descending ? (descending[j] != 0) : 0;
    c.sub[j] = sub[j];
    c.mask[j] = bits[j] >= 64 ? ~0ull : ((1ull << bits[j]) - 1ull);
    c.shift[j] = shift;
    shift += bits[j];
    JFX_CHECK(shift <= 64, JFX_ERR_UNSUPPORTED, "key ranges need more than 64 bits");
  }
  return JFX_OK;
}

}  // namespace

}  // namespace jfx

extern "C" {

int jfx_pack_keys(const void* const* keys, const int* key_types, int ncols, const uint64_t* sub,
                  const uint32_t* bits, const int* descending, uint64_t n, uint64_t* out,
                  void* stream) {
  jfx::PackCols c{};
  int rc = jfx::fill_cols(c, key_types, ncols, sub, bits, descending);
  if (rc != JFX_OK) return rc;
  if (n == 0) return JFX_OK;
  JFX_CHECK(keys && out, JFX_ERR_INVALID, "null argument");
  for (int j = 0; j < ncols; ++j) {
    JFX_CHECK(keys[j] != nullptr, JFX_ERR_INVALID, "null key column");
    c.col[j] = keys[j];
  }
  hipStream_t s = jfx::as_stream(stream);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(16384, (n + 255) / 256);
  JFX_LAUNCH("pack_keys", s, (jfx::k_pack_keys<<<blocks, 256, 0, s>>>(c, n, out)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int jfx_unpack_keys(const uint64_t* packed, uint64_t n, const int* key_types, int ncols,
                    const uint64_t* sub, const uint32_t* bits, const int* descending,
                    void* const* outs, void* stream) {
  jfx::PackCols c{};
  int rc = jfx::fill_cols(c, key_types, ncols, sub, bits, descending);
  if (rc != JFX_OK) return rc;
  if (n == 0) return JFX_OK;
  JFX_CHECK(packed && outs, JFX_ERR_INVALID, "null argument");
  for (int j = 0; j < ncols; ++j) {
    JFX_CHECK(outs[j] != nullptr, JFX_ERR_INVALID, "null output column");
    c.out[j] = outs[j];
  }
  hipStream_t s = jfx::as_stream(stream);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(16384, (n + 255) / 256);
  JFX_LAUNCH("unpack_keys", s, (jfx::k_unpack_keys<<<blocks, 256, 0, s>>>(c, packed, n)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // extern "C"
