// Stable stream compaction: the positions, in row order, of the rows that
// pass a predicate.
//
// Serves compact_repartitioned_rows (parallel_algorithms.py:277-293: the
// nonzero positions of a validity mask) and boolean filters (the frame
// layer's df[mask] and the filter step of filter -> groupby queries).
// Predicates: a byte mask (nonzero keeps the row) or `column <cmp> value`
// evaluated on the fly, so a filter reads its column once per pass and never
// materialises a mask.
//
// Three kernels, no cross-workgroup waiting:
//   k_compact_count    per chunk: number of passing rows (wave ballots)
//   k_compact_scan     exclusive scan of the chunk counts (one workgroup)
//   k_compact_scatter  per chunk, tile by tile: ballot + mbcnt ranks inside
//                      each wave, an LDS scan over the waves, and the running
//                      chunk offset give every passing row its output slot
#include <algorithm>

#include "blockscan.hpp"
#include "common.hpp"
#include "runtime.hpp"

namespace jfx {

constexpr int CP_BLOCK = 256;
constexpr int CP_W = CP_BLOCK / 64;
constexpr int CP_IPT = 16;
constexpr uint32_t CP_TILE = CP_BLOCK * CP_IPT;  // rows per workgroup iteration
constexpr uint32_t CP_MAXCHUNKS = 1024;

struct MaskPred {
  const uint8_t* m;
  __device__ __forceinline__ bool operator()(uint64_t i) const { return m[i] != 0; }
};

template <typename T>
struct CmpPred {
  const T* col;
  int cmp;
  T thr;
  // IEEE comparisons: NaN passes only !=, as in numpy / jax
  __device__ __forceinline__ bool operator()(uint64_t i) const {
    const T v = col[i];
    switch (cmp) {
      case JFX_CMP_GT: return v > thr;
      case JFX_CMP_GE: return v >= thr;
      case JFX_CMP_LT: return v < thr;
      case JFX_CMP_LE: return v <= thr;
      case JFX_CMP_EQ: return v == thr;
      default: return v != thr;
    }
  }
};

// Row of item j of lane `lane` of wave `w` in a tile: each wave owns a
// contiguous 64*CP_IPT-row segment, walked 64 rows at a time.
__device__ __forceinline__ uint64_t cp_row(uint64_t tbase, int w, int j, uint32_t lane) {
  return tbase + (uint64_t)w * 64u * CP_IPT + (uint64_t)j * 64u + lane;
}

template <typename Pred>
__global__ void __launch_bounds__(CP_BLOCK) k_compact_count(Pred p, uint64_t n, uint64_t chunk,
                                                            uint32_t* counts) {
  __shared__ uint32_t s_w[CP_W];
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  uint32_t cnt = 0;
  for (uint64_t t = cbeg; t < cend; t += CP_TILE) {
#pragma unroll 4
    for (int j = 0; j < CP_IPT; ++j) {
      const uint64_t r = cp_row(t, wid, j, lane);
      cnt += (uint32_t)__popcll(ballot64(r < cend && p(r)));
    }
  }
  if (lane == 0) s_w[wid] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < CP_W; ++w) s += s_w[w];
    counts[blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(1024) k_compact_scan(uint32_t* counts, uint32_t nchunks,
                                                       uint64_t* total) {
  __shared__ uint32_t s[CP_MAXCHUNKS];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  for (uint32_t i = threadIdx.x; i < CP_MAXCHUNKS; i += 1024) s[i] = i < nchunks ? counts[i] : 0u;
  __syncthreads();
  const uint32_t tot = block_excl_scan_lds<CP_MAXCHUNKS, 1024>(s, s_tmp);
  for (uint32_t i = threadIdx.x; i < nchunks; i += 1024) counts[i] = s[i];
  if (threadIdx.x == 0) *total = tot;
}

template <typename Pred>
__global__ void __launch_bounds__(CP_BLOCK) k_compact_scatter(Pred p, uint64_t n, uint64_t chunk,
                                                              const uint32_t* offsets,
                                                              uint32_t* out_rows) {
  __shared__ uint32_t s_w[CP_W + 1];
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  uint32_t run = offsets[blockIdx.x];
  for (uint64_t t = cbeg; t < cend; t += CP_TILE) {
    uint64_t m[CP_IPT];
    uint32_t wcnt = 0;
#pragma unroll
    for (int j = 0; j < CP_IPT; ++j) {
      const uint64_t r = cp_row(t, wid, j, lane);
      m[j] = ballot64(r < cend && p(r));
      wcnt += (uint32_t)__popcll(m[j]);
    }
    if (lane == 0) s_w[wid] = wcnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int w = 0; w < CP_W; ++w) {
        const uint32_t c = s_w[w];
        s_w[w] = s;
        s += c;
      }
      s_w[CP_W] = s;
    }
    __syncthreads();
    uint32_t pos = run + s_w[wid];
#pragma unroll
    for (int j = 0; j < CP_IPT; ++j) {
      if ((m[j] >> lane) & 1ull) out_rows[pos + mbcnt(m[j])] = (uint32_t)cp_row(t, wid, j, lane);
      pos += (uint32_t)__popcll(m[j]);
    }
    run += s_w[CP_W];
    __syncthreads();
  }
}

namespace {

struct CpGeom {
  uint64_t chunk;
  uint32_t nchunks;
};

CpGeom cp_geometry(uint64_t n) {
  const uint64_t tiles = (n + CP_TILE - 1) / CP_TILE;
  const uint64_t nch = std::max<uint64_t>(1, std::min<uint64_t>(CP_MAXCHUNKS, tiles));
  const uint64_t tpc = (tiles + nch - 1) / nch;
  CpGeom g;
  g.chunk = tpc * CP_TILE;
  g.nchunks = (uint32_t)((n + g.chunk - 1) / g.chunk);
  return g;
}

struct CpWs {
  uint32_t* counts;
  uint64_t* total;
};

CpWs carve_cp(Carver& c) {
  CpWs w;
  w.counts = c.take<uint32_t>(CP_MAXCHUNKS);
  w.total = c.take<uint64_t>(1);
  return w;
}

template <typename Pred>
int compact_run(const Pred& p, uint64_t n, uint32_t* out_rows, uint64_t* out_count, void* ws,
                uint64_t ws_bytes, hipStream_t s) {
  Carver cv(ws, ws_bytes);
  CpWs w = carve_cp(cv);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "compaction workspace too small");
  const CpGeom g = cp_geometry(n);
  JFX_LAUNCH("compact_count", s, (k_compact_count<Pred><<<g.nchunks, CP_BLOCK, 0, s>>>(
                                      p, n, g.chunk, w.counts)));
  JFX_LAUNCH("compact_scan", s, (k_compact_scan<<<1, 1024, 0, s>>>(w.counts, g.nchunks, w.total)));
  JFX_LAUNCH("compact_scatter", s, (k_compact_scatter<Pred><<<g.nchunks, CP_BLOCK, 0, s>>>(
                                        p, n, g.chunk, w.counts, out_rows)));
  JFX_HIP(hipGetLastError());
  if (out_count) {
    return read_back(s, w.total, sizeof(uint64_t), out_count);
  }
  return JFX_OK;
}

template <typename T>
int filter_typed(const void* col, int cmp, const void* thr, uint64_t n, uint32_t* out_rows,
                 uint64_t* out_count, void* ws, uint64_t ws_bytes, hipStream_t s) {
  CmpPred<T> p{static_cast<const T*>(col), cmp, *static_cast<const T*>(thr)};
  return compact_run(p, n, out_rows, out_count, ws, ws_bytes, s);
}

}  // namespace

}  // namespace jfx

extern "C" {

int jfx_compact_workspace_size(uint64_t n, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  (void)n;
  jfx::Carver c(nullptr, 0);
  jfx::carve_cp(c);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_compact_rows(const uint8_t* mask, uint64_t n, uint32_t* out_rows, uint64_t* out_count,
                     void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(n < (1ull << 32), JFX_ERR_UNSUPPORTED, "n must be below 2^32 rows");
  if (n == 0) {
    if (out_count) *out_count = 0;
    return JFX_OK;
  }
  JFX_CHECK(mask && out_rows && workspace, JFX_ERR_INVALID, "null argument");
  jfx::MaskPred p{mask};
  return jfx::compact_run(p, n, out_rows, out_count, workspace, ws_bytes, jfx::as_stream(stream));
}

int jfx_filter_rows(int col_type, const void* col, int cmp, const void* value, uint64_t n,
                    uint32_t* out_rows, uint64_t* out_count, void* workspace, uint64_t ws_bytes,
                    void* stream) {
  JFX_CHECK(n < (1ull << 32), JFX_ERR_UNSUPPORTED, "n must be below 2^32 rows");
  JFX_CHECK(cmp >= JFX_CMP_GT && cmp <= JFX_CMP_NE, JFX_ERR_INVALID, "bad comparison code");
  JFX_CHECK(value != nullptr, JFX_ERR_INVALID, "value is null");
  if (n == 0) {
    if (out_count) *out_count = 0;
    return JFX_OK;
  }
  JFX_CHECK(col && out_rows && workspace, JFX_ERR_INVALID, "null argument");
  hipStream_t s = jfx::as_stream(stream);
  switch (col_type) {
    case JFX_U32: return jfx::filter_typed<uint32_t>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    case JFX_I32: return jfx::filter_typed<int32_t>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    case JFX_F32: return jfx::filter_typed<float>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    case JFX_U64: return jfx::filter_typed<uint64_t>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    case JFX_I64: return jfx::filter_typed<int64_t>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    case JFX_F64: return jfx::filter_typed<double>(col, cmp, value, n, out_rows, out_count, workspace, ws_bytes, s);
    default: break;
  }
  JFX_CHECK(false, JFX_ERR_INVALID, "bad column type");
  return JFX_ERR_INVALID;
}

}  // extern "C"
