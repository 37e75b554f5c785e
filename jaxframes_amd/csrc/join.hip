// Sort-merge join on sorted key columns.
//
// Restates ParallelSortMergeJoin.join (parallel_algorithms.py:938-1075). The
// reference sorts both sides (:968-972) and then takes all (i, j) with
// sorted_left[i] == sorted_right[j] from an L x R boolean matrix (:989-994),
// which enumerates pairs in (sorted-left position, sorted-right position)
// order. Here both sides arrive sorted (the radix sort with a row-index
// payload) and the matches of sorted-left row i are the range
// [lower_bound(key_i), upper_bound(key_i)) of the sorted right keys:
//   k_join_windows  one thread per left tile: the right-key window covering
//                   the tile's first..last key (two global binary searches)
//   k_join_count    per tile: each row's bounds by binary search inside the
//                   window staged in LDS; the tile's pair count
//   k_join_scan     exclusive scan of tile counts (M read back by the host)
//   k_join_fill     per tile: bounds again, an in-tile scan of counts, then
//                   a load-balanced expansion: output position k of the tile
//                   finds its left row by binary search over the scan, so a
//                   row with many matches is spread over the workgroup
//   k_join_first    left/right joins: first match or -1 (:1000-1034)
//   k_gather        payload gather by uint32 row, fill for row == ~0
// Float keys: NaN never matches (== in the match matrix is false); -0.0
// matches +0.0. Comparisons run on the order-preserving unsigned image.
#include <algorithm>

#include "blockscan.hpp"
#include "common.hpp"
#include "runtime.hpp"

namespace jfx {

constexpr int JT = 2048;        // left rows per tile
constexpr int JB = 256;         // threads per workgroup
constexpr int JIPT = JT / JB;   // left rows per thread
constexpr uint32_t JWMAX = 2048; // right keys staged in LDS per tile

template <typename K>
__device__ __forceinline__ bool is_nan_key(K k, int oc) {
  if (oc != ORD_FLOAT) return false;
  if constexpr (sizeof(K) == 4) return (k & 0x7FFFFFFFu) > 0x7F800000u;
  else return (k & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull;
}

// First index in [lo, hi) of sorted (ordered-image) array a with image >= x
// (upper = false) or > x (upper = true).
template <typename K>
__device__ __forceinline__ uint64_t gbound(const K* a, uint64_t lo, uint64_t hi, K x, int oc,
                                           bool upper) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const K v = ordered(a[mid], oc);
    if (upper ? (v <= x) : (v < x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K>
__device__ __forceinline__ uint32_t lbound(const K* s, uint32_t n, K x, bool upper) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const K v = s[mid];
    if (upper ? (v <= x) : (v < x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename K>
__global__ void __launch_bounds__(256)
k_join_windows(const K* sl, uint64_t L, const K* sr, uint64_t R, int oc, uint32_t ntiles,
               uint64_t* wlo, uint64_t* whi) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ntiles) return;
  const uint64_t a = (uint64_t)t * JT;
  const uint64_t b = (a + JT < L ? a + JT : L) - 1;
  wlo[t] = gbound(sr, 0, R, ordered(sl[a], oc), oc, false);
  whi[t] = gbound(sr, 0, R, ordered(sl[b], oc), oc, true);
}

// Bounds of the tile's rows: lo[j], cnt[j] for j in this thread's rows.
template <typename K>
__device__ __forceinline__ void tile_bounds(const K* sl, uint64_t L, const K* sr, int oc, uint64_t tb,
                                            uint64_t w0, uint64_t w1, const K* s_w, bool staged,
                                            uint64_t* lo, uint32_t* cnt) {
#pragma unroll
  for (int j = 0; j < JIPT; ++j) {
    const uint64_t i = tb + (uint64_t)threadIdx.x * JIPT + j;
    lo[j] = 0;
    cnt[j] = 0;
    if (i < L) {
      const K raw = sl[i];
      if (is_nan_key(raw, oc)) continue;
      const K x = ordered(raw, oc);
      uint64_t a, b;
      if (staged) {
        const uint32_t wn = (uint32_t)(w1 - w0);
        a = w0 + lbound(s_w, wn, x, false);
        b = w0 + lbound(s_w, wn, x, true);
      } else {
        a = gbound(sr, w0, w1, x, oc, false);
        b = gbound(sr, a, w1, x, oc, true);
      }
      lo[j] = a;
      cnt[j] = (uint32_t)(b - a);
    }
  }
}

template <typename K>
__device__ __forceinline__ bool stage_window(const K* sr, int oc, uint64_t w0, uint64_t w1, K* s_w) {
  if (w1 - w0 > JWMAX) return false;
  for (uint64_t r = w0 + threadIdx.x; r < w1; r += JB) s_w[r - w0] = ordered(sr[r], oc);
  return true;
}

template <typename K>
__global__ void __launch_bounds__(JB)
k_join_count(const K* sl, uint64_t L, const K* sr, int oc, const uint64_t* wlo, const uint64_t* whi,
             uint64_t* tile_cnt) {
  __shared__ K s_w[JWMAX];
  __shared__ uint64_t s_red[JB / 64];
  const uint32_t t = blockIdx.x;
  const uint64_t tb = (uint64_t)t * JT;
  const uint64_t w0 = wlo[t], w1 = whi[t];
  const bool staged = stage_window(sr, oc, w0, w1, s_w);
  __syncthreads();
  uint64_t lo[JIPT];
  uint32_t cnt[JIPT];
  tile_bounds(sl, L, sr, oc, tb, w0, w1, s_w, staged, lo, cnt);
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < JIPT; ++j) sum += cnt[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (lane_id() == 0) s_red[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t tot = 0;
    for (int w = 0; w < JB / 64; ++w) tot += s_red[w];
    tile_cnt[t] = tot;
  }
}

// Exclusive scan of n uint64 counts in place by one workgroup; writes the
// total to *total.
__global__ void __launch_bounds__(1024) k_join_scan(uint64_t* c, uint32_t n, uint64_t* total) {
  __shared__ uint64_t s_w[17];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t a = tid * per, b = (a + per) < n ? (a + per) : n;
  uint64_t sum = 0;
  for (uint32_t i = a; i < b; ++i) sum += c[i];
  const uint64_t incl = wave_incl_scan64(sum);
  if (lane_id() == 63) s_w[tid >> 6] = incl;
  __syncthreads();
  if (tid == 0) {
    uint64_t acc = 0;
    for (int w = 0; w < 16; ++w) {
      const uint64_t v = s_w[w];
      s_w[w] = acc;
      acc += v;
    }
    s_w[16] = acc;
  }
  __syncthreads();
  uint64_t run = s_w[tid >> 6] + incl - sum;
  for (uint32_t i = a; i < b; ++i) {
    const uint64_t v = c[i];
    c[i] = run;
    run += v;
  }
  if (tid == 0) *total = s_w[16];
}

// One payload column per side written by the fill itself (the value gather
// of parallel_algorithms.py:1052-1073 fused into pair enumeration): src is
// indexed by the side's source row (by_row) or by its sorted position.
template <typename LT, typename RT>
struct FillPay {
  const LT* lv;
  const RT* rv;
  LT* out_lv;
  RT* out_rv;
  int l_by_row;
  int r_by_row;
};

template <typename K, typename LT, typename RT>
__global__ void __launch_bounds__(JB)
k_join_fill(const K* sl, const uint32_t* lperm, uint64_t L, const K* sr, const uint32_t* rperm,
            int oc, const uint64_t* wlo, const uint64_t* whi, const uint64_t* tile_off,
            uint32_t* out_l, uint32_t* out_r, K* out_key, FillPay<LT, RT> pay) {
  __shared__ K s_w[JWMAX];
  __shared__ uint32_t s_incl[JT];   // inclusive scan of the tile's counts
  __shared__ uint32_t s_lo[JT];
  __shared__ uint32_t s_tmp[JB / 64 + 1];
  const uint32_t t = blockIdx.x;
  const uint64_t tb = (uint64_t)t * JT;
  const uint64_t w0 = wlo[t], w1 = whi[t];
  const bool staged = stage_window(sr, oc, w0, w1, s_w);
  __syncthreads();
  uint64_t lo[JIPT];
  uint32_t cnt[JIPT];
  tile_bounds(sl, L, sr, oc, tb, w0, w1, s_w, staged, lo, cnt);
#pragma unroll
  for (int j = 0; j < JIPT; ++j) {
    s_incl[threadIdx.x * JIPT + j] = cnt[j];
    s_lo[threadIdx.x * JIPT + j] = (uint32_t)lo[j];
  }
  __syncthreads();
  const uint32_t tot = block_excl_scan_lds<JT, JB>(s_incl, s_tmp);  // now exclusive
  const uint64_t base = tile_off[t];
  for (uint32_t k = threadIdx.x; k < tot; k += JB) {
    // row r: last row whose exclusive start <= k
    uint32_t a = 0, b = JT;
    while (a + 1 < b) {
      const uint32_t m = (a + b) >> 1;
      if (s_incl[m] <= k) a = m;
      else b = m;
    }
    const uint32_t r = a;
    const uint64_t i = tb + r;
    const uint64_t rr = (uint64_t)s_lo[r] + (k - s_incl[r]);
    // left rows: through lperm, or the sorted position when lperm is null
    const uint32_t lrow = lperm ? lperm[i] : (uint32_t)i;
    const uint32_t rrow = rperm ? rperm[rr] : (uint32_t)rr;
    if (out_l) out_l[base + k] = lrow;
    if (out_r) out_r[base + k] = rrow;
    if (out_key) out_key[base + k] = sl[i];
    if (pay.out_lv) pay.out_lv[base + k] = pay.lv[pay.l_by_row ? lrow : i];
    if (pay.out_rv) pay.out_rv[base + k] = pay.rv[pay.r_by_row ? rrow : rr];
  }
}

// First match of every sorted probe row in the sorted build side: build row
// (through perm) or ~0.
template <typename K>
__global__ void __launch_bounds__(256)
k_join_first(const K* sp, uint64_t P, const K* sb, const uint32_t* bperm, uint64_t B, int oc,
             uint32_t* out) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < P; i += stride) {
    const K raw = sp[i];
    uint32_t res = 0xFFFFFFFFu;
    if (!is_nan_key(raw, oc)) {
      const K x = ordered(raw, oc);
      const uint64_t a = gbound(sb, 0, B, x, oc, false);
      if (a < B && ordered(sb[a], oc) == x) res = bperm[a];
    }
    out[i] = res;
  }
}

// dst[i] = src[idx[i]] (elem bytes 4 or 8); idx == ~0 gives `fill`.
template <typename T>
__global__ void __launch_bounds__(256)
k_gather(const T* src, const uint32_t* idx, uint64_t n, T* dst, T fill) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const uint32_t j = idx[i];
    dst[i] = j == 0xFFFFFFFFu ? fill : src[j];
  }
}

namespace {

uint64_t jtiles(uint64_t L) { return (L + JT - 1) / JT; }

struct JoinWs {
  uint64_t* wlo;
  uint64_t* whi;
  uint64_t* tcnt;
  uint64_t* total;
};

JoinWs carve_join(Carver& c, uint64_t L) {
  JoinWs w{};
  const uint64_t t = jtiles(L);
  w.wlo = c.take<uint64_t>(t);
  w.whi = c.take<uint64_t>(t);
  w.tcnt = c.take<uint64_t>(t);
  w.total = c.take<uint64_t>(2);
  return w;
}

template <typename K>
int join_count_typed(int oc, const void* sl, uint64_t L, const void* sr, uint64_t R,
                     const JoinWs& w, uint64_t* out_m, hipStream_t s) {
  const uint32_t nt = (uint32_t)jtiles(L);
  JFX_LAUNCH("join_windows", s, (k_join_windows<K><<<(nt + 255) / 256, 256, 0, s>>>(
                                     (const K*)sl, L, (const K*)sr, R, oc, nt, w.wlo, w.whi)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("join_count", s, (k_join_count<K><<<nt, JB, 0, s>>>((const K*)sl, L, (const K*)sr, oc,
                                                                  w.wlo, w.whi, w.tcnt)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("join_scan", s, (k_join_scan<<<1, 1024, 0, s>>>(w.tcnt, nt, w.total)));
  JFX_HIP(hipGetLastError());
  return read_back(s, w.total, sizeof(uint64_t), out_m);
}

template <typename K, typename LT = uint32_t, typename RT = uint32_t>
int join_fill_typed(int oc, const void* sl, const uint32_t* lperm, uint64_t L, const void* sr,
                    const uint32_t* rperm, const JoinWs& w, uint32_t* out_l, uint32_t* out_r,
                    void* out_key, hipStream_t s, FillPay<LT, RT> pay = {}) {
  const uint32_t nt = (uint32_t)jtiles(L);
  JFX_LAUNCH("join_fill", s, (k_join_fill<K, LT, RT><<<nt, JB, 0, s>>>(
                                 (const K*)sl, lperm, L, (const K*)sr, rperm, oc, w.wlo, w.whi,
                                 w.tcnt, out_l, out_r, (K*)out_key, pay)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K, typename LT, typename RT>
int join_fill_pay3(int oc, const void* sl, const uint32_t* lperm, uint64_t L, const void* sr,
                   const uint32_t* rperm, const JoinWs& w, uint32_t* out_l, uint32_t* out_r,
                   void* out_key, const void* lv, int l_by_row, void* out_lv, const void* rv,
                   int r_by_row, void* out_rv, hipStream_t s) {
  FillPay<LT, RT> pay{(const LT*)lv, (const RT*)rv, (LT*)out_lv, (RT*)out_rv, l_by_row, r_by_row};
  return join_fill_typed<K, LT, RT>(oc, sl, lperm, L, sr, rperm, w, out_l, out_r, out_key, s, pay);
}

template <typename K, typename LT>
int join_fill_pay2(int rb, int oc, const void* sl, const uint32_t* lperm, uint64_t L,
                   const void* sr, const uint32_t* rperm, const JoinWs& w, uint32_t* out_l,
                   uint32_t* out_r, void* out_key, const void* lv, int l_by_row, void* out_lv,
                   const void* rv, int r_by_row, void* out_rv, hipStream_t s) {
  if (rb == 8)
    return join_fill_pay3<K, LT, uint64_t>(oc, sl, lperm, L, sr, rperm, w, out_l, out_r, out_key,
                                           lv, l_by_row, out_lv, rv, r_by_row, out_rv, s);
  return join_fill_pay3<K, LT, uint32_t>(oc, sl, lperm, L, sr, rperm, w, out_l, out_r, out_key, lv,
                                         l_by_row, out_lv, rv, r_by_row, out_rv, s);
}

template <typename K>
int join_fill_pay(int lb, int rb, int oc, const void* sl, const uint32_t* lperm, uint64_t L,
                  const void* sr, const uint32_t* rperm, const JoinWs& w, uint32_t* out_l,
                  uint32_t* out_r, void* out_key, const void* lv, int l_by_row, void* out_lv,
                  const void* rv, int r_by_row, void* out_rv, hipStream_t s) {
  if (lb == 8)
    return join_fill_pay2<K, uint64_t>(rb, oc, sl, lperm, L, sr, rperm, w, out_l, out_r, out_key,
                                       lv, l_by_row, out_lv, rv, r_by_row, out_rv, s);
  return join_fill_pay2<K, uint32_t>(rb, oc, sl, lperm, L, sr, rperm, w, out_l, out_r, out_key, lv,
                                     l_by_row, out_lv, rv, r_by_row, out_rv, s);
}

}  // namespace

}  // namespace jfx

extern "C" {

int jfx_join_workspace_size(uint64_t n_left, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  jfx::Carver c(nullptr, 0);
  jfx::carve_join(c, n_left);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_join_count(int key_type, const void* sorted_left, uint64_t n_left, const void* sorted_right,
                   uint64_t n_right, uint64_t* out_pairs, void* workspace, uint64_t ws_bytes,
                   void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(out_pairs != nullptr, JFX_ERR_INVALID, "out_pairs is null");
  JFX_CHECK(n_left < (1ull << 32) && n_right < (1ull << 32), JFX_ERR_UNSUPPORTED,
            "each side must have fewer than 2^32 rows");
  if (n_left == 0 || n_right == 0) {
    *out_pairs = 0;
    return JFX_OK;
  }
  JFX_CHECK(sorted_left && sorted_right && workspace, JFX_ERR_INVALID, "null input or workspace");
  jfx::Carver c(workspace, ws_bytes);
  jfx::JoinWs w = jfx::carve_join(c, n_left);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "join workspace too small");
  const int oc = jfx::order_code(key_type);
  hipStream_t s = jfx::as_stream(stream);
  if (jfx::key_bytes(key_type) == 4)
    return jfx::join_count_typed<uint32_t>(oc, sorted_left, n_left, sorted_right, n_right, w, out_pairs, s);
  return jfx::join_count_typed<uint64_t>(oc, sorted_left, n_left, sorted_right, n_right, w, out_pairs, s);
}

int jfx_join_fill(int key_type, const void* sorted_left, const uint32_t* left_rows, uint64_t n_left,
                  const void* sorted_right, const uint32_t* right_rows, uint64_t n_right,
                  uint32_t* out_left_rows, uint32_t* out_right_rows, void* out_keys,
                  void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  if (n_left == 0 || n_right == 0) return JFX_OK;
  JFX_CHECK(sorted_left && left_rows && sorted_right && right_rows && out_left_rows &&
                out_right_rows && workspace,
            JFX_ERR_INVALID, "null argument");
  jfx::Carver c(workspace, ws_bytes);
  jfx::JoinWs w = jfx::carve_join(c, n_left);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "join workspace too small");
  const int oc = jfx::order_code(key_type);
  hipStream_t s = jfx::as_stream(stream);
  (void)n_right;
  if (jfx::key_bytes(key_type) == 4)
    return jfx::join_fill_typed<uint32_t>(oc, sorted_left, left_rows, n_left, sorted_right,
                                          right_rows, w, out_left_rows, out_right_rows, out_keys, s);
  return jfx::join_fill_typed<uint64_t>(oc, sorted_left, left_rows, n_left, sorted_right,
                                        right_rows, w, out_left_rows, out_right_rows, out_keys, s);
}

int jfx_join_fill_values(int key_type, const void* sorted_left, const uint32_t* left_rows,
                         uint64_t n_left, const void* sorted_right, const uint32_t* right_rows,
                         uint64_t n_right, int left_elem_bytes, const void* left_values,
                         int left_by_row, int right_elem_bytes, const void* right_values,
                         int right_by_row, uint32_t* out_left_rows, uint32_t* out_right_rows,
                         void* out_keys, void* out_left_values, void* out_right_values,
                         void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK((left_elem_bytes == 4 || left_elem_bytes == 8) &&
                (right_elem_bytes == 4 || right_elem_bytes == 8),
            JFX_ERR_UNSUPPORTED, "payload elements must be 4 or 8 bytes");
  if (n_left == 0 || n_right == 0) return JFX_OK;
  JFX_CHECK(sorted_left && sorted_right && workspace, JFX_ERR_INVALID, "null argument");
  JFX_CHECK(!out_left_values || left_values, JFX_ERR_INVALID, "left payload output without input");
  JFX_CHECK(!out_right_values || right_values, JFX_ERR_INVALID,
            "right payload output without input");
  JFX_CHECK(!(out_left_rows || (out_left_values && left_by_row)) || left_rows, JFX_ERR_INVALID,
            "left rows needed for row outputs or row-indexed payloads");
  JFX_CHECK(!(out_right_rows || (out_right_values && right_by_row)) || right_rows, JFX_ERR_INVALID,
            "right rows needed for row outputs or row-indexed payloads");
  jfx::Carver c(workspace, ws_bytes);
  jfx::JoinWs w = jfx::carve_join(c, n_left);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "join workspace too small");
  const int oc = jfx::order_code(key_type);
  hipStream_t s = jfx::as_stream(stream);
  if (jfx::key_bytes(key_type) == 4)
    return jfx::join_fill_pay<uint32_t>(left_elem_bytes, right_elem_bytes, oc, sorted_left,
                                        left_rows, n_left, sorted_right, right_rows, w,
                                        out_left_rows, out_right_rows, out_keys, left_values,
                                        left_by_row, out_left_values, right_values, right_by_row,
                                        out_right_values, s);
  return jfx::join_fill_pay<uint64_t>(left_elem_bytes, right_elem_bytes, oc, sorted_left, left_rows,
                                      n_left, sorted_right, right_rows, w, out_left_rows,
                                      out_right_rows, out_keys, left_values, left_by_row,
                                      out_left_values, right_values, right_by_row,
                                      out_right_values, s);
}

int jfx_join_first_match(int key_type, const void* sorted_probe, uint64_t n_probe,
                         const void* sorted_build, const uint32_t* build_rows, uint64_t n_build,
                         uint32_t* out_rows, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  if (n_probe == 0) return JFX_OK;
  JFX_CHECK(sorted_probe && out_rows, JFX_ERR_INVALID, "null argument");
  JFX_CHECK(n_build == 0 || (sorted_build && build_rows), JFX_ERR_INVALID, "null build side");
  const int oc = jfx::order_code(key_type);
  hipStream_t s = jfx::as_stream(stream);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(8192, (n_probe + 255) / 256);
  if (jfx::key_bytes(key_type) == 4) {
    JFX_LAUNCH("join_first", s, (jfx::k_join_first<uint32_t><<<blocks, 256, 0, s>>>(
                                     (const uint32_t*)sorted_probe, n_probe,
                                     (const uint32_t*)sorted_build, build_rows, n_build, oc, out_rows)));
  } else {
    JFX_LAUNCH("join_first", s, (jfx::k_join_first<uint64_t><<<blocks, 256, 0, s>>>(
                                     (const uint64_t*)sorted_probe, n_probe,
                                     (const uint64_t*)sorted_build, build_rows, n_build, oc, out_rows)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int jfx_gather(int elem_bytes, const void* src, const uint32_t* rows, uint64_t n, void* dst,
               uint64_t fill_bits, void* stream) {
  JFX_CHECK(elem_bytes == 4 || elem_bytes == 8, JFX_ERR_UNSUPPORTED, "elem_bytes must be 4 or 8");
  if (n == 0) return JFX_OK;
  JFX_CHECK(src && rows && dst, JFX_ERR_INVALID, "null argument");
  hipStream_t s = jfx::as_stream(stream);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(16384, (n + 255) / 256);
  if (elem_bytes == 4) {
    JFX_LAUNCH("gather", s, (jfx::k_gather<uint32_t><<<blocks, 256, 0, s>>>(
                                 (const uint32_t*)src, rows, n, (uint32_t*)dst, (uint32_t)fill_bits)));
  } else {
    JFX_LAUNCH("gather", s, (jfx::k_gather<uint64_t><<<blocks, 256, 0, s>>>(
                                 (const uint64_t*)src, rows, n, (uint64_t*)dst, fill_bits)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // extern "C"
