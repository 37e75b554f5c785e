// Indexed filter -> groupby for resident frames.
//
// A resident frame is indexed once: rows ordered by (key, value) -- groups
// contiguous, values ascending inside each group in the order-preserving
// unsigned image (NaN last) -- and the float64 prefix sums of the ordered
// values restarted at every group (NaN and infinities counted as 0,
// infinities counted per group). A query `groupby(key).agg(sum, count, mean) of
// the rows with value <cmp> t` then needs, per group, one or two binary
// searches for t inside the group's values and a difference of two prefix
// sums: the rows that pass form at most two runs of the group. The result of
// a group is the exact sum of its passing float32 values up to the float64
// rounding of the group's own prefix sums (about log2(rows) * 2^-53 *
// sum |x| over the group, independent of the other groups), rounded to the
// value type; groups where nothing passes are dropped by the caller (mask output).
// Semantics are those of filtering the rows and running
// SortBasedGroupBy.groupby_aggregate (parallel_algorithms.py:733-803) on them:
// IEEE comparisons (a NaN value passes only !=, and then makes its group's sum
// NaN), -0.0 == +0.0.
#include <algorithm>
#include <type_traits>

#include "blockscan.hpp"
#include "common.hpp"
#include "runtime.hpp"
#include "sortctx.hpp"

namespace jfx {

template <typename V>
struct ServedArgs {
  const V* vals;              // values in (key, value) order
  const double* prefix;       // n + 1 per-group prefix sums of vals (NaN and +-inf as 0)
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

// Evaluation of one group: passing-row count and the float64 sum of the
// passing values (NaN / +-inf rules as stated above).
struct SvEval {
  uint64_t cnt;
  double s;
};

template <typename V>
__device__ __forceinline__ SvEval served_eval(const V* vals, const double* prefix,
                                              const uint32_t* starts, const uint32_t* nan_start,
                                              const uint32_t* ninf_end, const uint32_t* pinf_start,
                                              uint32_t g, int cmp, V thr) {
  const uint32_t lo = starts[g], hi = starts[g + 1], hn = nan_start[g];
  const bool tnan = thr != thr;
  const auto x = ord_val(thr);
  // up to two passing runs [r0, r1) and [r2, r3) among the non-NaN rows
  uint32_t r0 = lo, r1 = lo, r2 = hn, r3 = hn;
  bool nan_pass = false;
  if (!tnan) {
    switch (cmp) {
      case JFX_CMP_GT: r0 = sv_bound(vals, lo, hn, x, true); r1 = hn; break;
      case JFX_CMP_GE: r0 = sv_bound(vals, lo, hn, x, false); r1 = hn; break;
      case JFX_CMP_LT: r0 = lo; r1 = sv_bound(vals, lo, hn, x, false); break;
      case JFX_CMP_LE: r0 = lo; r1 = sv_bound(vals, lo, hn, x, true); break;
      case JFX_CMP_EQ:
        r0 = sv_bound(vals, lo, hn, x, false);
        r1 = sv_bound(vals, r0, hn, x, true);
        break;
      default:  // NE: everything but the equal run, NaN rows included
        r0 = lo;
        r1 = sv_bound(vals, lo, hn, x, false);
        r2 = sv_bound(vals, r1, hn, x, true);
        r3 = hn;
        nan_pass = true;
        break;
    }
  } else if (cmp == JFX_CMP_NE) {  // t is NaN: every row passes !=
    r0 = lo;
    r1 = hn;
    nan_pass = true;
  }
  const uint32_t nn = nan_pass ? hi - hn : 0u;
  SvEval e;
  e.cnt = (uint64_t)(r1 - r0) + (uint64_t)(r3 - r2) + nn;
  // finite rows through the group's prefix sums (prefix[i] = sum of the
  // group's rows before row i for i > lo; 0 at the group's first row);
  // -inf rows lead and +inf rows end the group's non-NaN run, so the passing
  // runs' overlaps with them count them
  auto pre = [&](uint32_t i) -> double { return i > lo ? prefix[i] : 0.0; };
  double s = (pre(r1) - pre(r0)) + (pre(r3) - pre(r2));
  const uint32_t ne = ninf_end[g], ps = pinf_start[g];
  auto overlap = [](uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1) {
    const uint32_t b = x0 > y0 ? x0 : y0, en = x1 < y1 ? x1 : y1;
    return en > b ? en - b : 0u;
  };
  const uint32_t nneg = overlap(r0, r1, lo, ne) + overlap(r2, r3, lo, ne);
  const uint32_t npos = overlap(r0, r1, ps, hn) + overlap(r2, r3, ps, hn);
  if (npos) s += __builtin_inf();
  if (nneg) s -= __builtin_inf();
  if (nn) s = __builtin_nan("");
  e.s = s;
  return e;
}

template <typename V>
__global__ void __launch_bounds__(256) k_served_groups(ServedArgs<V> a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G) return;
  const SvEval e = served_eval(a.vals, a.prefix, a.starts, a.nan_start, a.ninf_end, a.pinf_start,
                               g, a.cmp, a.thr);
  a.out_count[g] = (int64_t)e.cnt;
  if (a.out_sum) a.out_sum[g] = (V)e.s;
  if (a.out_mean) a.out_mean[g] = e.cnt ? (V)(e.s / (double)e.cnt) : (V)0;
  a.out_mask[g] = e.cnt ? 1 : 0;
}

// ---- a batch of queries in one launch --------------------------------------
// The groups of all queries form one index space (query-major); a workgroup
// takes the next tile of SVB groups by ticket, evaluates them, and writes the
// groups that keep rows compacted in that order, its output position coming
// from a decoupled lookback over the tiles' pass counts. The first group of
// each query records where the query's output starts.
constexpr int SVB = 256;
constexpr uint32_t SV_QMAX = 1024;

struct SvDev {
  jfx_served_query q;
  uint64_t goff;  // first index of the query's groups in the batch
};

template <typename V, typename KT>
__global__ void __launch_bounds__(SVB)
k_served_batch(const SvDev* qs, uint32_t Q, uint64_t tot, uint32_t* desc, uint32_t* counter,
               uint32_t* err, KT* out_keys, int64_t* out_count, V* out_sum, V* out_mean,
               uint64_t* out_start) {
  __shared__ uint64_t s_off[SV_QMAX];
  __shared__ uint32_t s_wcnt[SVB / 64];
  __shared__ uint32_t s_tile, s_excl;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  for (uint32_t i = tid; i < Q; i += SVB) s_off[i] = qs[i].goff;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t g = (uint64_t)tile * SVB + tid;
  bool pass = false;
  SvEval e{0, 0.0};
  uint32_t q = 0;
  uint32_t gl = 0;
  if (g < tot) {
    uint32_t lo = 0, hi = Q;  // last query whose first group is <= g
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= g) lo = mid;
      else hi = mid;
    }
    q = lo;
    gl = (uint32_t)(g - s_off[q]);
    const jfx_served_query& qq = qs[q].q;
    e = served_eval(static_cast<const V*>(qq.values), qq.prefix, qq.group_starts, qq.nan_start,
                    qq.ninf_end, qq.pinf_start, gl, qq.cmp, (V)qq.threshold);
    pass = e.cnt != 0;
  }
  const uint64_t m = ballot64(pass);
  const uint32_t below = mbcnt(m);
  const uint32_t wave = tid >> 6;
  if ((tid & 63) == 0) s_wcnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t woff = 0, T = 0;
#pragma unroll
  for (int w = 0; w < SVB / 64; ++w) {
    if ((uint32_t)w < wave) woff += s_wcnt[w];
    T += s_wcnt[w];
  }
  if (tid == 0) lookback_publish(desc, tile, 1u, 0u, T);
  if (tid < 64) {
    const uint32_t ex = wave_lookback_u32(desc, tile, T, err);
    if (tid == 0) s_excl = ex;
  }
  __syncthreads();
  if (g >= tot) return;
  const uint64_t pos = (uint64_t)s_excl + woff + below;
  if (gl == 0) {  // this query's output starts here, as do those of empty queries before it
    out_start[q] = pos;
    for (uint32_t p = q; p > 0 && s_off[p - 1] == s_off[q]; --p) out_start[p - 1] = pos;
  }
  if (g == tot - 1)  // the end of the output, and the starts of trailing empty queries
    for (uint32_t p = q + 1; p <= Q; ++p) out_start[p] = pos + (pass ? 1u : 0u);
  if (pass) {
    out_keys[pos] = static_cast<const KT*>(qs[q].q.group_keys)[gl];
    out_count[pos] = (int64_t)e.cnt;
    out_sum[pos] = (V)e.s;
    out_mean[pos] = (V)(e.s / (double)e.cnt);
  }
}

__global__ void k_served_zero(uint32_t* w, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) w[i] = 0u;
}

namespace {

struct SvWs {
  SvDev* table;
  uint64_t* start;
  uint32_t* desc;    // one descriptor per tile, then the ticket counter
};

SvWs carve_served(Carver& c, uint32_t Q, uint64_t tot) {
  SvWs w{};
  w.table = c.take<SvDev>(Q);
  w.start = c.take<uint64_t>((uint64_t)Q + 1);
  w.desc = c.take<uint32_t>((tot + SVB - 1) / SVB + 1);
  return w;
}

template <typename V, typename KT>
int served_batch_typed(const SvWs& w, uint32_t Q, uint64_t tot, void* ok, int64_t* oc, void* os,
                       void* om, hipStream_t s) {
  const uint64_t nt = (tot + SVB - 1) / SVB;
  uint32_t* err = device_error_word();
  JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  JFX_LAUNCH("served_zero", s, (k_served_zero<<<(uint32_t)std::min<uint64_t>(64, (nt + 256) / 256),
                                                 256, 0, s>>>(w.desc, nt + 1)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("served_batch", s, (k_served_batch<V, KT><<<(uint32_t)nt, SVB, 0, s>>>(
                                    w.table, Q, tot, w.desc, w.desc + nt, err,
                                    static_cast<KT*>(ok), oc, static_cast<V*>(os),
                                    static_cast<V*>(om), w.start)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // namespace

// ---- served join: a resident frame's group index against a dimension table --
// The frame is indexed by its key groups (group_keys ascending, group g's rows
// at sorted positions [starts[g], starts[g+1])), so the join runs over groups
// and dimension rows instead of frame rows:
//   k_sj_code   dimension row -> its group (binary search in group_keys), or G
//   (radix sort of the codes with the row index, fixed plan: codes < 2^bits)
//   k_sj_runs   first and end sorted dimension position of every group's run
//   k_sj_tiles  per tile of SJ_TILE groups: pairs = group rows x dimension rows
//   k_join_scan tile offsets and the pair count M
//   k_sj_offs   every group's first output position (off[G] = M)
//   k_sj_fill   SJ_CHUNK outputs per workgroup: output k of group g is left
//               position starts[g] + q / c, dimension row of sorted position
//               dfirst[g] + q % c (q = k - off[g], c = the group's dimension
//               rows): (key, left row, right row) order, the reference's
//               pair order (parallel_algorithms.py:987-994).
constexpr uint32_t SJ_TILE = 1024;
constexpr uint32_t SJ_CHUNK = 2048;

template <typename K>
__global__ void __launch_bounds__(256)
k_sj_code(const K* gk, uint32_t G, const K* dk, uint64_t R, int oc, uint32_t* code) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < R; i += stride) {
    const K x = ordered(dk[i], oc);
    uint32_t lo = 0, hi = G;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ordered(gk[mid], oc) < x) lo = mid + 1;
      else hi = mid;
    }
    code[i] = (lo < G && ordered(gk[lo], oc) == x) ? lo : G;
  }
}

__global__ void __launch_bounds__(256)
k_sj_runs(const uint32_t* sc, uint64_t R, uint32_t G, uint32_t* dfirst, uint32_t* dend) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < R; j += stride) {
    const uint32_t c = sc[j];
    if (c >= G) continue;
    if (j == 0 || sc[j - 1] != c) dfirst[c] = (uint32_t)j;
    if (j + 1 == R || sc[j + 1] != c) dend[c] = (uint32_t)(j + 1);
  }
}

__device__ __forceinline__ uint64_t sj_pairs(const uint32_t* starts, const uint32_t* dfirst,
                                             const uint32_t* dend, uint32_t g) {
  return (uint64_t)(starts[g + 1] - starts[g]) * (uint64_t)(dend[g] - dfirst[g]);
}

// per-thread groups tid*4 .. tid*4+3 of the tile; returns the thread's sum
__device__ __forceinline__ uint64_t sj_thread_pairs(const uint32_t* starts, const uint32_t* dfirst,
                                                    const uint32_t* dend, uint32_t G, uint32_t g0,
                                                    uint64_t* p) {
  uint64_t t = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t g = g0 + threadIdx.x * 4 + j;
    p[j] = g < G ? sj_pairs(starts, dfirst, dend, g) : 0ull;
    t += p[j];
  }
  return t;
}

__global__ void __launch_bounds__(256)
k_sj_tiles(const uint32_t* starts, const uint32_t* dfirst, const uint32_t* dend, uint32_t G,
           uint64_t* tile_sum) {
  __shared__ uint64_t s_w[4];
  uint64_t p[4];
  uint64_t t = sj_thread_pairs(starts, dfirst, dend, G, blockIdx.x * SJ_TILE, p);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ void __launch_bounds__(256)
k_sj_offs(const uint32_t* starts, const uint32_t* dfirst, const uint32_t* dend, uint32_t G,
          const uint64_t* tile_off, const uint64_t* total, uint64_t* off) {
  __shared__ uint64_t s_w[5];
  uint64_t p[4];
  const uint32_t g0 = blockIdx.x * SJ_TILE;
  const uint64_t t = sj_thread_pairs(starts, dfirst, dend, G, g0, p);
  const uint64_t incl = wave_incl_scan64(t);
  if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint64_t wbase = 0;
  for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) wbase += s_w[w];
  uint64_t run = tile_off[blockIdx.x] + wbase + incl - t;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t g = g0 + threadIdx.x * 4 + j;
    if (g < G) off[g] = run;
    run += p[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) off[G] = *total;
}

// Exclusive scan of n tile sums in place by one workgroup of 1024 threads;
// the total goes to *total.
__global__ void __launch_bounds__(1024) k_sj_scan(uint64_t* c, uint32_t n, uint64_t* total) {
  __shared__ uint64_t s_w[17];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t a = tid * per, b = (a + per) < n ? (a + per) : n;
  uint64_t sum = 0;
  for (uint32_t i = a; i < b; ++i) sum += c[i];
  const uint64_t incl = wave_incl_scan64(sum);
  if ((tid & 63) == 63) s_w[tid >> 6] = incl;
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

template <typename K, typename LT, typename RT>
__global__ void __launch_bounds__(256)
k_sj_fill(const K* gk, const uint32_t* starts, const uint32_t* dfirst, const uint32_t* dend,
          const uint32_t* dperm, const uint64_t* off, uint32_t G, uint64_t M, const LT* lv,
          const RT* rv, K* out_key, LT* out_lv, RT* out_rv) {
  __shared__ uint32_t s_g[2];
  const uint64_t k0 = (uint64_t)blockIdx.x * SJ_CHUNK;
  const uint64_t k1 = k0 + SJ_CHUNK < M ? k0 + SJ_CHUNK : M;
  if (threadIdx.x < 2) {  // last group whose first output is <= k0 (resp. k1 - 1)
    const uint64_t kk = threadIdx.x == 0 ? k0 : k1 - 1;
    uint32_t lo = 0, hi = G + 1;
    while (lo + 1 < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (off[mid] <= kk) lo = mid;
      else hi = mid;
    }
    s_g[threadIdx.x] = lo;
  }
  __syncthreads();
  const uint32_t ga = s_g[0], gb = s_g[1];
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += 256) {
    uint32_t g = ga;
    if (ga != gb) {
      uint32_t lo = ga, hi = gb + 1;
      while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= k) lo = mid;
        else hi = mid;
      }
      g = lo;
    }
    const uint64_t q = k - off[g];
    const uint32_t df = dfirst[g], c = dend[g] - df;
    uint64_t pidx = q, d = 0;
    if (c != 1) {
      pidx = q / c;
      d = q - pidx * c;
    }
    out_key[k] = gk[g];
    out_lv[k] = lv[starts[g] + pidx];
    out_rv[k] = rv[dperm[df + d]];
  }
}

namespace {

struct SjWs {
  uint32_t* code;
  uint32_t* scode;
  uint32_t* dperm;
  uint32_t* dfirst;   // dfirst[G] then dend[G], zeroed together
  uint32_t* dend;
  uint64_t* tile_off;
  uint64_t* total;
  uint64_t* off;
  void* sort_ws;
  uint64_t sort_bytes;
};

uint32_t sj_bits(uint32_t G) {
  uint32_t b = 1;
  while (b < 32 && (G >> b) != 0) ++b;
  return b;
}

SjWs carve_sj(Carver& c, uint64_t R, uint32_t G) {
  SjWs w{};
  const uint64_t r = R ? R : 1;
  w.code = c.take<uint32_t>(r);
  w.scode = c.take<uint32_t>(r);
  w.dperm = c.take<uint32_t>(r);
  w.dfirst = c.take<uint32_t>(2ull * G + 2);
  w.dend = w.dfirst + G + 1;
  const uint64_t nt = (G + SJ_TILE - 1) / SJ_TILE;
  w.tile_off = c.take<uint64_t>(nt + 1);
  w.total = c.take<uint64_t>(2);
  w.off = c.take<uint64_t>((uint64_t)G + 1);
  w.sort_bytes = sort_ws_bytes(JFX_U32, 4, r);
  w.sort_ws = c.take<char>(w.sort_bytes);
  return w;
}

template <typename K>
int sj_count_typed(int oc, const void* gk, const uint32_t* starts, uint32_t G, const void* dk,
                   uint64_t R, const SjWs& w, uint64_t* out_m, hipStream_t s) {
  const uint32_t cb = (uint32_t)std::min<uint64_t>(4096, (R + 255) / 256);
  JFX_LAUNCH("sj_code", s, (k_sj_code<K><<<cb, 256, 0, s>>>((const K*)gk, G, (const K*)dk, R, oc,
                                                            w.code)));
  JFX_HIP(hipGetLastError());
  SortCtx ctx;
  int rc = sort_prepare_known(ctx, JFX_U32, w.code, R, 4, 0, sj_bits(G), w.sort_ws, w.sort_bytes, s);
  if (rc != JFX_OK) return rc;
  rc = sort_execute(ctx, w.code, w.scode, 1, nullptr, w.dperm, 1, 0);
  if (rc != JFX_OK) return rc;
  JFX_LAUNCH("sj_zero", s, (k_served_zero<<<64, 256, 0, s>>>(w.dfirst, 2ull * G + 2)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("sj_runs", s, (k_sj_runs<<<cb, 256, 0, s>>>(w.scode, R, G, w.dfirst, w.dend)));
  JFX_HIP(hipGetLastError());
  const uint32_t nt = (G + SJ_TILE - 1) / SJ_TILE;
  JFX_LAUNCH("sj_tiles", s, (k_sj_tiles<<<nt, 256, 0, s>>>(starts, w.dfirst, w.dend, G, w.tile_off)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("sj_scan", s, (k_sj_scan<<<1, 1024, 0, s>>>(w.tile_off, nt, w.total)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("sj_offs", s, (k_sj_offs<<<nt, 256, 0, s>>>(starts, w.dfirst, w.dend, G, w.tile_off,
                                                         w.total, w.off)));
  JFX_HIP(hipGetLastError());
  return read_back(s, w.total, sizeof(uint64_t), out_m);
}

template <typename K, typename LT, typename RT>
int sj_fill_typed(const void* gk, const uint32_t* starts, uint32_t G, const SjWs& w, uint64_t M,
                  const void* lv, const void* rv, void* ok, void* ol, void* orr, hipStream_t s) {
  const uint64_t nb = (M + SJ_CHUNK - 1) / SJ_CHUNK;
  JFX_LAUNCH("sj_fill", s, (k_sj_fill<K, LT, RT><<<(uint32_t)nb, 256, 0, s>>>(
                               (const K*)gk, starts, w.dfirst, w.dend, w.dperm, w.off, G, M,
                               (const LT*)lv, (const RT*)rv, (K*)ok, (LT*)ol, (RT*)orr)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K>
int sj_fill_key(int lb, int rb, const void* gk, const uint32_t* starts, uint32_t G, const SjWs& w,
                uint64_t M, const void* lv, const void* rv, void* ok, void* ol, void* orr,
                hipStream_t s) {
  if (lb == 4 && rb == 4) return sj_fill_typed<K, uint32_t, uint32_t>(gk, starts, G, w, M, lv, rv, ok, ol, orr, s);
  if (lb == 4) return sj_fill_typed<K, uint32_t, uint64_t>(gk, starts, G, w, M, lv, rv, ok, ol, orr, s);
  if (rb == 4) return sj_fill_typed<K, uint64_t, uint32_t>(gk, starts, G, w, M, lv, rv, ok, ol, orr, s);
  return sj_fill_typed<K, uint64_t, uint64_t>(gk, starts, G, w, M, lv, rv, ok, ol, orr, s);
}

}  // namespace

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

int jfx_served_join_workspace_size(uint64_t n_dim, uint32_t ngroups, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  jfx::Carver c(nullptr, 0);
  jfx::carve_sj(c, n_dim, ngroups);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_served_join_count(int key_type, const void* group_keys, const uint32_t* group_starts,
                          uint32_t ngroups, const void* dim_keys, uint64_t n_dim,
                          uint64_t* out_pairs, void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type == JFX_U32 || key_type == JFX_I32 || key_type == JFX_U64 ||
                key_type == JFX_I64,
            JFX_ERR_UNSUPPORTED, "served joins take integer keys");
  JFX_CHECK(out_pairs != nullptr, JFX_ERR_INVALID, "out_pairs is null");
  JFX_CHECK(n_dim < (1ull << 32), JFX_ERR_UNSUPPORTED, "the dimension table must have < 2^32 rows");
  JFX_CHECK(ngroups < 0xFFFFFFFFu, JFX_ERR_UNSUPPORTED, "too many groups");
  *out_pairs = 0;
  if (n_dim == 0 || ngroups == 0) return JFX_OK;
  JFX_CHECK(group_keys && group_starts && dim_keys && workspace, JFX_ERR_INVALID, "null argument");
  jfx::Carver c(workspace, ws_bytes);
  jfx::SjWs w = jfx::carve_sj(c, n_dim, ngroups);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "served join workspace too small");
  const int oc = jfx::order_code(key_type);
  hipStream_t s = jfx::as_stream(stream);
  if (jfx::key_bytes(key_type) == 4)
    return jfx::sj_count_typed<uint32_t>(oc, group_keys, group_starts, ngroups, dim_keys, n_dim, w,
                                         out_pairs, s);
  return jfx::sj_count_typed<uint64_t>(oc, group_keys, group_starts, ngroups, dim_keys, n_dim, w,
                                       out_pairs, s);
}

int jfx_served_join_fill(int key_type, const void* group_keys, const uint32_t* group_starts,
                         uint32_t ngroups, int left_elem_bytes, const void* left_values,
                         uint64_t n_dim, int right_elem_bytes, const void* dim_values,
                         uint64_t n_pairs, void* out_keys, void* out_left, void* out_right,
                         void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type == JFX_U32 || key_type == JFX_I32 || key_type == JFX_U64 ||
                key_type == JFX_I64,
            JFX_ERR_UNSUPPORTED, "served joins take integer keys");
  JFX_CHECK((left_elem_bytes == 4 || left_elem_bytes == 8) &&
                (right_elem_bytes == 4 || right_elem_bytes == 8),
            JFX_ERR_UNSUPPORTED, "payload elements must be 4 or 8 bytes");
  if (n_pairs == 0) return JFX_OK;
  JFX_CHECK(group_keys && group_starts && left_values && dim_values && out_keys && out_left &&
                out_right && workspace,
            JFX_ERR_INVALID, "null argument");
  JFX_CHECK(n_pairs / jfx::SJ_CHUNK < 0xFFFFFFFFull, JFX_ERR_UNSUPPORTED, "too many pairs");
  jfx::Carver c(workspace, ws_bytes);
  jfx::SjWs w = jfx::carve_sj(c, n_dim, ngroups);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "served join workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  if (jfx::key_bytes(key_type) == 4)
    return jfx::sj_fill_key<uint32_t>(left_elem_bytes, right_elem_bytes, group_keys, group_starts,
                                      ngroups, w, n_pairs, left_values, dim_values, out_keys,
                                      out_left, out_right, s);
  return jfx::sj_fill_key<uint64_t>(left_elem_bytes, right_elem_bytes, group_keys, group_starts,
                                    ngroups, w, n_pairs, left_values, dim_values, out_keys, out_left,
                                    out_right, s);
}

int jfx_served_batch_workspace_size(uint32_t nqueries, uint64_t total_groups,
                                    uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  jfx::Carver c(nullptr, 0);
  jfx::carve_served(c, nqueries, total_groups);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_served_batch(int val_type, int key_bytes, const jfx_served_query* queries,
                     uint32_t nqueries, void* out_keys, int64_t* out_count, void* out_sum,
                     void* out_mean, uint64_t* out_starts, void* workspace, uint64_t ws_bytes,
                     void* stream) {
  JFX_CHECK(val_type == JFX_F32 || val_type == JFX_F64, JFX_ERR_UNSUPPORTED,
            "indexed filters take float32 / float64 values");
  JFX_CHECK(key_bytes == 4 || key_bytes == 8, JFX_ERR_UNSUPPORTED, "group keys must be 4 or 8 bytes");
  JFX_CHECK(out_starts != nullptr, JFX_ERR_INVALID, "out_starts is null");
  JFX_CHECK(nqueries <= jfx::SV_QMAX, JFX_ERR_UNSUPPORTED, "at most 1024 queries per batch");
  if (nqueries == 0) {
    out_starts[0] = 0;
    return JFX_OK;
  }
  JFX_CHECK(queries != nullptr, JFX_ERR_INVALID, "queries is null");
  uint64_t tot = 0;
  for (uint32_t i = 0; i < nqueries; ++i) {
    const jfx_served_query& q = queries[i];
    JFX_CHECK(q.cmp >= JFX_CMP_GT && q.cmp <= JFX_CMP_NE, JFX_ERR_INVALID, "bad comparison code");
    JFX_CHECK(q.ngroups == 0 || (q.values && q.prefix && q.group_starts && q.nan_start &&
                                 q.ninf_end && q.pinf_start && q.group_keys),
              JFX_ERR_INVALID, "null query column");
    tot += q.ngroups;
  }
  JFX_CHECK(tot < (1ull << 30), JFX_ERR_UNSUPPORTED, "a batch holds fewer than 2^30 groups");
  if (tot == 0) {
    for (uint32_t i = 0; i <= nqueries; ++i) out_starts[i] = 0;
    return JFX_OK;
  }
  JFX_CHECK(out_keys && out_count && out_sum && out_mean && workspace, JFX_ERR_INVALID,
            "null output or workspace");
  jfx::Carver c(workspace, ws_bytes);
  jfx::SvWs w = jfx::carve_served(c, nqueries, tot);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "served batch workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  // query table and the read-back of the output starts share one pinned block
  const size_t tb = sizeof(jfx::SvDev) * nqueries, sb = sizeof(uint64_t) * (nqueries + 1);
  char* h = static_cast<char*>(jfx::pinned_scratch(tb));
  JFX_CHECK(h != nullptr, JFX_ERR_HIP, "pinned host allocation failed");
  jfx::SvDev* ht = reinterpret_cast<jfx::SvDev*>(h);
  uint64_t off = 0;
  for (uint32_t i = 0; i < nqueries; ++i) {
    ht[i].q = queries[i];
    ht[i].goff = off;
    off += queries[i].ngroups;
  }
  JFX_HIP(hipMemcpyAsync(w.table, ht, tb, hipMemcpyHostToDevice, s));
  int rc;
  if (val_type == JFX_F32)
    rc = key_bytes == 4 ? jfx::served_batch_typed<float, uint32_t>(w, nqueries, tot, out_keys, out_count, out_sum, out_mean, s)
                        : jfx::served_batch_typed<float, uint64_t>(w, nqueries, tot, out_keys, out_count, out_sum, out_mean, s);
  else
    rc = key_bytes == 4 ? jfx::served_batch_typed<double, uint32_t>(w, nqueries, tot, out_keys, out_count, out_sum, out_mean, s)
                        : jfx::served_batch_typed<double, uint64_t>(w, nqueries, tot, out_keys, out_count, out_sum, out_mean, s);
  if (rc != JFX_OK) return rc;
  // the read-back runs after the table upload on the stream, so the staging
  // block is free again once it returns
  return jfx::read_back(s, w.start, sb, out_starts);
}

}  // extern "C"
