// Sort-based groupby: head-flag segmented reductions over sorted keys.
//
// Restates SortBasedGroupBy.groupby_aggregate (parallel_algorithms.py:733-803)
// and its aggregations (parallel_algorithms.py:626-660): heads where a sorted
// key differs from its predecessor (:767-770), one group per head in ascending
// key order, sum / mean / count / min / max per group. The reference assigns
// segment ids with cumsum and scatter-adds (:777, :630); here:
//   k_segreduce_tiles  one read of keys + values per tile. Head counts get
//                      global segment ids through a decoupled lookback;
//                      each thread reduces its 16 rows sequentially and the
//                      workgroup runs a wave64 segmented scan over the thread
//                      partials. Groups that start and end inside the tile
//                      are written here; each tile leaves its leading and
//                      trailing partial.
//   k_segreduce_fix    one workgroup scans the per-tile partials and writes
//                      the groups that cross tile boundaries, plus G.
#include <type_traits>

#include "blockscan.hpp"
#include "runtime.hpp"
#include "segreduce.hpp"

namespace jfx {

constexpr int SEG_BLOCK = 512;
constexpr int SEG_IPT = 16;
constexpr int SEG_TILE = SEG_BLOCK * SEG_IPT;
constexpr int SEG_W = SEG_BLOCK / 64;

__host__ __device__ inline uint32_t seg_pad(uint32_t e) { return e + (e >> 5); }

template <typename K, typename V, typename MT>
struct SegArgs {
  const K* keys;
  int oc;
  const V* vals;
  const uint32_t* perm;  // value of row i is vals[perm[i]] when non-null
  uint64_t n;
  K* out_keys;
  V* out_sum;
  int64_t* out_count;
  V* out_min;
  V* out_max;
  MT* out_mean;
  AggState<V>* tile_lead;   // partial before the tile's first head (whole tile if none)
  AggState<V>* tile_trail;  // partial from the tile's last head to its end
  uint2* tile_info;         // {heads before the tile, heads in the tile}
  uint32_t* desc;
  uint32_t* tile_counter;
  uint32_t* err;
  uint32_t* ngroups;        // device scalar G
  uint32_t num_tiles;
};

template <typename V, typename MT>
__device__ __forceinline__ MT mean_of(const AggState<V>& st) {
  if constexpr (ValTraits<V>::is_float) {
    return (MT)st.sum() / (MT)(st.n ? st.n : 1u);
  } else {
    return (MT)((double)st.s / (double)(st.n ? st.n : 1u));
  }
}

template <typename K, typename V, typename MT>
__device__ __forceinline__ void finalize(const SegArgs<K, V, MT>& a, uint32_t g, const AggState<V>& st) {
  if (a.out_sum) a.out_sum[g] = st.sum();
  if (a.out_count) a.out_count[g] = (int64_t)st.n;
  if (a.out_min) a.out_min[g] = st.mn;
  if (a.out_max) a.out_max[g] = st.mx;
  if (a.out_mean) a.out_mean[g] = mean_of<V, MT>(st);
}

// (flag, state) segmented combine: earlier then later.
template <typename V>
__device__ __forceinline__ void seg_combine(uint32_t& fa, AggState<V>& a, uint32_t fb, const AggState<V>& b) {
  if (fb) {
    a = b;
  } else {
    a.merge(b);
  }
  fa |= fb;
}

// Workgroup exclusive segmented scan of (f, st) in thread order. Returns the
// exclusive value in (ef, est) and the workgroup total in (tf, tst).
template <typename V, int W>
__device__ __forceinline__ void block_seg_scan(uint32_t f, const AggState<V>& st, uint32_t& ef,
                                               AggState<V>& est, uint32_t& tf, AggState<V>& tst,
                                               AggState<V>* s_wst, uint32_t* s_wf) {
  const uint32_t lane = lane_id();
  const int wid = threadIdx.x >> 6;
  uint32_t incf = f;
  AggState<V> inc = st;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    AggState<V> up = shfl_up_state(inc, off);
    const uint32_t fu = __shfl_up(incf, off, 64);
    if (lane >= (uint32_t)off) {
      if (!incf) {
        up.merge(inc);
        inc = up;
      }
      incf |= fu;
    }
  }
  // exclusive within the wave
  AggState<V> ex = shfl_up_state(inc, 1);
  uint32_t exf = __shfl_up(incf, 1, 64);
  if (lane == 0) {
    ex = AggState<V>::identity();
    exf = 0u;
  }
  if (lane == 63) {
    s_wst[wid] = inc;
    s_wf[wid] = incf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t cf = 0u;
    AggState<V> c = AggState<V>::identity();
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const AggState<V> t = s_wst[w];
      const uint32_t tfw = s_wf[w];
      s_wst[w] = c;
      s_wf[w] = cf;
      seg_combine(cf, c, tfw, t);
    }
    s_wst[W] = c;
    s_wf[W] = cf;
  }
  __syncthreads();
  ef = s_wf[wid];
  est = s_wst[wid];
  seg_combine(ef, est, exf, ex);
  tf = s_wf[W];
  tst = s_wst[W];
  __syncthreads();
}

template <typename K, typename V, typename MT>
__global__ void __launch_bounds__(SEG_BLOCK) k_segreduce_tiles(SegArgs<K, V, MT> a) {
  __shared__ K s_k[SEG_TILE + SEG_TILE / 32];
  __shared__ V s_v[SEG_TILE + SEG_TILE / 32];
  __shared__ uint32_t s_cnt[SEG_BLOCK];
  __shared__ uint32_t s_tmp[SEG_W + 1];
  __shared__ AggState<V> s_wst[SEG_W + 1];
  __shared__ uint32_t s_wf[SEG_W + 1];
  __shared__ uint32_t s_tile, s_excl;

  const int tid = threadIdx.x;
  if (tid == 0) s_tile = atomicAdd(a.tile_counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tbase = (uint64_t)tile * SEG_TILE;
  const uint32_t tn = (uint32_t)((a.n - tbase) < (uint64_t)SEG_TILE ? (a.n - tbase) : (uint64_t)SEG_TILE);

  // striped (coalesced) loads into LDS in row order
#pragma unroll
  for (int i = 0; i < SEG_IPT; ++i) {
    const uint32_t e = (uint32_t)i * SEG_BLOCK + tid;
    if (e < tn) {
      s_k[seg_pad(e)] = a.keys[tbase + e];
      s_v[seg_pad(e)] = a.perm ? a.vals[a.perm[tbase + e]] : a.vals[tbase + e];
    }
  }
  K prevk = K(0);
  if (tid == 0 && tbase > 0) prevk = a.keys[tbase - 1];
  __syncthreads();

  // blocked: thread t owns rows t*IPT .. t*IPT+IPT-1 of the tile
  const uint32_t e0 = (uint32_t)tid * SEG_IPT;
  if (tid > 0 && e0 - 1 < tn) prevk = s_k[seg_pad(e0 - 1)];
  K kj[SEG_IPT];
  V vj[SEG_IPT];
  uint32_t hmask = 0u;
#pragma unroll
  for (int j = 0; j < SEG_IPT; ++j) {
    const uint32_t e = e0 + j;
    if (e < tn) {
      kj[j] = s_k[seg_pad(e)];
      vj[j] = s_v[seg_pad(e)];
      const K pk = j ? kj[j - 1] : prevk;
      const bool h = (tbase + e == 0) || !key_equal(kj[j], pk, a.oc);
      hmask |= (uint32_t)h << j;
    }
  }
  s_cnt[tid] = (uint32_t)__popc(hmask);
  __syncthreads();
  const uint32_t T = block_excl_scan_lds<SEG_BLOCK, SEG_BLOCK>(s_cnt, s_tmp);
  const uint32_t toff = s_cnt[tid];
  if (tid == 0) {
    lookback_publish(a.desc, tile, 1u, 0u, T);
    s_excl = lookback_u32(a.desc, tile, 1u, 0u, T, a.err);
  }
  __syncthreads();
  const uint32_t texcl = s_excl;

  // sequential reduction of the thread's rows
  AggState<V> cur = AggState<V>::identity();
  AggState<V> lead = AggState<V>::identity();
  uint32_t seen = 0u;
  uint32_t gid = texcl + toff - 1u;  // segment open at the thread's start
#pragma unroll
  for (int j = 0; j < SEG_IPT; ++j) {
    const uint32_t e = e0 + j;
    if (e < tn) {
      if ((hmask >> j) & 1u) {
        if (seen) finalize(a, gid, cur);
        else lead = cur;
        seen = 1u;
        ++gid;
        if (a.out_keys) a.out_keys[gid] = kj[j];
        cur = AggState<V>::identity();
      }
      cur.add(vj[j]);
    }
  }
  if (!seen) lead = cur;

  uint32_t ef, tf;
  AggState<V> est, tst;
  block_seg_scan<V, SEG_W>(seen, seen ? cur : lead, ef, est, tf, tst, s_wst, s_wf);
  if (seen) {
    AggState<V> x = est;
    x.merge(lead);
    if (ef) finalize(a, texcl + toff - 1u, x);  // group began inside this tile
    else a.tile_lead[tile] = x;                  // group began before this tile
  }
  if (tid == 0) {
    if (T == 0u) a.tile_lead[tile] = tst;
    else a.tile_trail[tile] = tst;
    a.tile_info[tile] = make_uint2(texcl, T);
  }
}

template <typename K, typename V, typename MT>
__global__ void __launch_bounds__(1024) k_segreduce_fix(SegArgs<K, V, MT> a) {
  __shared__ AggState<V> s_wst[16 + 1];
  __shared__ uint32_t s_wf[16 + 1];
  const int tid = threadIdx.x;
  uint32_t cf = 0u;
  AggState<V> carry = AggState<V>::identity();
  for (uint32_t c0 = 0; c0 < a.num_tiles; c0 += 1024u) {
    const uint32_t u = c0 + tid;
    const bool valid = u < a.num_tiles;
    uint2 info = valid ? a.tile_info[u] : make_uint2(0u, 0u);
    const uint32_t hh = info.y > 0u ? 1u : 0u;
    AggState<V> lead = valid ? a.tile_lead[u] : AggState<V>::identity();
    AggState<V> el = hh ? a.tile_trail[u] : lead;
    uint32_t ef, tf;
    AggState<V> est, tst;
    block_seg_scan<V, 16>(valid ? hh : 0u, el, ef, est, tf, tst, s_wst, s_wf);
    uint32_t ff = cf;
    AggState<V> full = carry;
    seg_combine(ff, full, ef, est);
    if (valid && hh && info.x > 0u) {
      full.merge(lead);
      finalize(a, info.x - 1u, full);
    }
    seg_combine(cf, carry, tf, tst);
  }
  if (tid == 0) {
    const uint2 last = a.tile_info[a.num_tiles - 1];
    const uint32_t G = last.x + last.y;
    if (G > 0u) finalize(a, G - 1u, carry);
    *a.ngroups = G;
  }
}

namespace {

struct SegWs {
  void* lead;
  void* trail;
  uint2* info;
  uint32_t* desc;
  uint32_t* counter;
  uint32_t* ngroups;
  uint64_t bytes;
};

uint64_t seg_tiles(uint64_t n) { return (n + SEG_TILE - 1) / SEG_TILE; }

SegWs carve_seg(Carver& c, int vb, uint64_t n) {
  SegWs w{};
  const uint64_t t = seg_tiles(n);
  const uint64_t st = vb == 4 ? sizeof(AggState<float>) : sizeof(AggState<double>);
  w.lead = c.take<char>(t * st);
  w.trail = c.take<char>(t * st);
  w.info = c.take<uint2>(t);
  w.desc = c.take<uint32_t>(t);
  w.counter = c.take<uint32_t>(4);
  w.ngroups = c.take<uint32_t>(4);
  return w;
}

__global__ void k_seg_zero(uint32_t* desc, uint64_t ntiles, uint32_t* counter) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += stride) desc[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

template <typename K, typename V, typename MT>
int segreduce_typed(const SegWs& w, const void* keys, int oc, const void* vals, const uint32_t* perm,
                    uint64_t n, void* out_keys, void* out_sum, int64_t* out_count, void* out_min,
                    void* out_max, void* out_mean, hipStream_t s) {
  SegArgs<K, V, MT> a{};
  a.keys = static_cast<const K*>(keys);
  a.oc = oc;
  a.vals = static_cast<const V*>(vals);
  a.perm = perm;
  a.n = n;
  a.out_keys = static_cast<K*>(out_keys);
  a.out_sum = static_cast<V*>(out_sum);
  a.out_count = out_count;
  a.out_min = static_cast<V*>(out_min);
  a.out_max = static_cast<V*>(out_max);
  a.out_mean = static_cast<MT*>(out_mean);
  a.tile_lead = static_cast<AggState<V>*>(w.lead);
  a.tile_trail = static_cast<AggState<V>*>(w.trail);
  a.tile_info = w.info;
  a.desc = w.desc;
  a.tile_counter = w.counter;
  a.err = device_error_word();
  a.ngroups = w.ngroups;
  a.num_tiles = (uint32_t)seg_tiles(n);
  JFX_CHECK(a.err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  k_seg_zero<<<64, 256, 0, s>>>(w.desc, a.num_tiles, w.counter);
  JFX_HIP(hipGetLastError());
  k_segreduce_tiles<K, V, MT><<<a.num_tiles, SEG_BLOCK, 0, s>>>(a);
  JFX_HIP(hipGetLastError());
  k_segreduce_fix<K, V, MT><<<1, 1024, 0, s>>>(a);
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K>
int segreduce_key(const SegWs& w, int vt, const void* keys, int oc, const void* vals,
                  const uint32_t* perm, uint64_t n, void* out_keys, void* out_sum,
                  int64_t* out_count, void* out_min, void* out_max, void* out_mean, hipStream_t s) {
  switch (vt) {
    case JFX_F32:
      return segreduce_typed<K, float, float>(w, keys, oc, vals, perm, n, out_keys, out_sum, out_count,
                                              out_min, out_max, out_mean, s);
    case JFX_F64:
      return segreduce_typed<K, double, double>(w, keys, oc, vals, perm, n, out_keys, out_sum,
                                                out_count, out_min, out_max, out_mean, s);
    case JFX_I32:
      return segreduce_typed<K, int32_t, float>(w, keys, oc, vals, perm, n, out_keys, out_sum,
                                                out_count, out_min, out_max, out_mean, s);
    case JFX_I64:
      return segreduce_typed<K, int64_t, double>(w, keys, oc, vals, perm, n, out_keys, out_sum,
                                                 out_count, out_min, out_max, out_mean, s);
    case JFX_U32:
      return segreduce_typed<K, uint32_t, float>(w, keys, oc, vals, perm, n, out_keys, out_sum,
                                                 out_count, out_min, out_max, out_mean, s);
    case JFX_U64:
      return segreduce_typed<K, uint64_t, double>(w, keys, oc, vals, perm, n, out_keys, out_sum,
                                                  out_count, out_min, out_max, out_mean, s);
  }
  return set_error(JFX_ERR_INVALID, "bad value type");
}

int segreduce_dispatch(const SegWs& w, int kt, const void* keys, int vt, const void* vals,
                       const uint32_t* perm, uint64_t n, void* out_keys, void* out_sum,
                       int64_t* out_count, void* out_min, void* out_max, void* out_mean, hipStream_t s) {
  const int oc = order_code(kt);
  if (key_bytes(kt) == 4)
    return segreduce_key<uint32_t>(w, vt, keys, oc, vals, perm, n, out_keys, out_sum, out_count,
                                   out_min, out_max, out_mean, s);
  return segreduce_key<uint64_t>(w, vt, keys, oc, vals, perm, n, out_keys, out_sum, out_count,
                                 out_min, out_max, out_mean, s);
}

int read_ngroups(const SegWs& w, uint64_t* out, hipStream_t s) {
  uint32_t* h = static_cast<uint32_t*>(pinned_scratch(sizeof(uint32_t)));
  JFX_CHECK(h != nullptr, JFX_ERR_HIP, "pinned host allocation failed");
  JFX_HIP(hipMemcpyAsync(h, w.ngroups, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  JFX_HIP(hipStreamSynchronize(s));
  *out = *h;
  return JFX_OK;
}

bool valid_vt(int vt) { return vt >= JFX_U32 && vt <= JFX_F64; }

}  // namespace

int sort_dispatch(int kt, const void* keys_in, void* keys_out, int want_keys, int pay_bytes,
                  const void* pay_in, void* pay_out, uint64_t n, int descending, int digit_bits,
                  void* ws, uint64_t ws_bytes, hipStream_t s);
uint64_t sort_ws_bytes(int kt, int pay_bytes, uint64_t n);
int default_digit_bits(int kt, uint64_t n);

}  // namespace jfx

extern "C" {

int jfx_segreduce_workspace_size(int val_type, uint64_t n, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  jfx::Carver c(nullptr, 0);
  jfx::carve_seg(c, jfx::key_bytes(val_type), n);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_segreduce(int key_type, const void* sorted_keys, int val_type, const void* vals,
                  const uint32_t* perm, uint64_t n, void* out_keys, void* out_sum,
                  int64_t* out_count, void* out_min, void* out_max, void* out_mean,
                  uint64_t* out_ngroups, void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  JFX_CHECK(n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be below 2^30 rows");
  if (n == 0) {
    if (out_ngroups) *out_ngroups = 0;
    return JFX_OK;
  }
  JFX_CHECK(sorted_keys != nullptr && vals != nullptr && workspace != nullptr, JFX_ERR_INVALID,
            "null input or workspace");
  jfx::Carver c(workspace, ws_bytes);
  jfx::SegWs w = jfx::carve_seg(c, jfx::key_bytes(val_type), n);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "segreduce workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  int rc = jfx::segreduce_dispatch(w, key_type, sorted_keys, val_type, vals, perm, n, out_keys,
                                   out_sum, out_count, out_min, out_max, out_mean, s);
  if (rc != JFX_OK) return rc;
  if (out_ngroups) return jfx::read_ngroups(w, out_ngroups, s);
  return JFX_OK;
}

int jfx_groupby_workspace_size(int key_type, int val_type, uint64_t n, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  const int kb = jfx::key_bytes(key_type), vb = jfx::key_bytes(val_type);
  jfx::Carver c(nullptr, 0);
  c.take<char>(n * kb);
  c.take<char>(n * vb);
  jfx::carve_seg(c, vb, n);
  c.take<char>(jfx::sort_ws_bytes(key_type, vb, n));
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_groupby(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                void* out_keys, void* out_sum, int64_t* out_count, void* out_min, void* out_max,
                void* out_mean, uint64_t* out_ngroups, int digit_bits, void* workspace,
                uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  JFX_CHECK(n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be below 2^30 rows");
  JFX_CHECK(out_ngroups != nullptr, JFX_ERR_INVALID, "out_ngroups is null");
  if (n == 0) {
    *out_ngroups = 0;
    return JFX_OK;
  }
  JFX_CHECK(keys != nullptr && vals != nullptr && workspace != nullptr, JFX_ERR_INVALID,
            "null input or workspace");
  const int kb = jfx::key_bytes(key_type), vb = jfx::key_bytes(val_type);
  jfx::Carver c(workspace, ws_bytes);
  void* skeys = c.take<char>(n * kb);
  void* svals = c.take<char>(n * vb);
  jfx::SegWs w = jfx::carve_seg(c, vb, n);
  const uint64_t sws = jfx::sort_ws_bytes(key_type, vb, n);
  void* sort_ws = c.take<char>(sws);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "groupby workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  if (digit_bits == 0) digit_bits = jfx::default_digit_bits(key_type, n);
  int rc = jfx::sort_dispatch(key_type, keys, skeys, 1, vb, vals, svals, n, 0, digit_bits, sort_ws,
                              sws, s);
  if (rc != JFX_OK) return rc;
  rc = jfx::segreduce_dispatch(w, key_type, skeys, val_type, svals, nullptr, n, out_keys, out_sum,
                               out_count, out_min, out_max, out_mean, s);
  if (rc != JFX_OK) return rc;
  return jfx::read_ngroups(w, out_ngroups, s);
}

}  // extern "C"
