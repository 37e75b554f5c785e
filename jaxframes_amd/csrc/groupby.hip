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
//                      the groups that cross tile boundaries, plus G; above
//                      1024 tiles its chunks of 1024 run in parallel
//                      (k_segfix_chunks, k_segfix_carry, k_segfix_apply).
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "blockscan.hpp"
#include "runtime.hpp"
#include "segreduce.hpp"
#include "sortctx.hpp"
#include "tilereduce.hpp"
#include "wavetable.hpp"
#include "onepass.hpp"
#include "batched.hpp"

namespace jfx {

constexpr int SEG_BLOCK = 512;
constexpr int SEG_IPT = 16;
constexpr int SEG_TILE = SEG_BLOCK * SEG_IPT;
constexpr int SEG_W = SEG_BLOCK / 64;

template <typename K, typename V, typename MT, bool MM>
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
  AggState<V, MM>* tile_lead;   // partial before the tile's first head (whole tile if none)
  AggState<V, MM>* tile_trail;  // partial from the tile's last head to its end
  uint2* tile_info;         // {heads before the tile, heads in the tile}
  uint32_t* desc;
  uint32_t* tile_counter;
  uint32_t* err;
  uint32_t* ngroups;        // device scalar G
  uint32_t num_tiles;
  const uint32_t* seg_start; // known-segment mode: G+1 starts (last = n)
  uint32_t nseg;             // known-segment mode: G (<= 2048)
  int vec;                   // keys, values and perm are 16-byte aligned
};

// First index i in s[0..m) with s[i] >= x (s strictly increasing).
__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t* s, uint32_t m, uint64_t x) {
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)s[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// mean from a group's finished sum and row count
template <typename V, typename MT>
__device__ __forceinline__ MT mean_parts(V sum, uint32_t n) {
  if constexpr (ValTraits<V>::is_float) {
    return (MT)sum / (MT)(n ? n : 1u);
  } else {
    return (MT)((double)sum / (double)(n ? n : 1u));
  }
}

template <typename V, typename MT, bool MM>
__device__ __forceinline__ MT mean_of(const AggState<V, MM>& st) {
  return mean_parts<V, MT>(st.sum(), st.n);
}

template <typename K, typename V, typename MT, bool MM>
__device__ __forceinline__ void finalize(const SegArgs<K, V, MT, MM>& a, uint32_t g,
                                         const AggState<V, MM>& st) {
  if (a.out_sum) a.out_sum[g] = st.sum();
  if (a.out_count) a.out_count[g] = (int64_t)st.n;
  if constexpr (MM) {
    if (a.out_min) a.out_min[g] = st.mn;
    if (a.out_max) a.out_max[g] = st.mx;
  }
  if (a.out_mean) a.out_mean[g] = mean_of<V, MT, MM>(st);
}

// (flag, state) segmented combine: earlier then later.
template <typename V, bool MM>
__device__ __forceinline__ void seg_combine(uint32_t& fa, AggState<V, MM>& a, uint32_t fb,
                                            const AggState<V, MM>& b) {
  if (fb) {
    a = b;
  } else {
    a.merge(b);
  }
  fa |= fb;
}

// Workgroup exclusive segmented scan of (f, st) in thread order. Returns the
// exclusive value in (ef, est) and the workgroup total in (tf, tst).
template <typename V, bool MM, int W>
__device__ __forceinline__ void block_seg_scan(uint32_t f, const AggState<V, MM>& st, uint32_t& ef,
                                               AggState<V, MM>& est, uint32_t& tf,
                                               AggState<V, MM>& tst, AggState<V, MM>* s_wst,
                                               uint32_t* s_wf) {
  const uint32_t lane = lane_id();
  const int wid = threadIdx.x >> 6;
  uint32_t incf = f;
  AggState<V, MM> inc = st;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    AggState<V, MM> up = shfl_up_state(inc, off);
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
  AggState<V, MM> ex = shfl_up_state(inc, 1);
  uint32_t exf = __shfl_up(incf, 1, 64);
  if (lane == 0) {
    ex = AggState<V, MM>::identity();
    exf = 0u;
  }
  if (lane == 63) {
    s_wst[wid] = inc;
    s_wf[wid] = incf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t cf = 0u;
    AggState<V, MM> c = AggState<V, MM>::identity();
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const AggState<V, MM> t = s_wst[w];
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

// Per-wave output staging. A thread finds its groups in its 16 consecutive
// rows, so written in place one store instruction covers 64 lanes 16 outputs
// apart (measured ≈1 TB/s when nearly every row starts a group). The ids a
// wave writes are one contiguous range [lo, hi); it is written through a
// window of SEG_WIN entries in LDS, each window stored with consecutive lanes
// on consecutive ids.
constexpr int SEG_WIN = 256;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// out[id] = val_of(j) for each j in omask (id = id_of(j)), plus out[xid] =
// xval when has_x; every id in [lo, hi) is owned by exactly one lane of the
// wave. lo and hi are wave-uniform; the whole wave calls this.
template <typename E, typename IdOf, typename ValOf>
__device__ __forceinline__ void wave_stage_store(E* out, uint32_t lo, uint32_t hi, uint32_t omask,
                                                 IdOf id_of, ValOf val_of, bool has_x, uint32_t xid,
                                                 E xval, E* stage, uint32_t lane) {
  for (uint32_t w0 = lo; w0 < hi; w0 += SEG_WIN) {
#pragma unroll
    for (int j = 0; j < SEG_IPT; ++j) {
      if ((omask >> j) & 1u) {
        const uint32_t d = id_of(j) - w0;
        if (d < (uint32_t)SEG_WIN) stage[d] = val_of(j);
      }
    }
    if (has_x) {
      const uint32_t d = xid - w0;
      if (d < (uint32_t)SEG_WIN) stage[d] = xval;
    }
    wave_sync();
    for (uint32_t k = lane; k < (uint32_t)SEG_WIN && w0 + k < hi; k += 64) out[w0 + k] = stage[k];
    wave_sync();
  }
}

// KNOWN: the groups are given as sorted row ranges (seg_start) instead of
// being detected from the keys, so no keys are read and no lookback is needed.
//
// Each thread loads its own 16 consecutive rows straight into registers
// (16-byte vector loads when the inputs are aligned); the key before its
// first row comes from the previous lane (a shuffle) or, for a wave's first
// lane, from the previous wave's last key in LDS. With no tile staged in LDS
// the workgroup's LDS is a few KB and occupancy is set by registers (tiles
// staged through LDS, 68-100 KB per workgroup, ran 1-2 tiles per CU and were
// bound by each tile's load / barrier / lookback latency).
// Groups detected from the keys without min / max (STAGE) are written through
// the per-wave staging above, unless the caller knows they are few (DIRECT);
// known segments (at most 2048 groups) and min / max states are written in
// place. The staging's per-row sums cost registers: for 32-bit keys and float
// values 88 VGPRs against 70, five waves per SIMD against seven.
template <typename K, typename V, typename MT, bool MM, bool KNOWN, bool DIRECT>
__global__ void __launch_bounds__(SEG_BLOCK) k_segreduce_tiles(SegArgs<K, V, MT, MM> a) {
  constexpr bool STAGE = !KNOWN && !MM && !DIRECT;
  __shared__ uint64_t s_stage[STAGE ? SEG_W * SEG_WIN : 1];
  __shared__ uint32_t s_starts[KNOWN ? 2049 : 1];
  __shared__ K s_wlast[SEG_W];
  __shared__ uint32_t s_cnt[SEG_BLOCK];
  __shared__ uint32_t s_tmp[SEG_W + 1];
  __shared__ AggState<V, MM> s_wst[SEG_W + 1];
  __shared__ uint32_t s_wf[SEG_W + 1];
  __shared__ uint32_t s_tile, s_excl;

  const int tid = threadIdx.x;
  // Tiles that look back take their index from an atomic counter (every
  // earlier tile is then already running); known-segment tiles do not wait on
  // each other and use the workgroup index.
  if constexpr (!KNOWN) {
    if (tid == 0) s_tile = atomicAdd(a.tile_counter, 1u);
    __syncthreads();
  }
  const uint32_t tile = KNOWN ? blockIdx.x : s_tile;
  const uint64_t tbase = (uint64_t)tile * SEG_TILE;
  const uint32_t tn = (uint32_t)((a.n - tbase) < (uint64_t)SEG_TILE ? (a.n - tbase) : (uint64_t)SEG_TILE);

  // blocked: thread t owns rows t*IPT .. t*IPT+IPT-1 of the tile
  const uint32_t e0 = (uint32_t)tid * SEG_IPT;
  const uint32_t lane = lane_id();
  const int wid = tid >> 6;
  K kj[SEG_IPT];
  V vj[SEG_IPT];
  K prevk = K(0);
  if constexpr (!KNOWN) {
    if (tid == 0 && tbase > 0) prevk = a.keys[tbase - 1];
  }
  if (a.vec && e0 + SEG_IPT <= tn) {
    if constexpr (!KNOWN) wt_load_rows(a.keys + tbase + e0, kj);
    if (a.perm) {
      uint32_t pj[SEG_IPT];
      wt_load_rows(a.perm + tbase + e0, pj);
#pragma unroll
      for (int j = 0; j < SEG_IPT; ++j) vj[j] = a.vals[pj[j]];
    } else {
      wt_load_rows(a.vals + tbase + e0, vj);
    }
  } else {
#pragma unroll
    for (int j = 0; j < SEG_IPT; ++j) {
      const uint32_t e = e0 + j;
      kj[j] = K(0);
      vj[j] = V(0);
      if (e < tn) {
        if constexpr (!KNOWN) kj[j] = a.keys[tbase + e];
        vj[j] = a.perm ? a.vals[a.perm[tbase + e]] : a.vals[tbase + e];
      }
    }
  }
  if constexpr (KNOWN) {
    for (uint32_t i = tid; i < a.nseg; i += SEG_BLOCK) s_starts[i] = a.seg_start[i];
  } else {
    // the key before the thread's first row (every thread before a thread
    // with rows has all its rows, so the neighbour's last key is valid)
    const K up = __shfl_up(kj[SEG_IPT - 1], 1, 64);
    if (lane == 63) s_wlast[wid] = kj[SEG_IPT - 1];
    if (tid > 0) prevk = up;
  }
  __syncthreads();
  if constexpr (!KNOWN) {
    if (lane == 0 && wid > 0) prevk = s_wlast[wid - 1];
  }

  uint32_t hmask = 0u;
  uint32_t T, toff, texcl;
  if constexpr (KNOWN) {
    texcl = lds_lower_bound(s_starts, a.nseg, tbase);
    T = lds_lower_bound(s_starts, a.nseg, tbase + tn) - texcl;
    uint32_t nxt = lds_lower_bound(s_starts, a.nseg, tbase + e0);
    toff = nxt - texcl;
#pragma unroll
    for (int j = 0; j < SEG_IPT; ++j) {
      const uint32_t e = e0 + j;
      if (e < tn) {
        if (nxt < a.nseg && (uint64_t)s_starts[nxt] == tbase + e) {
          hmask |= 1u << j;
          ++nxt;
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < SEG_IPT; ++j) {
      const uint32_t e = e0 + j;
      if (e < tn) {
        const K pk = j ? kj[j - 1] : prevk;
        const bool h = (tbase + e == 0) || !key_equal(kj[j], pk, a.oc);
        hmask |= (uint32_t)h << j;
      }
    }
    s_cnt[tid] = (uint32_t)__popc(hmask);
    __syncthreads();
    T = block_excl_scan_lds<SEG_BLOCK, SEG_BLOCK>(s_cnt, s_tmp);
    toff = s_cnt[tid];
    if (tid == 0) lookback_publish(a.desc, tile, 1u, 0u, T);
    if (tid < 64) {
      const uint32_t ex = wave_lookback_u32(a.desc, tile, T, a.err);
      if (tid == 0) s_excl = ex;
    }
    __syncthreads();
    texcl = s_excl;
  }

  // sequential reduction of the thread's rows; with STAGE the sum of the
  // group closed at head j is kept in fs[j] until the staged write (its row
  // count is j minus the thread's previous head)
  AggState<V, MM> cur = AggState<V, MM>::identity();
  AggState<V, MM> lead = AggState<V, MM>::identity();
  uint32_t seen = 0u;
  uint32_t gid = texcl + toff - 1u;  // segment open at the thread's start
  V fs[STAGE ? SEG_IPT : 1] = {};
#pragma unroll
  for (int j = 0; j < SEG_IPT; ++j) {
    const uint32_t e = e0 + j;
    if (e < tn) {
      if ((hmask >> j) & 1u) {
        if (seen) {
          if constexpr (STAGE) {
            fs[j] = cur.sum();
          } else {
            finalize(a, gid, cur);
          }
        } else {
          lead = cur;
        }
        seen = 1u;
        ++gid;
        if constexpr (!KNOWN && !STAGE) {
          if (a.out_keys) a.out_keys[gid] = kj[j];
        }
        cur = AggState<V, MM>::identity();
      }
      cur.add(vj[j]);
    }
  }
  if (!seen) lead = cur;

  uint32_t ef, tf;
  AggState<V, MM> est, tst;
  block_seg_scan<V, MM, SEG_W>(seen, seen ? cur : lead, ef, est, tf, tst, s_wst, s_wf);
  AggState<V, MM> x = est;
  if (seen) {
    x.merge(lead);
    if (!ef) a.tile_lead[tile] = x;  // group began before this tile
    else if constexpr (!STAGE) finalize(a, texcl + toff - 1u, x);  // began inside it
  }
  if constexpr (STAGE) {
    // tile-local ids: the key of the thread's h-th head is toff + h; the group
    // closed at its h-th head (h >= 1) is toff + h - 1, and the group closed
    // by its first head, when it began inside the tile, is toff - 1
    const uint32_t nh = (uint32_t)__popc(hmask);
    const bool has_x = seen && ef;
    const uint32_t xid = toff - 1u;
    const uint64_t lanes = __ballot(seen != 0u);
    const uint32_t a0 = has_x ? toff - 1u : toff;
    const uint32_t a1 = toff + nh - 1u;
    uint32_t alo = 0u, ahi = 0u;
    if (lanes) {
      alo = __shfl(a0, __ffsll((unsigned long long)lanes) - 1, 64);
      ahi = __shfl(a1, 63 - __clzll((long long)lanes), 64);
    }
    const uint32_t omask = hmask & (hmask - 1u);  // heads after the thread's first
    auto before = [&](int j) { return (uint32_t)__popc(hmask & ((1u << j) - 1u)); };
    auto agg_id = [&](int j) { return toff + before(j) - 1u; };
    // rows of the group closed at head j (j in omask: a head lies below j)
    auto rows_j = [&](int j) { return (uint32_t)j - (31u - (uint32_t)__clz(hmask & ((1u << j) - 1u))); };
    if (a.out_keys) {
      const uint32_t klo = __shfl(toff, 0, 64);
      const uint32_t khi = __shfl(toff + nh, 63, 64);
      wave_stage_store<K>(a.out_keys + texcl, klo, khi, hmask, [&](int j) { return toff + before(j); },
                          [&](int j) { return kj[j]; }, false, 0u, K(0),
                          reinterpret_cast<K*>(s_stage + wid * SEG_WIN), lane);
    }
    if (a.out_sum)
      wave_stage_store<V>(a.out_sum + texcl, alo, ahi, omask, agg_id, [&](int j) { return fs[j]; },
                          has_x, xid, x.sum(), reinterpret_cast<V*>(s_stage + wid * SEG_WIN), lane);
    if (a.out_count)
      wave_stage_store<int64_t>(a.out_count + texcl, alo, ahi, omask, agg_id,
                                [&](int j) { return (int64_t)rows_j(j); }, has_x, xid, (int64_t)x.n,
                                reinterpret_cast<int64_t*>(s_stage + wid * SEG_WIN), lane);
    if (a.out_mean) {
      auto mean_j = [&](int j) { return mean_parts<V, MT>(fs[j], rows_j(j)); };
      wave_stage_store<MT>(a.out_mean + texcl, alo, ahi, omask, agg_id, mean_j, has_x, xid,
                           mean_of<V, MT, MM>(x), reinterpret_cast<MT*>(s_stage + wid * SEG_WIN), lane);
    }
  }
  if (tid == 0) {
    if (T == 0u) a.tile_lead[tile] = tst;
    else a.tile_trail[tile] = tst;
    a.tile_info[tile] = make_uint2(texcl, T);
  }
}

// One workgroup over the tiles in chunks of 1024: thread t takes tile
// c * 1024 + t of chunk c, so every load is coalesced, and the next chunk's
// loads are in flight during the current chunk's scan. A tile is one (flag,
// state) element: (1, its trail) when it holds heads, else (0, its lead). A
// workgroup segmented scan per chunk, carried from chunk to chunk, gives each
// tile the state of the group open at its start; a tile with heads finalises
// the group that ends at its first head, and the last group takes the total.
// (Threads that each walked a run of 12 tiles made uncoalesced, dependent
// loads: 75-88 us at 12,208 tiles.) Trail is left unwritten by a tile without
// heads and then not used; states are chosen with plain branches, since a
// conditional expression over two states selects an address and keeps them
// in scratch memory.
template <typename K, typename V, typename MT, bool MM>
__global__ void __launch_bounds__(1024) k_segreduce_fix(SegArgs<K, V, MT, MM> a) {
  __shared__ AggState<V, MM> s_wst[16 + 1];
  __shared__ uint32_t s_wf[16 + 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t nt = a.num_tiles;
  uint32_t cf = 0u;
  AggState<V, MM> carry = AggState<V, MM>::identity();
  uint2 info = make_uint2(0u, 0u);
  AggState<V, MM> lead = AggState<V, MM>::identity();
  AggState<V, MM> trail = AggState<V, MM>::identity();
  if (tid < nt) {
    info = a.tile_info[tid];
    lead = a.tile_lead[tid];
    trail = a.tile_trail[tid];
  }
  for (uint32_t c0 = 0; c0 < nt; c0 += 1024u) {
    const uint32_t u = c0 + tid;
    const uint32_t un = u + 1024u;
    uint2 ninfo = make_uint2(0u, 0u);
    AggState<V, MM> nlead = AggState<V, MM>::identity();
    AggState<V, MM> ntrail = AggState<V, MM>::identity();
    if (un < nt) {
      ninfo = a.tile_info[un];
      nlead = a.tile_lead[un];
      ntrail = a.tile_trail[un];
    }
    const uint32_t h = (u < nt && info.y > 0u) ? 1u : 0u;
    AggState<V, MM> el = lead;
    if (h) el = trail;
    uint32_t ef, tf;
    AggState<V, MM> est, tst;
    block_seg_scan<V, MM, 16>(h, el, ef, est, tf, tst, s_wst, s_wf);
    // state of the group open at the tile's start: the carry, then est
    uint32_t pf = cf;
    AggState<V, MM> pre = carry;
    seg_combine(pf, pre, ef, est);
    if (h && info.x > 0u) {
      pre.merge(lead);
      finalize(a, info.x - 1u, pre);
    }
    seg_combine(cf, carry, tf, tst);
    info = ninfo;
    lead = nlead;
    trail = ntrail;
  }
  if (tid == 0) {
    const uint2 last = a.tile_info[nt - 1];
    const uint32_t G = last.x + last.y;
    if (G > 0u) finalize(a, G - 1u, carry);
    *a.ngroups = G;
  }
}

// Above one chunk of 1024 tiles the fix-up runs its chunks in parallel:
// k_segfix_chunks gives each chunk's (flag, state) as one element (the
// workgroup segmented scan of k_segreduce_fix without a carry),
// k_segfix_carry scans those elements in one workgroup (at most 1024 chunks:
// 2^30 rows), leaving each chunk the state of the group open at its start
// and the total after the last chunk, and k_segfix_apply does
// k_segreduce_fix's per-chunk work from that carry. The order of the state
// merges is fixed, so results repeat bit for bit from run to run.
template <typename K, typename V, typename MT, bool MM>
__device__ __forceinline__ void segfix_tile(const SegArgs<K, V, MT, MM>& a, uint32_t u,
                                            uint2& info, AggState<V, MM>& lead, uint32_t& h,
                                            AggState<V, MM>& el) {
  info = make_uint2(0u, 0u);
  lead = AggState<V, MM>::identity();
  el = AggState<V, MM>::identity();
  h = 0u;
  if (u < a.num_tiles) {
    info = a.tile_info[u];
    lead = a.tile_lead[u];
    if (info.y > 0u) {
      h = 1u;
      el = a.tile_trail[u];
    } else {
      el = lead;
    }
  }
}

template <typename K, typename V, typename MT, bool MM>
__global__ void __launch_bounds__(1024)
k_segfix_chunks(SegArgs<K, V, MT, MM> a, uint32_t* cflag, AggState<V, MM>* cstate) {
  __shared__ AggState<V, MM> s_wst[16 + 1];
  __shared__ uint32_t s_wf[16 + 1];
  uint2 info;
  AggState<V, MM> lead, el;
  uint32_t h;
  segfix_tile(a, blockIdx.x * 1024u + threadIdx.x, info, lead, h, el);
  uint32_t ef, tf;
  AggState<V, MM> est, tst;
  block_seg_scan<V, MM, 16>(h, el, ef, est, tf, tst, s_wst, s_wf);
  if (threadIdx.x == 0) {
    cflag[blockIdx.x] = tf;
    cstate[blockIdx.x] = tst;
  }
}

template <typename V, bool MM>
__global__ void __launch_bounds__(1024)
k_segfix_carry(uint32_t nch, uint32_t* cflag, AggState<V, MM>* cstate) {
  __shared__ AggState<V, MM> s_wst[16 + 1];
  __shared__ uint32_t s_wf[16 + 1];
  const uint32_t t = threadIdx.x;
  uint32_t f = 0u;
  AggState<V, MM> st = AggState<V, MM>::identity();
  if (t < nch) {
    f = cflag[t];
    st = cstate[t];
  }
  uint32_t ef, tf;
  AggState<V, MM> est, tst;
  block_seg_scan<V, MM, 16>(f, st, ef, est, tf, tst, s_wst, s_wf);
  if (t < nch) {  // each thread rewrites only the entry it read
    cflag[t] = ef;
    cstate[t] = est;
  }
  if (t == 0) {
    cflag[nch] = tf;
    cstate[nch] = tst;
  }
}

template <typename K, typename V, typename MT, bool MM>
__global__ void __launch_bounds__(1024)
k_segfix_apply(SegArgs<K, V, MT, MM> a, const uint32_t* cflag, const AggState<V, MM>* cstate,
               uint32_t nch) {
  __shared__ AggState<V, MM> s_wst[16 + 1];
  __shared__ uint32_t s_wf[16 + 1];
  const uint32_t c = blockIdx.x;
  uint2 info;
  AggState<V, MM> lead, el;
  uint32_t h;
  segfix_tile(a, c * 1024u + threadIdx.x, info, lead, h, el);
  uint32_t ef, tf;
  AggState<V, MM> est, tst;
  block_seg_scan<V, MM, 16>(h, el, ef, est, tf, tst, s_wst, s_wf);
  // state of the group open at the tile's start: the chunk's carry, then est
  uint32_t pf = cflag[c];
  AggState<V, MM> pre = cstate[c];
  seg_combine(pf, pre, ef, est);
  if (h && info.x > 0u) {
    pre.merge(lead);
    finalize(a, info.x - 1u, pre);
  }
  if (c == nch - 1u && threadIdx.x == 0) {
    const uint2 last = a.tile_info[a.num_tiles - 1];
    const uint32_t G = last.x + last.y;
    if (G > 0u) finalize(a, G - 1u, cstate[nch]);
    *a.ngroups = G;
  }
}

// Known-segment groups of a single range-reduced counting-sort pass over
// integer keys: digit d = ordered(key) - min is one group exactly when its
// total is non-zero. Writes G, each group's first row and its key.
template <typename K>
__global__ void __launch_bounds__(1024)
k_seg_segments(const uint32_t* total, uint32_t R, uint64_t mn, int oc, uint64_t n,
               uint32_t* seg_start, K* out_keys, uint32_t* ngroups, uint32_t* dgid,
               uint32_t* skew, uint32_t lanes_per_table) {
  __shared__ uint32_t s_base[2048];
  __shared__ uint32_t s_gid[2048];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  __shared__ uint32_t s_max;
  const int tid = threadIdx.x;
  if (tid == 0) s_max = 0u;
  __syncthreads();
  uint32_t mx = 0u;
  for (int d = tid; d < 2048; d += 1024) {
    const uint32_t t = (uint32_t)d < R ? total[d] : 0u;
    s_base[d] = t;
    s_gid[d] = t ? 1u : 0u;
    mx = t > mx ? t : mx;
  }
  atomicMax(&s_max, mx);
  __syncthreads();
  // skew selector for the small-range kernels: 1 when the most frequent key
  // is expected in more than 2.7 of the lanes sharing one wave table per
  // slot (8/3 * 64 = 170.7: s_max * lanes * 3 > 8 n)
  if (skew && tid == 0) *skew = (uint64_t)s_max * lanes_per_table * 3u > n * 8u ? 1u : 0u;
  block_excl_scan_lds<2048, 1024>(s_base, s_tmp);
  const uint32_t G = block_excl_scan_lds<2048, 1024>(s_gid, s_tmp);
  for (int d = tid; d < (int)R; d += 1024) {
    const uint32_t t = (d + 1 < 2048 ? s_gid[d + 1] : G) - s_gid[d];
    if (dgid) dgid[d] = t ? s_gid[d] : 0xFFFFFFFFu;
    if (t) {
      seg_start[s_gid[d]] = s_base[d];
      if (out_keys) {
        K o = (K)(mn + (uint64_t)d);
        if (oc == ORD_SIGNED) o ^= (K)1 << (sizeof(K) * 8 - 1);
        out_keys[s_gid[d]] = o;
      }
    }
  }
  if (tid == 0) {
    seg_start[G] = (uint32_t)n;
    *ngroups = G;
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
  uint32_t* seg_start;
  uint32_t* dgid;   // [2048] group index of each digit (small-range path)
  void* part;       // [chunks][2048] chunk accumulators (small-range path)
  uint64_t* op_mm;  // [SORT_MAXCHUNKS * 2][2] one-pass: block min / max of the ordered keys
  uint32_t* op_ctl; // [4] one-pass: [0] stop word, [1] residues per table
  uint64_t part_bytes = 0;  // bytes at part
  // one-pass result word in device memory (allocation-free calls: read
  // with a stream-ordered copy); null: the calling thread's pinned slot
  uint32_t* op_res = nullptr;
  // chunk elements and carries of the parallel fix-up ([chunks + 1])
  uint32_t* fix_flag = nullptr;
  void* fix_state = nullptr;
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
  w.seg_start = c.take<uint32_t>(2049);
  w.dgid = c.take<uint32_t>(2048);
  w.part_bytes = (uint64_t)SORT_MAXCHUNKS * 2048 * st;
  w.part = c.take<char>(w.part_bytes);
  w.op_mm = c.take<uint64_t>((uint64_t)SORT_MAXCHUNKS * 2 * 2);
  w.op_ctl = c.take<uint32_t>(4);
  const uint64_t nch = (t + 1023) / 1024 + 1;
  w.fix_flag = c.take<uint32_t>(nch);
  w.fix_state = c.take<char>(nch * st);
  return w;
}

__global__ void k_seg_zero(uint32_t* desc, uint64_t ntiles, uint32_t* counter) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += stride) desc[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
}

struct SegCall {
  const void* keys;       // sorted keys (head mode) or null (known mode)
  int oc;
  const void* vals;
  const uint32_t* perm;
  uint64_t n;
  void* out_keys;
  void* out_sum;
  int64_t* out_count;
  void* out_min;
  void* out_max;
  void* out_mean;
  bool known;
  uint32_t nseg;          // known mode: G
  // at most n / 16 groups (the key span says so): outputs are written in
  // place, without the per-wave staging that costs registers
  bool few_groups = false;
};

template <typename K, typename V, typename MT, bool MM>
int segreduce_mm(const SegWs& w, const SegCall& c, hipStream_t s) {
  SegArgs<K, V, MT, MM> a{};
  a.keys = static_cast<const K*>(c.keys);
  a.oc = c.oc;
  a.vals = static_cast<const V*>(c.vals);
  a.perm = c.perm;
  a.n = c.n;
  a.out_keys = static_cast<K*>(c.out_keys);
  a.out_sum = static_cast<V*>(c.out_sum);
  a.out_count = c.out_count;
  a.out_min = static_cast<V*>(c.out_min);
  a.out_max = static_cast<V*>(c.out_max);
  a.out_mean = static_cast<MT*>(c.out_mean);
  a.tile_lead = static_cast<AggState<V, MM>*>(w.lead);
  a.tile_trail = static_cast<AggState<V, MM>*>(w.trail);
  a.tile_info = w.info;
  a.desc = w.desc;
  a.tile_counter = w.counter;
  a.err = device_error_word();
  a.ngroups = w.ngroups;
  a.num_tiles = (uint32_t)seg_tiles(c.n);
  a.seg_start = w.seg_start;
  a.nseg = c.nseg;
  a.vec = ((c.known || (uintptr_t)c.keys % 16 == 0) && (uintptr_t)c.vals % 16 == 0 &&
           (uintptr_t)c.perm % 16 == 0) ? 1 : 0;
  JFX_CHECK(a.err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  JFX_LAUNCH("seg_zero", s, (k_seg_zero<<<64, 256, 0, s>>>(w.desc, a.num_tiles, w.counter)));
  JFX_HIP(hipGetLastError());
  if (c.known) {
    JFX_LAUNCH("segreduce_tiles", s, (k_segreduce_tiles<K, V, MT, MM, true, false><<<a.num_tiles, SEG_BLOCK, 0, s>>>(a)));
  } else if (MM || !c.few_groups) {
    JFX_LAUNCH("segreduce_tiles", s, (k_segreduce_tiles<K, V, MT, MM, false, false><<<a.num_tiles, SEG_BLOCK, 0, s>>>(a)));
  } else if constexpr (!MM) {
    JFX_LAUNCH("segreduce_tiles", s, (k_segreduce_tiles<K, V, MT, MM, false, true><<<a.num_tiles, SEG_BLOCK, 0, s>>>(a)));
  }
  JFX_HIP(hipGetLastError());
  const uint32_t nch = (a.num_tiles + 1023u) / 1024u;
  if (nch <= 1u) {
    JFX_LAUNCH("segreduce_fix", s, (k_segreduce_fix<K, V, MT, MM><<<1, 1024, 0, s>>>(a)));
  } else {
    JFX_CHECK(nch <= 1024u, JFX_ERR_UNSUPPORTED, "segmented reduction: more than 2^30 rows");
    auto* fs = static_cast<AggState<V, MM>*>(w.fix_state);
    JFX_LAUNCH("segreduce_fix", s, (k_segfix_chunks<K, V, MT, MM><<<nch, 1024, 0, s>>>(a, w.fix_flag, fs)));
    JFX_LAUNCH("segreduce_fix", s, (k_segfix_carry<V, MM><<<1, 1024, 0, s>>>(nch, w.fix_flag, fs)));
    JFX_LAUNCH("segreduce_fix", s, (k_segfix_apply<K, V, MT, MM><<<nch, 1024, 0, s>>>(a, w.fix_flag, fs, nch)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K, typename V, typename MT>
int segreduce_typed(const SegWs& w, const SegCall& c, hipStream_t s) {
  if (c.out_min || c.out_max) return segreduce_mm<K, V, MT, true>(w, c, s);
  return segreduce_mm<K, V, MT, false>(w, c, s);
}

template <typename K>
int segreduce_key(const SegWs& w, int vt, const SegCall& c, hipStream_t s) {
  switch (vt) {
    case JFX_F32: return segreduce_typed<K, float, float>(w, c, s);
    case JFX_F64: return segreduce_typed<K, double, double>(w, c, s);
    case JFX_I32: return segreduce_typed<K, int32_t, float>(w, c, s);
    case JFX_I64: return segreduce_typed<K, int64_t, double>(w, c, s);
    case JFX_U32: return segreduce_typed<K, uint32_t, float>(w, c, s);
    case JFX_U64: return segreduce_typed<K, uint64_t, double>(w, c, s);
  }
  return set_error(JFX_ERR_INVALID, "bad value type");
}

int segreduce_dispatch(const SegWs& w, int kt, int vt, const SegCall& c, hipStream_t s) {
  if (key_bytes(kt) == 4) return segreduce_key<uint32_t>(w, vt, c, s);
  return segreduce_key<uint64_t>(w, vt, c, s);
}

int read_ngroups(const SegWs& w, uint64_t* out, hipStream_t s) {
  uint32_t g = 0;
  const int rc = read_back(s, w.ngroups, sizeof(uint32_t), &g);
  *out = g;
  return rc;
}

bool valid_vt(int vt) { return vt >= JFX_U32 && vt <= JFX_F64; }
bool int_key(int kt) { return order_code(kt) != ORD_FLOAT; }

template <typename K>
int segments_launch(const SortCtx& c, const SegWs& w, void* out_keys, hipStream_t s) {
  JFX_LAUNCH("seg_segments", s, (k_seg_segments<K><<<1, 1024, 0, s>>>(
                                     c.total, c.range_radix(), c.mn, c.oc, c.n, w.seg_start,
                                     static_cast<K*>(out_keys), w.ngroups, w.dgid,
                                     w.ngroups + 1,
                                     64u / wt_lane_groups(c.range_radix(),
                                                          c.range_radix() <= 1024 ? 1024u : 2048u))));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}


// Kernel for the small-range groupby. By default both kernels are launched
// and each reads the skew selector written by k_seg_segments, returning at
// once unless it is the chosen one: the wave-table kernel for ordinary key
// frequencies, the tile-sorted kernel when one key is frequent enough to make
// the lanes of a wave collide on it. JFX_SMALL_RANGE=wave or =tile forces one.
enum SmallRangePath { SR_AUTO = 0, SR_WAVE_TABLE = 1, SR_TILE_SORT = 2 };
inline SmallRangePath small_range_path() {
  const char* e = getenv("JFX_SMALL_RANGE");
  if (e && strcmp(e, "tile") == 0) return SR_TILE_SORT;
  if (e && strcmp(e, "wave") == 0) return SR_WAVE_TABLE;
  return SR_AUTO;
}

// Waves per block (1, 2 or 4: the sort chunk must split evenly) so that a
// block's tables take at most 64 KB when one table fits.
constexpr int wt_waves(int te, size_t esz) {
  return 65536 / (te * esz) >= 4 ? 4 : 65536 / (te * esz) >= 2 ? 2 : 1;
}

template <typename K, typename V, bool MM, int TE>
int wave_table_te(const SortCtx& c, const SegWs& w, const void* keys, const void* vals,
                  const uint32_t* sel, hipStream_t s, uint32_t& nparts) {
  // sums of float32 or integer values: LDS-atomic tables (one word per digit)
  constexpr bool ATOMIC = !MM && !std::is_same<V, double>::value;
  constexpr size_t ESZ = ATOMIC ? sizeof(typename WaAcc<V>::T) : sizeof(WAcc<V, MM>) + sizeof(uint32_t);
  constexpr int W = wt_waves(TE, ESZ);
  constexpr uint64_t STEP = 64u * WT_RPL;
  WaveTabArgs<K, V, MM> a{};
  a.keys = static_cast<const K*>(keys);
  a.vals = static_cast<const V*>(vals);
  // raw bits of the minimum key (the range fast path only takes integer keys)
  a.sub = c.oc == ORD_SIGNED ? (K)((K)c.mn ^ ((K)1 << (sizeof(K) * 8 - 1))) : (K)c.mn;
  a.R = c.range_radix();
  a.lg = wt_lane_groups(a.R, TE);
  a.rt = (uint32_t)TE / a.lg;
  a.n = c.n;
  // As many blocks as fit in LDS at once (160 KB per CU, 256 CUs), each wave
  // a whole number of steps, within the partial buffer (SORT_MAXCHUNKS x 2048).
  const uint64_t per_cu = std::max<uint64_t>(1, (160u * 1024u) / ((uint64_t)W * TE * ESZ));
  const uint64_t cap = std::min<uint64_t>(256 * per_cu, (uint64_t)SORT_MAXCHUNKS * 2048 / a.R);
  const uint64_t want = (c.n + (uint64_t)W * STEP - 1) / ((uint64_t)W * STEP);
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, want));
  a.rows_per_wave = ((c.n + (uint64_t)grid * W - 1) / ((uint64_t)grid * W) + STEP - 1) / STEP * STEP;
  grid = (uint32_t)((c.n + a.rows_per_wave * W - 1) / (a.rows_per_wave * W));
  a.sel = sel;
  a.vec = ((uintptr_t)keys % 16 == 0 && (uintptr_t)vals % 16 == 0) ? 1 : 0;
  a.part = static_cast<AggState<V, MM>*>(w.part);
  a.err = c.err;
  const char* nm = TE == 1024 ? "wave_table_1k" : "wave_table_2k";
  if constexpr (ATOMIC) {
    if (a.vec) {
      JFX_LAUNCH(nm, s, (k_wave_atomic<K, V, TE, W, true><<<grid, W * 64, 0, s>>>(a)));
    } else {
      JFX_LAUNCH(nm, s, (k_wave_atomic<K, V, TE, W, false><<<grid, W * 64, 0, s>>>(a)));
    }
  } else if (a.vec) {
    JFX_LAUNCH(nm, s, (k_wave_table<K, V, MM, TE, W, true><<<grid, W * 64, 0, s>>>(a)));
  } else {
    JFX_LAUNCH(nm, s, (k_wave_table<K, V, MM, TE, W, false><<<grid, W * 64, 0, s>>>(a)));
  }
  JFX_HIP(hipGetLastError());
  nparts = grid;
  return JFX_OK;
}

// Sets nparts to the number of partials the kernel writes (its grid size).
template <typename K, typename V, bool MM>
int wave_table_launch(const SortCtx& c, const SegWs& w, const void* keys, const void* vals,
                      const uint32_t* sel, hipStream_t s, uint32_t& nparts) {
  if (c.range_radix() <= 1024) return wave_table_te<K, V, MM, 1024>(c, w, keys, vals, sel, s, nparts);
  return wave_table_te<K, V, MM, 2048>(c, w, keys, vals, sel, s, nparts);
}

template <typename K, typename V, typename MT, bool MM>
int tile_reduce_mm(const SortCtx& c, const SegWs& w, const void* keys, const void* vals,
                   const SegCall& sc, hipStream_t s) {
  TileRedArgs<K, V, MM> a{};
  a.keys = static_cast<const K*>(keys);
  a.oc = c.oc;
  a.sub = (K)c.mn;
  a.R = c.range_radix();
  a.vals = static_cast<const V*>(vals);
  a.n = c.n;
  a.chunk = c.chunk;
  a.part = static_cast<AggState<V, MM>*>(w.part);
  SmallRangePath path = small_range_path();
  // no key can make the lanes of a wave collide often enough to pick the
  // tile-sorted kernel when even the bound on the largest key count stays
  // under the selector's threshold: launch the wave-table kernel alone
  const uint32_t te = c.range_radix() <= 1024 ? 1024u : 2048u;
  const uint64_t lanes = 64u / wt_lane_groups(c.range_radix(), te);
  if (path == SR_AUTO && c.maxbin_sum != ~0ull && c.maxbin_sum * lanes * 3u <= c.n * 8u)
    path = SR_WAVE_TABLE;
  const uint32_t* sel = path == SR_AUTO ? w.ngroups + 1 : nullptr;
  uint32_t wave_parts = 0;
  if (path != SR_TILE_SORT) {
    const int rc = wave_table_launch<K, V, MM>(c, w, keys, vals, sel, s, wave_parts);
    if (rc != JFX_OK) return rc;
  }
  a.sel = sel;
  if (path == SR_WAVE_TABLE) {
    // the wave-table kernel alone
  } else if (a.R <= 256) {
    JFX_LAUNCH("tile_reduce_d8", s, (k_tile_reduce<K, V, MM, 8, 8, 8><<<c.nchunks, 512, 0, s>>>(a)));
  } else if (a.R <= 1024) {
    JFX_LAUNCH("tile_reduce_d10", s, (k_tile_reduce<K, V, MM, 10, 8, 8><<<c.nchunks, 512, 0, s>>>(a)));
  } else {
    JFX_LAUNCH("tile_reduce_d11", s, (k_tile_reduce<K, V, MM, 11, 4, 8><<<c.nchunks, 256, 0, s>>>(a)));
  }
  JFX_HIP(hipGetLastError());
  GroupOut<V, MT> o{static_cast<V*>(sc.out_sum), sc.out_count, static_cast<V*>(sc.out_min),
                    static_cast<V*>(sc.out_max), static_cast<MT*>(sc.out_mean)};
  JFX_LAUNCH("tile_combine", s, (k_tile_combine<V, MT, MM><<<(a.R + 3) / 4, 256, 0, s>>>(
                                     a.part, path == SR_TILE_SORT ? c.nchunks : wave_parts,
                                     a.R, w.dgid, o, c.total, sel, c.nchunks)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K, typename V, typename MT>
int tile_reduce_typed(const SortCtx& c, const SegWs& w, const void* keys, const void* vals,
                      const SegCall& sc, hipStream_t s) {
  if (sc.out_min || sc.out_max) return tile_reduce_mm<K, V, MT, true>(c, w, keys, vals, sc, s);
  return tile_reduce_mm<K, V, MT, false>(c, w, keys, vals, sc, s);
}

template <typename K>
int tile_reduce_key(const SortCtx& c, const SegWs& w, int vt, const void* keys, const void* vals,
                    const SegCall& sc, hipStream_t s) {
  switch (vt) {
    case JFX_F32: return tile_reduce_typed<K, float, float>(c, w, keys, vals, sc, s);
    case JFX_F64: return tile_reduce_typed<K, double, double>(c, w, keys, vals, sc, s);
    case JFX_I32: return tile_reduce_typed<K, int32_t, float>(c, w, keys, vals, sc, s);
    case JFX_I64: return tile_reduce_typed<K, int64_t, double>(c, w, keys, vals, sc, s);
    case JFX_U32: return tile_reduce_typed<K, uint32_t, float>(c, w, keys, vals, sc, s);
    case JFX_U64: return tile_reduce_typed<K, uint64_t, double>(c, w, keys, vals, sc, s);
  }
  return set_error(JFX_ERR_INVALID, "bad value type");
}

// ---- one-pass path ------------------------------------------------------------
// Integer keys, float32 / float64 / integer values (sum / count / mean / min /
// max), at least 2^16 rows: a miss costs 10-14 us (the aborted table launch,
// two small kernels and the result exchange) before the statistics-first
// path, against 22-25 us saved by a hit (measured at 64K to 2M rows).
// JFX_ONEPASS=0 turns the path off, =always tries it at any size.
// res receives the final kernel's result (status, value).

std::atomic<uint32_t> g_op_epoch{0};
constexpr uint64_t OP_MIN_ROWS = 1ull << 16;

bool onepass_wanted(int kt, int vt, uint64_t n) {
  (void)vt;
  if (!int_key(kt)) return false;
  const char* e = getenv("JFX_ONEPASS");
  if (e && strcmp(e, "0") == 0) return false;
  if (e && strcmp(e, "always") == 0) return true;
  return n >= OP_MIN_ROWS;
}

template <typename K, typename V, typename MT, bool MM, uint32_t TE>
int onepass_run(int oc, const void* keys, const void* vals, uint64_t n, const SegWs& w,
                void* out_keys, const SegCall& sc, uint64_t capacity, uint32_t rt_min,
                uint32_t* res, hipStream_t s) {
  constexpr uint64_t STEP = 64u * OP_RPL;
  // LDS per block: a sum table per wave, one count table and (MM) one min and
  // one max table
  constexpr uint64_t BLK = (uint64_t)OP_W * TE * sizeof(typename OpT<V>::T) * (op_dd<V> ? 2u : 1u) +
                           TE * 4u + (MM ? 2u * TE * sizeof(typename ValTraits<V>::U) : 0u);
  OnePassArgs<K, V, MM> a{};
  a.keys = static_cast<const K*>(keys);
  a.vals = static_cast<const V*>(vals);
  a.n = n;
  a.oc = oc;
  a.epoch = ++g_op_epoch;
  if (a.epoch == 0u) a.epoch = ++g_op_epoch;  // 0 is the value of fresh scratch memory
  a.rt_min = rt_min;
  a.part = static_cast<OpPart<V, MM>*>(w.part);
  a.mm = w.op_mm;
  a.ctl = w.op_ctl;
  // as many blocks as fit in LDS at once, at most OP_GROUP x OP_L1 (the
  // final merge's reach)
  const uint64_t per_cu = std::max<uint64_t>(1, (160u * 1024u) / BLK);
  const uint64_t cap = std::min<uint64_t>(256 * per_cu, (uint64_t)OP_GROUP * OP_L1);
  const uint64_t want = (n + (uint64_t)OP_W * STEP - 1) / ((uint64_t)OP_W * STEP);
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, want));
  uint32_t group = OP_GROUP;
  const bool vec = (uintptr_t)keys % 16 == 0 && (uintptr_t)vals % 16 == 0;
  const char* nm = TE == OP_TE ? "onepass_table" : "onepass_table_2k";
  // float32 sums alone: the count-free tables (group presence from the sums)
  constexpr bool NC_OK = std::is_same<V, float>::value && !MM;
  const bool nc = NC_OK && sc.out_count == nullptr && sc.out_mean == nullptr;
  if (nc) {
    if constexpr (NC_OK) {
      // 32-bit keys, TE = 1024: 16 waves per block (128 KB of wave tables,
      // one block per CU): the same 4096 waves as 1024 blocks of 4, each
      // wave reading the same steps, but a quarter of the block tables to
      // write and merge (3.1 MB instead of 12.6 MB). The 2048-residue tables
      // (256 KB at 16 waves) and 64-bit keys (over the 128 VGPRs that 16
      // waves per CU leave) keep 4 waves.
      constexpr int WN = (TE == OP_TE && sizeof(K) == 4) ? 16 : OP_W;
      if constexpr (WN != OP_W) {
        const uint64_t want16 = (n + (uint64_t)WN * STEP - 1) / ((uint64_t)WN * STEP);
        grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(256, want16));
        group = (grid + OP_L1 - 1) / OP_L1;
      }
      if (vec) {
        JFX_LAUNCH(nm, s, (k_onepass_table<K, V, MM, true, TE, true, WN><<<grid, WN * 64, 0, s>>>(a)));
      } else {
        JFX_LAUNCH(nm, s, (k_onepass_table<K, V, MM, false, TE, true, WN><<<grid, WN * 64, 0, s>>>(a)));
      }
    }
  } else if (vec) {
    JFX_LAUNCH(nm, s, (k_onepass_table<K, V, MM, true, TE><<<grid, OP_W * 64, 0, s>>>(a)));
  } else {
    JFX_LAUNCH(nm, s, (k_onepass_table<K, V, MM, false, TE><<<grid, OP_W * 64, 0, s>>>(a)));
  }
  JFX_HIP(hipGetLastError());
  GroupOut<V, MT> o{static_cast<V*>(sc.out_sum), sc.out_count, static_cast<V*>(sc.out_min),
                    static_cast<V*>(sc.out_max), static_cast<MT*>(sc.out_mean)};
  // level-1 merges go to the part buffer after the block tables (it holds
  // SORT_MAXCHUNKS x 2048 min / max states, more than both need)
  const uint32_t ng1 = (grid + group - 1) / group;
  JFX_CHECK(ng1 <= OP_L1, JFX_ERR_UNSUPPORTED, "one-pass grid exceeds the final merge");
  JFX_CHECK(((uint64_t)grid + ng1) * TE * sizeof(OpPart<V, MM>) <= w.part_bytes,
            JFX_ERR_WORKSPACE, "one-pass partials exceed the partial buffer");
  OpPart<V, MM>* l1 = a.part + (size_t)grid * TE;
  static_assert(TE % OP_MT == 0, "level-1 merge blocks cover whole tables");
  JFX_LAUNCH("onepass_merge", s, (k_onepass_merge<V, MM, TE><<<dim3(ng1, TE / OP_MT), OP_MT, 0, s>>>(
                                      a.part, grid, group, a.ctl, a.epoch, l1)));
  JFX_HIP(hipGetLastError());
  // the final kernel publishes its result word to the host itself (no
  // separate read-back kernel), or to a device word that is copied back
  ReadBackSlot rb{};
  if (w.op_res) {
    rb.dst = w.op_res;
    rb.seq_slot = reinterpret_cast<uint64_t*>(w.op_res + 2);
    rb.seq = a.epoch;
  } else {
    const int rc = read_back_slot(rb);
    if (rc != JFX_OK) return rc;
  }
  JFX_LAUNCH("onepass_final", s, (k_onepass_final<K, V, MT, MM, TE><<<1, 1024, 0, s>>>(
                                      l1, ng1, a.mm, grid, a.ctl, a.epoch, oc,
                                      static_cast<K*>(out_keys), o, capacity, rb)));
  JFX_HIP(hipGetLastError());
  if (w.op_res) {
    uint64_t word = 0;
    JFX_HIP(hipMemcpyAsync(&word, rb.seq_slot, sizeof(word), hipMemcpyDeviceToHost, s));
    JFX_HIP(hipStreamSynchronize(s));
    JFX_CHECK(rb_decode(word, rb.seq, res), JFX_ERR_HIP, "one-pass result word missing");
    return JFX_OK;
  }
  return read_back_wait_word(s, rb, res);
}

// te = 1024 or 2048 residues; rt_min = te turns the lane-group tables off;
// min / max tables when a min or max output is asked for.
template <typename K, typename V, typename MT>
int onepass_typed(int oc, const void* keys, const void* vals, uint64_t n, const SegWs& w,
                  void* out_keys, const SegCall& sc, uint64_t capacity, uint32_t te,
                  uint32_t rt_min, uint32_t* res, hipStream_t s) {
  const bool mm = sc.out_min || sc.out_max;
  if (te == 2 * OP_TE) {
    if (mm) {
      // float64 sums with compensation and 64-bit min / max codes at 2048
      // residues exceed the 160 KB of LDS: no retry (res keeps the miss)
      if constexpr (op_dd<V>) return JFX_OK;
      else return onepass_run<K, V, MT, true, 2 * OP_TE>(oc, keys, vals, n, w, out_keys, sc, capacity, rt_min, res, s);
    }
    return onepass_run<K, V, MT, false, 2 * OP_TE>(oc, keys, vals, n, w, out_keys, sc, capacity, rt_min, res, s);
  }
  if (mm) return onepass_run<K, V, MT, true, OP_TE>(oc, keys, vals, n, w, out_keys, sc, capacity, rt_min, res, s);
  return onepass_run<K, V, MT, false, OP_TE>(oc, keys, vals, n, w, out_keys, sc, capacity, rt_min, res, s);
}

template <typename K>
int onepass_key(int oc, int vt, const void* keys, const void* vals, uint64_t n, const SegWs& w,
                void* out_keys, const SegCall& sc, uint64_t capacity, uint32_t te,
                uint32_t rt_min, uint32_t* res, hipStream_t s) {
  switch (vt) {
    case JFX_F32: return onepass_typed<K, float, float>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
    case JFX_F64: return onepass_typed<K, double, double>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
    case JFX_I32: return onepass_typed<K, int32_t, float>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
    case JFX_I64: return onepass_typed<K, int64_t, double>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
    case JFX_U32: return onepass_typed<K, uint32_t, float>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
    case JFX_U64: return onepass_typed<K, uint64_t, double>(oc, keys, vals, n, w, out_keys, sc, capacity, te, rt_min, res, s);
  }
  return set_error(JFX_ERR_INVALID, "bad value type for the one-pass groupby");
}

// The one-pass try of jfx_groupby_ex: tables of 1024 residues, and on a miss
// whose span seen so far still fits, one retry (1024 residues without lane
// groups, or 2048). res[0] = OP_HIT / OP_MISS / OP_CAPACITY, res[1] = the
// group count, the capacity needed or a lower bound of the key span.
int onepass_try(int key_type, int val_type, const void* keys, const void* vals, uint64_t n,
                const SegWs& w, void* out_keys, void* out_sum, int64_t* out_count,
                void* out_min, void* out_max, void* out_mean, uint64_t capacity, uint32_t* res,
                hipStream_t s) {
  const SegCall oc{nullptr, order_code(key_type), vals, nullptr, n, out_keys, out_sum,
                   out_count, out_min, out_max, out_mean, true, 0u};
  const bool k4 = key_bytes(key_type) == 4;
  auto attempt = [&](uint32_t te, uint32_t rt_min) -> int {
    return k4 ? onepass_key<uint32_t>(oc.oc, val_type, keys, vals, n, w, out_keys, oc, capacity,
                                      te, rt_min, res, s)
              : onepass_key<uint64_t>(oc.oc, val_type, keys, vals, n, w, out_keys, oc, capacity,
                                      te, rt_min, res, s);
  };
  res[0] = OP_MISS;
  res[1] = 0u;
  int rc = attempt(OP_TE, 16u);
  if (rc != JFX_OK) return rc;
  if (res[0] == OP_MISS && res[1] < 2 * OP_TE) {
    const uint32_t te = res[1] < OP_TE ? OP_TE : 2 * OP_TE;
    rc = attempt(te, te);
  }
  return rc;
}

}  // namespace

}  // namespace jfx

// ---- allocation-free one-pass groupby ------------------------------------------
// Scratch of jfx_groupby_noalloc, one copy per device, held until the process
// ends: the one-pass block and level-1 tables (at most 780 tables of 1024
// residues x 32 bytes, 25.6 MB, for 64-bit integer values with min / max),
// the block min / max, the control words and the result word. Host code uses
// the address from hipGetSymbolAddress, never the array's name.
constexpr uint64_t JFX_NA_PART_BYTES = 26ull << 20;
constexpr uint64_t JFX_NA_SCRATCH_BYTES = 27ull << 20;
static __device__ __align__(256) unsigned char scratch[JFX_NA_SCRATCH_BYTES];

namespace jfx {
namespace {

// Device address of this device's scratch, resolved once per device with
// that device current.
int noalloc_scratch(void** out) {
  static std::mutex mu;
  static std::vector<void*> addr;
  int dev = 0;
  JFX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    JFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    JFX_CHECK(p != nullptr, JFX_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    addr[dev] = p;
  }
  *out = addr[dev];
  return JFX_OK;
}

}  // namespace
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
  jfx::Carver cv(workspace, ws_bytes);
  jfx::SegWs w = jfx::carve_seg(cv, jfx::key_bytes(val_type), n);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "segreduce workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  jfx::SegCall c{sorted_keys, jfx::order_code(key_type), vals, perm, n, out_keys, out_sum,
                 out_count, out_min, out_max, out_mean, false, 0u};
  int rc = jfx::segreduce_dispatch(w, key_type, val_type, c, s);
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
  return jfx_groupby_ex(key_type, keys, val_type, vals, n, out_keys, out_sum, out_count, out_min,
                        out_max, out_mean, out_ngroups, n, digit_bits, workspace, ws_bytes, stream);
}

int jfx_groupby_ex(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                   void* out_keys, void* out_sum, int64_t* out_count, void* out_min,
                   void* out_max, void* out_mean, uint64_t* out_ngroups, uint64_t out_capacity,
                   int digit_bits, void* workspace, uint64_t ws_bytes, void* stream) {
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
  JFX_CHECK(digit_bits == 0 || digit_bits == 8 || digit_bits == 11, JFX_ERR_INVALID,
            "digit_bits must be 0, 8 or 11");
  const int kb = jfx::key_bytes(key_type), vb = jfx::key_bytes(val_type);
  jfx::Carver cv(workspace, ws_bytes);
  void* skeys = cv.take<char>(n * kb);
  void* svals = cv.take<char>(n * vb);
  jfx::SegWs w = jfx::carve_seg(cv, vb, n);
  const uint64_t sws = jfx::sort_ws_bytes(key_type, vb, n);
  void* sort_ws = cv.take<char>(sws);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "groupby workspace too small");
  hipStream_t s = jfx::as_stream(stream);
  // a key span known to be at least 2048: the sort-based path will run
  bool wide = false;
  if (jfx::onepass_wanted(key_type, val_type, n)) {
    // one read of keys and values when the key range turns out below 1024
    // (2048 on the retry); on a miss the statistics-first path below runs
    uint32_t res[2] = {0u, 0u};
    const int rc = jfx::onepass_try(key_type, val_type, keys, vals, n, w, out_keys, out_sum,
                                    out_count, out_min, out_max, out_mean, out_capacity, res, s);
    if (rc != JFX_OK) return rc;
    if (res[0] == jfx::OP_HIT) {
      *out_ngroups = res[1];
      return JFX_OK;
    }
    if (res[0] == jfx::OP_CAPACITY) {
      *out_ngroups = res[1];
      return jfx::set_error(JFX_ERR_CAPACITY, "outputs hold fewer entries than the groups this "
                                              "call produces; *out_ngroups holds the capacity needed");
    }
    wide = res[1] >= 2048u;
    if (wide && out_capacity < n) {
      *out_ngroups = n;
      return jfx::set_error(JFX_ERR_CAPACITY, "outputs hold fewer entries than the groups this "
                                              "call may produce; *out_ngroups holds the capacity needed");
    }
  }
  jfx::SortCtx ctx;
  // float keys never take the small-range path, nor keys whose span is known
  // to be at least 2048: their statistics pass also makes the layout
  // histograms of the multi-pass sort
  int rc = jfx::sort_prepare(ctx, key_type, keys, n, vb, digit_bits, sort_ws, sws, s,
                             !jfx::int_key(key_type) || wide);
  if (rc != JFX_OK) return rc;
  const int oc = jfx::order_code(key_type);
  const bool small = jfx::int_key(key_type) && ctx.single_pass();
  // groups are at most the rows, and at most the key range on the small path
  const uint64_t need = small ? std::min<uint64_t>(ctx.range_radix(), n) : n;
  if (need > out_capacity) {
    *out_ngroups = need;
    return jfx::set_error(JFX_ERR_CAPACITY, "outputs hold fewer entries than the groups this "
                                            "call may produce; *out_ngroups holds the capacity needed");
  }
  if (small) {
    // Small key range: every distinct key is one digit of a single counting
    // sort; each tile is sorted and reduced in LDS (tile_reduce), so neither
    // sorted keys nor sorted values are written to HBM.
    rc = jfx::sort_digit_totals(ctx);
    if (rc != JFX_OK) return rc;
    rc = kb == 4 ? jfx::segments_launch<uint32_t>(ctx, w, out_keys, s)
                 : jfx::segments_launch<uint64_t>(ctx, w, out_keys, s);
    if (rc != JFX_OK) return rc;
    jfx::SegCall c{nullptr, oc, vals, nullptr, n, nullptr, out_sum, out_count, out_min, out_max,
                   out_mean, true, 0u};
    rc = kb == 4 ? jfx::tile_reduce_key<uint32_t>(ctx, w, val_type, keys, vals, c, s)
                 : jfx::tile_reduce_key<uint64_t>(ctx, w, val_type, keys, vals, c, s);
    if (rc != JFX_OK) return rc;
    return jfx::read_ngroups(w, out_ngroups, s);
  }
  rc = jfx::sort_execute(ctx, keys, skeys, 1, vals, svals, 0, 0);
  if (rc != JFX_OK) return rc;
  jfx::SegCall c{skeys, oc, svals, nullptr, n, out_keys, out_sum, out_count, out_min, out_max,
                 out_mean, false, 0u};
  // distinct keys have distinct ordered codes, so G <= max - min + 1
  c.few_groups = ctx.mx - ctx.mn < n / 16;
  rc = jfx::segreduce_dispatch(w, key_type, val_type, c, s);
  if (rc != JFX_OK) return rc;
  return jfx::read_ngroups(w, out_ngroups, s);
}

int jfx_groupby_noalloc(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                        void* out_keys, void* out_sum, int64_t* out_count, void* out_min,
                        void* out_max, void* out_mean, uint64_t* out_ngroups,
                        uint64_t out_capacity, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(jfx::int_key(key_type), JFX_ERR_UNSUPPORTED,
            "the allocation-free groupby takes integer keys only");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  JFX_CHECK(n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be below 2^30 rows");
  JFX_CHECK(out_ngroups != nullptr, JFX_ERR_INVALID, "out_ngroups is null");
  if (n == 0) {
    *out_ngroups = 0;
    return JFX_OK;
  }
  JFX_CHECK(keys != nullptr && vals != nullptr, JFX_ERR_INVALID, "null input");
  void* base = nullptr;
  int rc = jfx::noalloc_scratch(&base);
  if (rc != JFX_OK) return rc;
  jfx::Carver cv(base, JFX_NA_SCRATCH_BYTES);
  jfx::SegWs w{};
  w.part_bytes = JFX_NA_PART_BYTES;
  w.part = cv.take<char>(JFX_NA_PART_BYTES);
  w.op_mm = cv.take<uint64_t>((uint64_t)jfx::SORT_MAXCHUNKS * 2 * 2);
  w.op_ctl = cv.take<uint32_t>(4);
  w.op_res = cv.take<uint32_t>(4);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "allocation-free groupby scratch layout too large");
  hipStream_t s = jfx::as_stream(stream);
  uint32_t res[2] = {0u, 0u};
  rc = jfx::onepass_try(key_type, val_type, keys, vals, n, w, out_keys, out_sum, out_count,
                        out_min, out_max, out_mean, out_capacity, res, s);
  if (rc != JFX_OK) return rc;
  *out_ngroups = res[1];
  if (res[0] == jfx::OP_HIT) return JFX_OK;
  if (res[0] == jfx::OP_CAPACITY)
    return jfx::set_error(JFX_ERR_CAPACITY, "outputs hold fewer entries than the groups this "
                                            "call produces; *out_ngroups holds the capacity needed");
  *out_ngroups = 0;
  return jfx::set_error(JFX_ERR_UNSUPPORTED,
                        "key span (max - min) is 2048 or more (1024 or more for float64 values "
                        "with min / max): beyond the allocation-free groupby; use jfx_groupby_ex "
                        "with a workspace");
}

// ---- batched groupby ---------------------------------------------------------

int jfx_batched_groupby_workspace_size(uint32_t total_bits, int val_type, uint64_t n,
                                       uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  JFX_CHECK(total_bits >= 1 && total_bits <= 64, JFX_ERR_INVALID, "total_bits must be 1..64");
  const int cb = total_bits <= 32 ? 4 : 8, vb = jfx::key_bytes(val_type);
  jfx::Carver c(nullptr, 0);
  c.take<char>(n * cb);
  c.take<char>(n * cb);
  c.take<char>(n * vb);
  jfx::carve_seg(c, vb, n);
  c.take<char>(jfx::sort_ws_bytes(cb == 4 ? JFX_U32 : JFX_U64, vb, n, true));
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_batched_groupby(int key_type, const void* keys, int val_type, const void* vals, uint64_t n,
                        const int64_t* frame_offsets, uint32_t nframes, uint64_t key_min,
                        uint32_t key_bits, int digit_bits, int sort_shape, void* out_keys,
                        void* out_sum, int64_t* out_count, void* out_min, void* out_max,
                        void* out_mean, uint32_t* out_ngroups, void* workspace, uint64_t ws_bytes,
                        void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64 && jfx::int_key(key_type), JFX_ERR_INVALID,
            "batched keys must be integers");
  JFX_CHECK(jfx::valid_vt(val_type), JFX_ERR_INVALID, "bad value type");
  JFX_CHECK(n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be below 2^30 rows");
  JFX_CHECK(nframes >= 1, JFX_ERR_INVALID, "at least one frame");
  uint32_t fbits = 0;
  while ((1ull << fbits) < nframes) ++fbits;
  const uint32_t total = fbits + key_bits;
  JFX_CHECK(key_bits >= 1 && total <= 64, JFX_ERR_UNSUPPORTED,
            "frame and key bits exceed 64 (narrow the key range)");
  JFX_CHECK(out_ngroups != nullptr, JFX_ERR_INVALID, "out_ngroups is null");
  JFX_CHECK(digit_bits == 0 || digit_bits == 8 || digit_bits == 11, JFX_ERR_INVALID,
            "digit_bits must be 0, 8 or 11");
  JFX_CHECK(sort_shape >= 0 && sort_shape < jfx::SORT_SHAPES, JFX_ERR_INVALID,
            "sort_shape must be 0 to 3");
  hipStream_t s = jfx::as_stream(stream);
  if (n == 0) {
    JFX_LAUNCH("copy_ngroups", s, (jfx::k_zero_u32<<<1, 64, 0, s>>>(out_ngroups)));
    JFX_HIP(hipGetLastError());
    return JFX_OK;
  }
  JFX_CHECK(keys && vals && frame_offsets && workspace, JFX_ERR_INVALID, "null input or workspace");
  const int cb = total <= 32 ? 4 : 8, vb = jfx::key_bytes(val_type);
  const int ckt = cb == 4 ? JFX_U32 : JFX_U64;
  jfx::Carver cv(workspace, ws_bytes);
  void* comp = cv.take<char>(n * cb);
  void* scomp = cv.take<char>(n * cb);
  void* svals = cv.take<char>(n * vb);
  jfx::SegWs w = jfx::carve_seg(cv, vb, n);
  const uint64_t sws = jfx::sort_ws_bytes(ckt, vb, n, true);
  void* sort_ws = cv.take<char>(sws);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "batched groupby workspace too small");
  uint32_t* err = jfx::device_error_word();
  JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  const int oc = jfx::order_code(key_type);
  // ordered image of the minimum (integer orders: flip the sign bit of signed keys)
  uint64_t omin = jfx::key_bytes(key_type) == 4 ? (key_min & 0xFFFFFFFFull) : key_min;
  if (oc == jfx::ORD_SIGNED) omin ^= jfx::key_bytes(key_type) == 4 ? 0x80000000ull : 0x8000000000000000ull;
  const uint32_t blocks = (uint32_t)((n + jfx::BC_ROWS - 1) / jfx::BC_ROWS);
  if (jfx::key_bytes(key_type) == 4) {
    if (cb == 4) {
      JFX_LAUNCH("batch_compose", s, (jfx::k_batch_compose<uint32_t, uint32_t><<<blocks, jfx::BC_BLOCK, 0, s>>>(
                                         (const uint32_t*)keys, oc, omin, key_bits, frame_offsets,
                                         nframes, n, (uint32_t*)comp, err)));
    } else {
      JFX_LAUNCH("batch_compose", s, (jfx::k_batch_compose<uint32_t, uint64_t><<<blocks, jfx::BC_BLOCK, 0, s>>>(
                                         (const uint32_t*)keys, oc, omin, key_bits, frame_offsets,
                                         nframes, n, (uint64_t*)comp, err)));
    }
  } else {
    if (cb == 4) {
      JFX_LAUNCH("batch_compose", s, (jfx::k_batch_compose<uint64_t, uint32_t><<<blocks, jfx::BC_BLOCK, 0, s>>>(
                                         (const uint64_t*)keys, oc, omin, key_bits, frame_offsets,
                                         nframes, n, (uint32_t*)comp, err)));
    } else {
      JFX_LAUNCH("batch_compose", s, (jfx::k_batch_compose<uint64_t, uint64_t><<<blocks, jfx::BC_BLOCK, 0, s>>>(
                                         (const uint64_t*)keys, oc, omin, key_bits, frame_offsets,
                                         nframes, n, (uint64_t*)comp, err)));
    }
  }
  JFX_HIP(hipGetLastError());
  jfx::SortCtx ctx;
  int rc = jfx::sort_prepare_known(ctx, ckt, comp, n, vb, digit_bits, total, sort_ws, sws, s,
                                   sort_shape, true);
  if (rc != JFX_OK) return rc;
  rc = jfx::sort_execute(ctx, comp, scomp, 1, vals, svals, 0, 0);
  if (rc != JFX_OK) return rc;
  jfx::SegCall c{scomp, jfx::ORD_UNSIGNED, svals, nullptr, n, out_keys, out_sum, out_count,
                 out_min, out_max, out_mean, false, 0u};
  rc = jfx::segreduce_dispatch(w, ckt, val_type, c, s);
  if (rc != JFX_OK) return rc;
  JFX_LAUNCH("copy_ngroups", s, (jfx::k_copy_u32<<<1, 64, 0, s>>>(w.ngroups, out_ngroups)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // extern "C"

namespace jfx {
namespace {

// One narrow table entry of a value type, with or without min / max codes.
size_t bn_part_bytes(int vt, bool mm) {
  switch (vt) {
    case JFX_F32: return mm ? sizeof(OpPart<float, true>) : sizeof(OpPart<float, false>);
    case JFX_I32: return mm ? sizeof(OpPart<int32_t, true>) : sizeof(OpPart<int32_t, false>);
    case JFX_U32: return mm ? sizeof(OpPart<uint32_t, true>) : sizeof(OpPart<uint32_t, false>);
    case JFX_I64: return mm ? sizeof(OpPart<int64_t, true>) : sizeof(OpPart<int64_t, false>);
    case JFX_U64: return mm ? sizeof(OpPart<uint64_t, true>) : sizeof(OpPart<uint64_t, false>);
  }
  return 0;
}

// Workspace of the routed pipeline (the same carving sizes it and uses it).
struct RoutedWs {
  void *comp = nullptr, *scomp = nullptr, *cvals = nullptr, *svals = nullptr;
  SegWs w{};
  void* sort_ws = nullptr;
  uint64_t sws = 0;
  void *tkeys = nullptr, *tsum = nullptr, *tmin = nullptr, *tmax = nullptr, *tmean = nullptr;
  int64_t* tcount = nullptr;
  void *part = nullptr, *mtab = nullptr;
  uint32_t *gn = nullptr, *base = nullptr, *wstart = nullptr;
};

void carve_routed(Carver& c, uint32_t total, int vt, uint32_t nframes, const jfx_batched_routing& r,
                  int outputs, RoutedWs& o) {
  const int cb = total <= 32 ? 4 : 8, vb = key_bytes(vt);
  const int mtb = (vt == JFX_I64 || vt == JFX_U64 || vt == JFX_F64) ? 8 : 4;
  const uint64_t nw = r.n_wide_rows;
  if (nw) {
    o.comp = c.take<char>(nw * cb);
    o.scomp = c.take<char>(nw * cb);
    o.cvals = c.take<char>(nw * vb);
    o.svals = c.take<char>(nw * vb);
    o.w = carve_seg(c, vb, nw);
    o.sws = sort_ws_bytes(cb == 4 ? JFX_U32 : JFX_U64, vb, nw, true);
    o.sort_ws = c.take<char>(o.sws);
    o.tkeys = c.take<char>(nw * cb);
    if (outputs & 1) o.tsum = c.take<char>(nw * vb);
    if (outputs & 2) o.tcount = c.take<int64_t>(nw);
    if (outputs & 4) o.tmin = c.take<char>(nw * vb);
    if (outputs & 8) o.tmax = c.take<char>(nw * vb);
    if (outputs & 16) o.tmean = c.take<char>(nw * mtb);
  }
  const size_t pb = bn_part_bytes(vt, (outputs & 12) != 0);
  o.part = c.take<char>(r.part_entries * pb);
  o.mtab = c.take<char>(r.merged_entries * pb);
  o.gn = c.take<uint32_t>((uint64_t)r.n_narrow + 1u);
  o.base = c.take<uint32_t>(nframes);
  o.wstart = c.take<uint32_t>(nframes);
}

template <typename K, typename V, typename MT, bool MM, typename CK>
int routed_run(int oc, int vt, const void* keys, const void* vals, const int64_t* offsets,
               uint32_t nframes, uint64_t omin, uint32_t key_bits, uint32_t total, int digit_bits,
               int sort_shape, const jfx_batched_routing& r, const RoutedWs& ws, void* out_keys,
               const GroupOut<V, MT>& out, uint32_t* out_ngroups, uint32_t* err, hipStream_t s) {
  using P = OpPart<V, MM>;
  using VU = typename std::conditional<sizeof(V) == 4, uint32_t, uint64_t>::type;
  using MU = typename std::conditional<sizeof(MT) == 4, uint32_t, uint64_t>::type;
  if (r.n_narrow) {
    JFX_LAUNCH("bn_table", s, (k_bn_table<K, V, MM><<<r.n_chunks, BR_W * 64, 0, s>>>(
                                  static_cast<const K*>(keys), static_cast<const V*>(vals), oc, r,
                                  static_cast<P*>(ws.part), err)));
    JFX_HIP(hipGetLastError());
    JFX_LAUNCH("bn_merge", s, (k_bn_merge<V, MM><<<r.n_narrow, 1024, 0, s>>>(
                                  r, static_cast<const P*>(ws.part), static_cast<P*>(ws.mtab), ws.gn)));
    JFX_HIP(hipGetLastError());
  }
  const uint64_t nw = r.n_wide_rows;
  if (nw) {
    const uint32_t blocks = (uint32_t)((nw + BC_ROWS - 1) / BC_ROWS);
    JFX_LAUNCH("batch_compose", s, (k_batch_compose_wide<K, CK, VU><<<blocks, BC_BLOCK, 0, s>>>(
                                       static_cast<const K*>(keys), static_cast<const VU*>(vals), oc,
                                       omin, key_bits, offsets, r, static_cast<CK*>(ws.comp),
                                       static_cast<VU*>(ws.cvals), err)));
    JFX_HIP(hipGetLastError());
    const int ckt = sizeof(CK) == 4 ? JFX_U32 : JFX_U64;
    SortCtx ctx;
    int rc = sort_prepare_known(ctx, ckt, ws.comp, nw, (int)sizeof(V), digit_bits, total, ws.sort_ws,
                                ws.sws, s, sort_shape, true);
    if (rc != JFX_OK) return rc;
    rc = sort_execute(ctx, ws.comp, ws.scomp, 1, ws.cvals, ws.svals, 0, 0);
    if (rc != JFX_OK) return rc;
    SegCall c{ws.scomp, ORD_UNSIGNED, ws.svals, nullptr, nw, ws.tkeys,
              out.sum ? ws.tsum : nullptr, out.count ? ws.tcount : nullptr,
              out.mn ? ws.tmin : nullptr, out.mx ? ws.tmax : nullptr,
              out.mean ? ws.tmean : nullptr, false, 0u};
    rc = segreduce_dispatch(ws.w, ckt, vt, c, s);
    if (rc != JFX_OK) return rc;
  }
  JFX_LAUNCH("batch_counts", s, (k_batch_counts<CK><<<(nframes + 255) / 256, 256, 0, s>>>(
                                    r, nframes, static_cast<const CK*>(ws.tkeys),
                                    nw ? ws.w.ngroups : nullptr, key_bits, ws.gn, ws.base,
                                    ws.wstart)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("batch_bases", s, (k_batch_bases<<<1, 1024, 0, s>>>(nframes, ws.base, out_ngroups)));
  JFX_HIP(hipGetLastError());
  if (r.n_narrow) {
    JFX_LAUNCH("bn_write", s, (k_bn_write<V, MT, MM, CK><<<r.n_narrow, 1024, 0, s>>>(
                                  r, static_cast<const P*>(ws.mtab), ws.base, omin, key_bits,
                                  static_cast<CK*>(out_keys), out)));
    JFX_HIP(hipGetLastError());
  }
  if (nw) {
    JFX_LAUNCH("bw_copy", s, (k_bw_copy<CK, VU, MU><<<2048, 256, 0, s>>>(
                                 ws.w.ngroups, key_bits, ws.base, ws.wstart,
                                 static_cast<const CK*>(ws.tkeys), static_cast<const VU*>(ws.tsum),
                                 ws.tcount, static_cast<const VU*>(ws.tmin),
                                 static_cast<const VU*>(ws.tmax), static_cast<const MU*>(ws.tmean),
                                 static_cast<CK*>(out_keys), reinterpret_cast<VU*>(out.sum), out.count,
                                 reinterpret_cast<VU*>(out.mn), reinterpret_cast<VU*>(out.mx),
                                 reinterpret_cast<MU*>(out.mean))));
    JFX_HIP(hipGetLastError());
  }
  return JFX_OK;
}

template <typename K, typename V, typename MT>
int routed_vt(int oc, int vt, const void* keys, const void* vals, const int64_t* offsets,
              uint32_t nframes, uint64_t omin, uint32_t key_bits, uint32_t total, int digit_bits,
              int sort_shape, const jfx_batched_routing& r, const RoutedWs& ws, void* out_keys,
              void* out_sum, int64_t* out_count, void* out_min, void* out_max, void* out_mean,
              uint32_t* out_ngroups, uint32_t* err, hipStream_t s) {
  const GroupOut<V, MT> o{static_cast<V*>(out_sum), out_count, static_cast<V*>(out_min),
                          static_cast<V*>(out_max), static_cast<MT*>(out_mean)};
  const bool mm = out_min || out_max;
  const bool c4 = total <= 32;
#define JFX_ROUTED(MMV, CKT)                                                                     \
  routed_run<K, V, MT, MMV, CKT>(oc, vt, keys, vals, offsets, nframes, omin, key_bits, total,    \
                                 digit_bits, sort_shape, r, ws, out_keys, o, out_ngroups, err, s)
  if (mm) return c4 ? JFX_ROUTED(true, uint32_t) : JFX_ROUTED(true, uint64_t);
  return c4 ? JFX_ROUTED(false, uint32_t) : JFX_ROUTED(false, uint64_t);
#undef JFX_ROUTED
}

template <typename K>
int routed_key(int oc, int vt, const void* keys, const void* vals, const int64_t* offsets,
               uint32_t nframes, uint64_t omin, uint32_t key_bits, uint32_t total, int digit_bits,
               int sort_shape, const jfx_batched_routing& r, const RoutedWs& ws, void* out_keys,
               void* out_sum, int64_t* out_count, void* out_min, void* out_max, void* out_mean,
               uint32_t* out_ngroups, uint32_t* err, hipStream_t s) {
#define JFX_ROUTED_VT(VT, MTT)                                                                   \
  routed_vt<K, VT, MTT>(oc, vt, keys, vals, offsets, nframes, omin, key_bits, total, digit_bits, \
                        sort_shape, r, ws, out_keys, out_sum, out_count, out_min, out_max,       \
                        out_mean, out_ngroups, err, s)
  switch (vt) {
    case JFX_F32: return JFX_ROUTED_VT(float, float);
    case JFX_I32: return JFX_ROUTED_VT(int32_t, float);
    case JFX_U32: return JFX_ROUTED_VT(uint32_t, float);
    case JFX_I64: return JFX_ROUTED_VT(int64_t, double);
    case JFX_U64: return JFX_ROUTED_VT(uint64_t, double);
  }
#undef JFX_ROUTED_VT
  return set_error(JFX_ERR_UNSUPPORTED, "the routed batched groupby takes no float64 values");
}

}  // namespace
}  // namespace jfx

extern "C" {

int jfx_batched_key_ranges(int key_type, const void* keys, const uint64_t* chunk_row0,
                           const uint32_t* chunk_rows, uint32_t nchunks, uint64_t* out_minmax,
                           void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64 && jfx::int_key(key_type), JFX_ERR_INVALID,
            "batched keys must be integers");
  hipStream_t s = jfx::as_stream(stream);
  if (nchunks == 0) return JFX_OK;
  JFX_CHECK(keys && chunk_row0 && chunk_rows && out_minmax, JFX_ERR_INVALID, "null argument");
  const int oc = jfx::order_code(key_type);
  if (jfx::key_bytes(key_type) == 4) {
    JFX_LAUNCH("batch_ranges", s, (jfx::k_batch_ranges<uint32_t><<<nchunks, 256, 0, s>>>(
                                      static_cast<const uint32_t*>(keys), oc, chunk_row0, chunk_rows,
                                      out_minmax)));
  } else {
    JFX_LAUNCH("batch_ranges", s, (jfx::k_batch_ranges<uint64_t><<<nchunks, 256, 0, s>>>(
                                      static_cast<const uint64_t*>(keys), oc, chunk_row0, chunk_rows,
                                      out_minmax)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int jfx_batched_routed_workspace_size(uint32_t total_bits, int val_type, uint32_t nframes,
                                      const jfx_batched_routing* route, int outputs,
                                      uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr && route != nullptr, JFX_ERR_INVALID, "null argument");
  JFX_CHECK(jfx::valid_vt(val_type) && val_type != JFX_F64, JFX_ERR_UNSUPPORTED,
            "the routed batched groupby takes integer or float32 values");
  JFX_CHECK(total_bits >= 1 && total_bits <= 64, JFX_ERR_INVALID, "total_bits must be 1..64");
  jfx::Carver c(nullptr, 0);
  jfx::RoutedWs o;
  jfx::carve_routed(c, total_bits, val_type, nframes, *route, outputs, o);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_batched_groupby_routed(int key_type, const void* keys, int val_type, const void* vals,
                               uint64_t n, const int64_t* frame_offsets, uint32_t nframes,
                               uint64_t key_min, uint32_t key_bits, int digit_bits, int sort_shape,
                               const jfx_batched_routing* route, void* out_keys, void* out_sum,
                               int64_t* out_count, void* out_min, void* out_max, void* out_mean,
                               uint32_t* out_ngroups, void* workspace, uint64_t ws_bytes,
                               void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64 && jfx::int_key(key_type), JFX_ERR_INVALID,
            "batched keys must be integers");
  JFX_CHECK(jfx::valid_vt(val_type) && val_type != JFX_F64, JFX_ERR_UNSUPPORTED,
            "the routed batched groupby takes integer or float32 values");
  JFX_CHECK(route != nullptr, JFX_ERR_INVALID, "route is null");
  JFX_CHECK(n >= 1 && n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be 1 .. 2^30 - 1 rows");
  JFX_CHECK(nframes >= 1, JFX_ERR_INVALID, "at least one frame");
  const jfx_batched_routing& r = *route;
  JFX_CHECK((uint64_t)r.n_narrow + r.n_wide == nframes, JFX_ERR_INVALID,
            "every frame must be narrow or wide");
  JFX_CHECK(r.n_wide_rows <= n, JFX_ERR_INVALID, "wide rows exceed the table");
  JFX_CHECK(r.n_narrow == 0 || r.n_chunks >= r.n_narrow, JFX_ERR_INVALID,
            "every narrow frame needs a chunk");
  uint32_t fbits = 0;
  while ((1ull << fbits) < nframes) ++fbits;
  const uint32_t total = fbits + key_bits;
  JFX_CHECK(key_bits >= 1 && total <= 64, JFX_ERR_UNSUPPORTED,
            "frame and key bits exceed 64 (narrow the key range)");
  JFX_CHECK(out_keys && out_ngroups && frame_offsets && keys && vals && workspace, JFX_ERR_INVALID,
            "null input, output or workspace");
  JFX_CHECK(digit_bits == 0 || digit_bits == 8 || digit_bits == 11, JFX_ERR_INVALID,
            "digit_bits must be 0, 8 or 11");
  JFX_CHECK(sort_shape >= 0 && sort_shape < jfx::SORT_SHAPES, JFX_ERR_INVALID,
            "sort_shape must be 0 to 3");
  const int outputs = (out_sum ? 1 : 0) | (out_count ? 2 : 0) | (out_min ? 4 : 0) |
                      (out_max ? 8 : 0) | (out_mean ? 16 : 0);
  jfx::Carver cv(workspace, ws_bytes);
  jfx::RoutedWs ws;
  jfx::carve_routed(cv, total, val_type, nframes, r, outputs, ws);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "routed batched groupby workspace too small");
  uint32_t* err = jfx::device_error_word();
  JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  hipStream_t s = jfx::as_stream(stream);
  const int oc = jfx::order_code(key_type);
  uint64_t omin = jfx::key_bytes(key_type) == 4 ? (key_min & 0xFFFFFFFFull) : key_min;
  if (oc == jfx::ORD_SIGNED) omin ^= jfx::key_bytes(key_type) == 4 ? 0x80000000ull : 0x8000000000000000ull;
  if (jfx::key_bytes(key_type) == 4)
    return jfx::routed_key<uint32_t>(oc, val_type, keys, vals, frame_offsets, nframes, omin, key_bits,
                                     total, digit_bits, sort_shape, r, ws, out_keys, out_sum,
                                     out_count, out_min, out_max, out_mean, out_ngroups, err, s);
  return jfx::routed_key<uint64_t>(oc, val_type, keys, vals, frame_offsets, nframes, omin, key_bits,
                                   total, digit_bits, sort_shape, r, ws, out_keys, out_sum, out_count,
                                   out_min, out_max, out_mean, out_ngroups, err, s);
}

}  // extern "C"
