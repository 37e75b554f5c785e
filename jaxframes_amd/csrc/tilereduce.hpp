// Tile-sorted groupby for small key ranges (max - min < 2048).
//
// Each workgroup owns a contiguous chunk of rows. Per tile it stably counting-
// sorts the rows by digit = ordered(key) - min inside LDS (wave64 ballot
// multisplit ranking, the ranking of the radix scatter), then runs the
// head-flag segmented reduction over the LDS-sorted tile: every digit run of
// the tile is reduced in row order and merged into the chunk's accumulator
// for that digit. The chunk accumulators go to HBM ([chunk][digit]) and a
// combine kernel merges them in chunk order. The sorted order never leaves
// the chip: HBM traffic is one read of keys and values plus the accumulators,
// and every sum is combined in a fixed order, so results are reproducible.
#pragma once

#include <type_traits>

#include "blockscan.hpp"
#include "common.hpp"
#include "onesweep.hpp"
#include "segreduce.hpp"

namespace jfx {

template <typename K, typename V, bool MM>
struct TileRedArgs {
  const K* keys;
  int oc;
  K sub;                     // ordered min
  uint32_t R;                // digits in use (max - min + 1)
  const V* vals;
  uint64_t n;
  uint64_t chunk;            // rows per chunk (multiple of the tile)
  AggState<V, MM>* part;     // [nchunks][R] chunk accumulators
  const uint32_t* sel;       // skew selector: run only when it is nonzero (null: always run)
};

template <typename V, bool MM, int W>
__device__ __forceinline__ void tr_seg_scan(uint32_t f, const AggState<V, MM>& st, uint32_t& ef,
                                            AggState<V, MM>& est, AggState<V, MM>& tst,
                                            AggState<V, MM>* s_wst, uint32_t* s_wf) {
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
    for (int w = 0; w < W; ++w) {
      const AggState<V, MM> t = s_wst[w];
      const uint32_t tfw = s_wf[w];
      s_wst[w] = c;
      s_wf[w] = cf;
      if (tfw) c = t;
      else c.merge(t);
      cf |= tfw;
    }
    s_wst[W] = c;
    s_wf[W] = cf;
  }
  __syncthreads();
  ef = s_wf[wid];
  est = s_wst[wid];
  if (exf) est = ex;
  else est.merge(ex);
  ef |= exf;
  tst = s_wst[W];
  __syncthreads();
}

// Register budget: two 8-wave workgroups per CU (at most 128 VGPRs).
template <typename K, typename V, bool MM, int D, int W, int KPT>
__global__ void __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_tile_reduce(TileRedArgs<K, V, MM> a) {
  constexpr int RADIX = 1 << D;
  constexpr int BLOCK = W * 64;
  constexpr int TILE = BLOCK * KPT;
  constexpr uint32_t MASK = RADIX - 1;
  using HT = typename std::conditional<(D > 8), uint16_t, uint32_t>::type;
  using St = AggState<V, MM>;

  __shared__ HT s_whist[W * RADIX];
  __shared__ uint32_t s_tstart[RADIX];
  __shared__ V s_v[TILE];
  __shared__ uint16_t s_d[TILE];
  __shared__ St s_acc[RADIX];
  __shared__ St s_wst[W + 1];
  __shared__ uint32_t s_wf[W + 1];
  __shared__ uint32_t s_tmp[W + 1];

  if (a.sel && *a.sel == 0u) return;
  const int tid = threadIdx.x;
  const int wid = tid >> 6;
  const uint32_t lane = lane_id();
  const uint32_t R = a.R;
  const uint32_t RL = ((R + 63u) & ~63u) < (uint32_t)RADIX ? ((R + 63u) & ~63u) : (uint32_t)RADIX;
  const uint32_t b = blockIdx.x;
  const uint64_t cbeg = (uint64_t)b * a.chunk;
  const uint64_t cend = (cbeg + a.chunk) < a.n ? (cbeg + a.chunk) : a.n;
  for (uint32_t d = tid; d < RL; d += BLOCK) s_acc[d] = St::identity();

  HT* wh = s_whist + wid * RADIX;
  const uint32_t wb = (uint32_t)wid * 64u * KPT;
  for (uint64_t tbase = cbeg; tbase < cend; tbase += TILE) {
    const uint32_t tn = (uint32_t)((cend - tbase) < (uint64_t)TILE ? (cend - tbase) : (uint64_t)TILE);
    for (int w = 0; w < W; ++w)
      for (uint32_t d = tid; d < RL; d += BLOCK) s_whist[w * RADIX + d] = HT(0);
    uint32_t dg[KPT];
    uint32_t slot[KPT];
    V val[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      const K k = li < tn ? a.keys[tbase + li] : K(0);
      val[i] = li < tn ? a.vals[tbase + li] : V(0);
      dg[i] = (uint32_t)(ordered(k, a.oc) - a.sub) & MASK;
    }
    __syncthreads();
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
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    for (uint32_t d = tid; d < RADIX; d += BLOCK) {
      uint32_t sum = 0;
      if (d < RL) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const uint32_t c = s_whist[w * RADIX + d];
          s_whist[w * RADIX + d] = (HT)sum;
          sum += c;
        }
      }
      s_tstart[d] = sum;
    }
    __syncthreads();
    block_excl_scan_lds<RADIX, BLOCK>(s_tstart, s_tmp);
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const uint32_t li = wb + (uint32_t)i * 64u + lane;
      if (li < tn) {
        const uint32_t sl = slot[i] + s_tstart[dg[i]] + wh[dg[i]];
        s_v[sl] = val[i];
        s_d[sl] = (uint16_t)dg[i];
      }
    }
    __syncthreads();

    // head-flag segmented reduction over the sorted tile, KPT rows per thread
    const uint32_t e0 = (uint32_t)tid * KPT;
    St cur = St::identity();
    St lead = St::identity();
    uint32_t seen = 0u;
    uint32_t curd = 0u;
    uint32_t efh = 0u;  // first head position in this thread
    uint32_t prevd = (e0 > 0 && e0 - 1 < tn) ? s_d[e0 - 1] : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const uint32_t e = e0 + j;
      if (e < tn) {
        const uint32_t d = s_d[e];
        if (e == 0 || d != prevd) {
          if (seen) {
            s_acc[curd].merge(cur);  // a run wholly inside this thread
          } else {
            lead = cur;
            efh = e;
          }
          seen = 1u;
          curd = d;
          cur = St::identity();
        }
        cur.add(s_v[e]);
        prevd = d;
      }
    }
    if (!seen) lead = cur;
    uint32_t ef;
    St est, tst;
    tr_seg_scan<V, MM, W>(seen, seen ? cur : lead, ef, est, tst, s_wst, s_wf);
    if (seen && efh > 0) {
      // the run ending just before this thread's first head
      est.merge(lead);
      s_acc[s_d[efh - 1]].merge(est);
    }
    if (tid == 0 && tn > 0) s_acc[s_d[tn - 1]].merge(tst);  // the tile's last run
    __syncthreads();
  }
  for (uint32_t d = tid; d < R; d += BLOCK) a.part[(size_t)b * R + d] = s_acc[d];
}

// Output pointers of the groupby aggregations (null = not requested).
template <typename V, typename MT>
struct GroupOut {
  V* sum;
  int64_t* count;
  V* mn;
  V* mx;
  MT* mean;
};

// Merge the chunk accumulators of each digit in chunk order (one wave per
// digit, a fixed-shape tree across lanes) and write the non-empty digits'
// results at their group index.
template <typename V, typename MT, bool MM>
__global__ void __launch_bounds__(256)
k_tile_combine(const AggState<V, MM>* part, uint32_t nchunks, uint32_t R, const uint32_t* dgid,
               GroupOut<V, MT> out, const uint32_t* total, const uint32_t* sel,
               uint32_t nchunks_sel) {
  const uint32_t d = blockIdx.x * 4u + (threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  if (d >= R) return;
  // the partial count of the kernel that ran: nchunks when the selector is
  // zero (or absent), nchunks_sel when it is nonzero
  if (sel && *sel != 0u) nchunks = nchunks_sel;
  const uint32_t per = (nchunks + 63u) / 64u;
  const uint32_t c0 = lane * per;
  const uint32_t c1 = (c0 + per) < nchunks ? (c0 + per) : nchunks;
  AggState<V, MM> st = AggState<V, MM>::identity();
  for (uint32_t c = c0; c < c1; ++c) st.merge(part[(size_t)c * R + d]);
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    AggState<V, MM> o;
    o.s = __shfl_down(st.s, off, 64);
    o.c = __shfl_down(st.c, off, 64);
    o.n = __shfl_down(st.n, off, 64);
    o.mn = MM ? __shfl_down(st.mn, off, 64) : st.mn;
    o.mx = MM ? __shfl_down(st.mx, off, 64) : st.mx;
    if ((lane & (2u * off - 1u)) == 0u) st.merge(o);
  }
  if (lane == 0) {
    // row counts come from the digit histogram (exact; the wave-table kernel
    // does not count rows)
    if (total) st.n = total[d];
    const uint32_t g = dgid[d];
    if (g != 0xFFFFFFFFu) {
      if (out.sum) out.sum[g] = st.sum();
      if (out.count) out.count[g] = (int64_t)st.n;
      if constexpr (MM) {
        if (out.mn) out.mn[g] = st.mn;
        if (out.mx) out.mx[g] = st.mx;
      }
      if (out.mean) {
        if constexpr (ValTraits<V>::is_float) out.mean[g] = (MT)st.sum() / (MT)(st.n ? st.n : 1u);
        else out.mean[g] = (MT)((double)st.s / (double)(st.n ? st.n : 1u));
      }
    }
  }
}

}  // namespace jfx
