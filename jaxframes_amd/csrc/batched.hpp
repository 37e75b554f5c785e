// Batched groupby: many independent groupby instances of uneven row counts,
// concatenated into one table with frame offsets, run as ONE device pipeline.
//
// Every row gets the composite key (frame << key_bits) | (ordered(key) -
// ordered(min)); one stable sort of the composites (over exactly
// frame_bits + key_bits bits, no host read) puts the rows of each frame
// together in key order, and one head-flag segmented reduction over the
// whole table yields every frame's groups, frame-major. The composite is
// built here by k_batch_compose.
#pragma once

#include "common.hpp"
#include "onepass.hpp"

namespace jfx {

constexpr int BC_BLOCK = 256;
constexpr int BC_ROWS = 4096;        // rows per workgroup
constexpr int BC_LOCAL_FRAMES = 64;  // frame offsets staged in LDS per workgroup

// First index f in [0, m] with a[f] > x (a nondecreasing, m+1 entries).
__device__ __forceinline__ uint32_t bc_upper(const int64_t* a, uint32_t m, int64_t x) {
  uint32_t lo = 0, hi = m + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// offsets: nframes + 1 nondecreasing row offsets (offsets[0] = 0,
// offsets[nframes] = n). Keys outside [min, min + 2^key_bits) raise
// DEVERR_BOUNDS (the range was fixed when the pipeline was set up).
template <typename K, typename CK>
__global__ void __launch_bounds__(BC_BLOCK)
k_batch_compose(const K* __restrict__ keys, int oc, uint64_t omin, uint32_t key_bits,
                const int64_t* __restrict__ offsets, uint32_t nframes, uint64_t n,
                CK* __restrict__ out, uint32_t* err) {
  __shared__ int64_t s_off[BC_LOCAL_FRAMES + 1];
  __shared__ uint32_t s_f0, s_nf;
  const uint64_t r0 = (uint64_t)blockIdx.x * BC_ROWS;
  if (r0 >= n) return;
  const uint64_t r1 = (r0 + BC_ROWS) < n ? (r0 + BC_ROWS) : n;
  if (threadIdx.x == 0) {
    const uint32_t f0 = bc_upper(offsets, nframes, (int64_t)r0) - 1u;
    const uint32_t f1 = bc_upper(offsets, nframes, (int64_t)(r1 - 1)) - 1u;
    s_f0 = f0;
    s_nf = f1 - f0 + 1u;
  }
  __syncthreads();
  const uint32_t f0 = s_f0, nf = s_nf;
  const bool local = nf <= (uint32_t)BC_LOCAL_FRAMES;
  if (local)
    for (uint32_t i = threadIdx.x; i <= nf; i += BC_BLOCK) s_off[i] = offsets[f0 + i];
  __syncthreads();
  const uint64_t kmask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
  constexpr int R = BC_ROWS / BC_BLOCK;
  uint32_t bad = 0;
  uint32_t f = 0;  // frame index relative to f0; rows only increase per thread
  // all R keys loaded before the first composite is formed
  K kv[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t r = r0 + threadIdx.x + (uint64_t)k * BC_BLOCK;
    kv[k] = r < r1 ? keys[r] : K(0);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t r = r0 + threadIdx.x + (uint64_t)k * BC_BLOCK;
    if (r >= r1) continue;
    uint32_t fr;
    if (local) {
      while (f + 1 < nf && s_off[f + 1] <= (int64_t)r) ++f;
      fr = f0 + f;
    } else {
      fr = bc_upper(offsets, nframes, (int64_t)r) - 1u;
    }
    const uint64_t d = (uint64_t)(ordered(kv[k], oc) - (K)omin);
    bad |= (d & ~kmask) != 0 ? 1u : 0u;
    const uint64_t fk = key_bits >= 64 ? 0ull : ((uint64_t)fr << key_bits);
    out[r] = (CK)(fk | (d & kmask));
  }
  if (ballot64(bad != 0u) != 0ull && lane_id() == 0) atomicOr(err, DEVERR_BOUNDS);
}

static __global__ void k_copy_u32(const uint32_t* src, uint32_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

static __global__ void k_zero_u32(uint32_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = 0u;
}

// ---- routed pipeline: narrow frames on one-pass tables ---------------------
//
// A frame whose keys span fewer than BR_TE values (measured once by
// k_batch_ranges when the plan is made) needs no sort: digit d =
// ordered(key) - ordered(frame minimum) indexes LDS tables directly.
//   k_bn_table        one workgroup per chunk of a narrow frame: every wave
//                     adds its rows into its own sum table (wave-private:
//                     fixed float add order), counts (and min / max codes)
//                     go to one table per workgroup; the chunk's merged
//                     table (span + 1 digits) is written out
//   k_bn_merge        one workgroup per narrow frame: its chunk tables in
//                     chunk order per digit, and the frame's group count
//   k_batch_compose_wide  the wide frames' rows compacted: composite keys and
//                     their values, for the sort and segmented reduction
//   k_batch_counts    every frame's group count (narrow: from k_bn_merge;
//                     wide: two binary searches in the sorted composite keys
//                     of the reduction)
//   k_batch_bases     one workgroup: each frame's first output position
//                     (exclusive scan of the counts) and the total
//   k_bn_write / k_bw_copy  narrow groups and wide groups at their positions
constexpr uint32_t BR_TE = 1024;
constexpr int BR_W = 4;

// Min and max ordered key of each row chunk.
template <typename K>
__global__ void __launch_bounds__(256)
k_batch_ranges(const K* __restrict__ keys, int oc, const uint64_t* __restrict__ row0,
               const uint32_t* __restrict__ rows, uint64_t* __restrict__ out) {
  __shared__ K s_lo[4], s_hi[4];
  const uint32_t c = blockIdx.x;
  const uint64_t r0 = row0[c];
  const uint32_t m = rows[c];
  K lo = ~K(0), hi = K(0);
  for (uint32_t i = threadIdx.x; i < m; i += 256) {
    const K o = ordered(keys[r0 + i], oc);
    lo = o < lo ? o : lo;
    hi = o > hi ? o : hi;
  }
  lo = wave_min_k(lo);
  hi = wave_max_k(hi);
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_lo[wid] = lo;
    s_hi[wid] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    out[2 * (uint64_t)c] = (uint64_t)lo;
    out[2 * (uint64_t)c + 1] = (uint64_t)hi;
  }
}

// Tables of one chunk. Lane-group tables as in the one-pass kernel: rt is
// the smallest power of two above the frame's span (at least 16), the BR_TE
// entries hold BR_TE / rt tables of rt digits, and lane l uses table l mod
// (BR_TE / rt) with its digits XOR-swizzled by the table index, so a few
// distinct keys do not put a wave's lanes on one entry.
template <typename K, typename V, bool MM>
__global__ void __launch_bounds__(BR_W * 64)
k_bn_table(const K* __restrict__ keys, const V* __restrict__ vals, int oc, jfx_batched_routing r,
           OpPart<V, MM>* __restrict__ part, uint32_t* err) {
  static_assert(!op_dd<V>, "float64 values are not routed");
  using T = typename OpT<V>::T;
  using U = typename ValTraits<V>::U;
  using C = MmCode<V>;
  constexpr uint32_t STEP = 64u * 16u;
  __shared__ T s_acc[BR_W * BR_TE];
  __shared__ uint32_t s_cnt[BR_TE];
  __shared__ U s_min[MM ? BR_TE : 1], s_max[MM ? BR_TE : 1];
  const uint32_t c = blockIdx.x;
  const uint32_t j = r.chunk_narrow[c];
  const uint64_t r0 = r.chunk_row0[c];
  const uint32_t m = r.chunk_rows[c];
  const K omin = (K)r.narrow_omin[j];
  const uint32_t span = r.narrow_span[j];
  if (span >= BR_TE) {  // not a narrow frame: a malformed plan (block-uniform)
    if (threadIdx.x == 0) atomicOr(err, DEVERR_BOUNDS);
    return;
  }
  uint32_t rt = 16u;
  while (rt <= span) rt <<= 1;
  const uint32_t lg = BR_TE / rt;
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  for (uint32_t i = threadIdx.x; i < BR_TE; i += BR_W * 64) {
    s_cnt[i] = 0u;
    if constexpr (MM) {
      s_min[i] = ~U(0);
      s_max[i] = U(0);
    }
  }
  for (uint32_t i = lane; i < BR_TE; i += 64) s_acc[wid * BR_TE + i] = T(0);
  __syncthreads();
  const uint32_t g = lane & (lg - 1u);
  T* accs = s_acc + wid * BR_TE + g * rt;
  uint32_t* cnts = s_cnt + g * rt;
  U* mins = s_min + (MM ? g * rt : 0u);
  U* maxs = s_max + (MM ? g * rt : 0u);
  bool bad = false;
  // steps of 1024 rows, wave w taking steps w, w + BR_W, ...; lane l reads
  // rows step + i * 64 + l (one 64-row run per load instruction)
  for (uint32_t s0 = (uint32_t)wid * STEP; s0 < m; s0 += BR_W * STEP) {
    K kk[16];
    V vv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t e = s0 + (uint32_t)i * 64u + lane;
      kk[i] = K(0);
      vv[i] = V(0);
      if (e < m) {
        kk[i] = keys[r0 + e];
        vv[i] = vals[r0 + e];
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t e = s0 + (uint32_t)i * 64u + lane;
      const K dk = ordered(kk[i], oc) - omin;
      const bool in = dk <= (K)span;
      bad |= (e < m) && !in;
      if (e < m && in) {
        const uint32_t d = ((uint32_t)dk ^ g) & (rt - 1u);
        __hip_atomic_fetch_add(&accs[d], (T)vv[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_fetch_add(&cnts[d], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if constexpr (MM) {
          __hip_atomic_fetch_min(&mins[d], C::lo(vv[i]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_max(&maxs[d], C::hi(vv[i]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
  }
  if (ballot64(bad) != 0ull && lane == 0) atomicOr(err, DEVERR_BOUNDS);
  __syncthreads();
  // digit d: its lane-group tables in (wave, table) order
  OpPart<V, MM>* out = part + r.chunk_part[c];
  for (uint32_t d = threadIdx.x; d <= span; d += BR_W * 64) {
    OpAcc<V, MM> st;
    st.init();
    for (int w = 0; w < BR_W; ++w)
      for (uint32_t t = 0; t < lg; ++t) st.s += s_acc[w * BR_TE + t * rt + ((d ^ t) & (rt - 1u))];
    for (uint32_t t = 0; t < lg; ++t) {
      const uint32_t e = t * rt + ((d ^ t) & (rt - 1u));
      st.n += s_cnt[e];
      if constexpr (MM) {
        st.mn = s_min[e] < st.mn ? s_min[e] : st.mn;
        st.mx = s_max[e] > st.mx ? s_max[e] : st.mx;
      }
    }
    out[d] = st.pack();
  }
}

// Frame j's chunk tables merged per digit (thread = digit) in chunk order;
// gn[j] = digits with rows.
template <typename V, bool MM>
__global__ void __launch_bounds__(1024)
k_bn_merge(jfx_batched_routing r, const OpPart<V, MM>* __restrict__ part,
           OpPart<V, MM>* __restrict__ mtab, uint32_t* __restrict__ gn) {
  const uint32_t j = blockIdx.x;
  const uint32_t span = r.narrow_span[j];
  const uint32_t c0 = r.narrow_chunk0[j], c1 = r.narrow_chunk0[j + 1];
  const uint32_t d = threadIdx.x;
  bool present = false;
  if (span < BR_TE && d <= span) {
    OpAcc<V, MM> st;
    st.init();
    merge_seq<16>(st, part + r.chunk_part[c0] + d, (size_t)span + 1u, c1 - c0);
    mtab[r.narrow_mbase[j] + d] = st.pack();
    present = st.n != 0u;
  }
  const int cnt = __syncthreads_count(present);
  if (threadIdx.x == 0) gn[j] = (uint32_t)cnt;
}

// Compacted row i of wide frame j (wide_offsets[j] <= i < wide_offsets[j+1])
// is row offsets[wide_frame[j]] + i - wide_offsets[j]; its composite key and
// value (moved as raw bits: VB bytes).
template <typename K, typename CK, typename VU>
__global__ void __launch_bounds__(BC_BLOCK)
k_batch_compose_wide(const K* __restrict__ keys, const VU* __restrict__ vals, int oc, uint64_t omin,
                     uint32_t key_bits, const int64_t* __restrict__ offsets, jfx_batched_routing r,
                     CK* __restrict__ comp, VU* __restrict__ cvals, uint32_t* err) {
  __shared__ int64_t s_woff[BC_LOCAL_FRAMES + 1];
  __shared__ int64_t s_src[BC_LOCAL_FRAMES];
  __shared__ uint32_t s_fr[BC_LOCAL_FRAMES];
  __shared__ uint32_t s_j0, s_nj;
  const uint64_t n = r.n_wide_rows;
  const uint64_t i0 = (uint64_t)blockIdx.x * BC_ROWS;
  if (i0 >= n) return;
  const uint64_t i1 = (i0 + BC_ROWS) < n ? (i0 + BC_ROWS) : n;
  if (threadIdx.x == 0) {
    const uint32_t j0 = bc_upper(r.wide_offsets, r.n_wide, (int64_t)i0) - 1u;
    const uint32_t j1 = bc_upper(r.wide_offsets, r.n_wide, (int64_t)(i1 - 1)) - 1u;
    s_j0 = j0;
    s_nj = j1 - j0 + 1u;
  }
  __syncthreads();
  const uint32_t j0 = s_j0, nj = s_nj;
  const bool local = nj <= (uint32_t)BC_LOCAL_FRAMES;
  if (local) {
    for (uint32_t t = threadIdx.x; t <= nj; t += BC_BLOCK) s_woff[t] = r.wide_offsets[j0 + t];
    for (uint32_t t = threadIdx.x; t < nj; t += BC_BLOCK) {
      const uint32_t f = r.wide_frame[j0 + t];
      s_fr[t] = f;
      s_src[t] = offsets[f];
    }
  }
  __syncthreads();
  const uint64_t kmask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
  constexpr int R = BC_ROWS / BC_BLOCK;
  uint32_t bad = 0;
  uint32_t jj = 0;  // wide frame relative to j0; rows only increase per thread
  // every row's source first, so that all R key and value loads are in flight
  // together (one row at a time measured 1.08 ms for 2.4e8 rows)
  uint64_t src[R];
  uint32_t fr[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t i = i0 + threadIdx.x + (uint64_t)k * BC_BLOCK;
    src[k] = 0;
    fr[k] = 0;
    if (i < i1) {
      if (local) {
        while (jj + 1 < nj && s_woff[jj + 1] <= (int64_t)i) ++jj;
        fr[k] = s_fr[jj];
        src[k] = (uint64_t)(s_src[jj] + ((int64_t)i - s_woff[jj]));
      } else {
        const uint32_t j = bc_upper(r.wide_offsets, r.n_wide, (int64_t)i) - 1u;
        fr[k] = r.wide_frame[j];
        src[k] = (uint64_t)(offsets[fr[k]] + ((int64_t)i - r.wide_offsets[j]));
      }
    }
  }
  K kv[R];
  VU vv[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t i = i0 + threadIdx.x + (uint64_t)k * BC_BLOCK;
    if (i < i1) {
      kv[k] = keys[src[k]];
      vv[k] = vals[src[k]];
    }
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const uint64_t i = i0 + threadIdx.x + (uint64_t)k * BC_BLOCK;
    if (i < i1) {
      const uint64_t d = (uint64_t)(ordered(kv[k], oc) - (K)omin);
      bad |= (d & ~kmask) != 0 ? 1u : 0u;
      const uint64_t fk = key_bits >= 64 ? 0ull : ((uint64_t)fr[k] << key_bits);
      comp[i] = (CK)(fk | (d & kmask));
      cvals[i] = vv[k];
    }
  }
  if (ballot64(bad != 0u) != 0ull && lane_id() == 0) atomicOr(err, DEVERR_BOUNDS);
}

template <typename CK>
__device__ __forceinline__ uint32_t br_lower(const CK* a, uint32_t n, CK x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// cnt[f]: frame f's group count (narrow: from k_bn_merge; wide: two binary
// searches in the reduction's *gw ascending composite keys, gw null: no wide
// rows); wstart[f]: index of wide frame f's first group there. One thread
// per frame.
template <typename CK>
__global__ void __launch_bounds__(256)
k_batch_counts(jfx_batched_routing r, uint32_t nframes, const CK* __restrict__ wkeys,
               const uint32_t* gw_ptr, uint32_t key_bits, const uint32_t* __restrict__ gn,
               uint32_t* __restrict__ cnt, uint32_t* __restrict__ wstart) {
  const uint32_t f = blockIdx.x * 256u + threadIdx.x;
  if (f >= nframes) return;
  const uint32_t gw = gw_ptr ? *gw_ptr : 0u;
  const uint32_t jn = r.frame_narrow[f];
  uint32_t c = 0u, ws = 0u;
  if (jn != 0xFFFFFFFFu) {
    c = gn[jn];
  } else if (gw != 0u) {
    const uint64_t kmask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
    const uint64_t lo = key_bits >= 64 ? 0ull : ((uint64_t)f << key_bits);
    ws = br_lower(wkeys, gw, (CK)lo);
    // first key above the frame's: its largest composite + 1, unless that is
    // the top of the key type
    const uint64_t top = lo | kmask;
    const bool last = sizeof(CK) == 4 ? top >= 0xFFFFFFFFull : top == ~0ull;
    const uint32_t we = last ? gw : br_lower(wkeys, gw, (CK)(top + 1u));
    c = we - ws;
  }
  cnt[f] = c;
  wstart[f] = ws;
}

// base[f]: first output position of frame f's groups (exclusive scan of
// cnt, in place); *out_ngroups: the total. One workgroup.
__global__ void __launch_bounds__(1024)
k_batch_bases(uint32_t nframes, uint32_t* __restrict__ base, uint32_t* out_ngroups) {
  __shared__ uint32_t s_c[1024];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  const uint32_t tid = threadIdx.x;
  uint32_t carry = 0u;
  for (uint32_t f0 = 0; f0 < nframes; f0 += 1024) {
    const uint32_t f = f0 + tid;
    s_c[tid] = f < nframes ? base[f] : 0u;
    __syncthreads();
    const uint32_t tot = block_excl_scan_lds<1024, 1024>(s_c, s_tmp);
    if (f < nframes) base[f] = carry + s_c[tid];
    carry += tot;
    __syncthreads();
  }
  if (tid == 0) *out_ngroups = carry;
}

// Narrow frame j's groups (digits with rows) at base[frame] on, key order.
template <typename V, typename MT, bool MM, typename CK>
__global__ void __launch_bounds__(1024)
k_bn_write(jfx_batched_routing r, const OpPart<V, MM>* __restrict__ mtab,
           const uint32_t* __restrict__ base, uint64_t omin_g, uint32_t key_bits,
           CK* __restrict__ out_keys, GroupOut<V, MT> out) {
  using T = typename OpT<V>::T;
  __shared__ uint32_t s_p[1024];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  const uint32_t j = blockIdx.x;
  const uint32_t f = r.narrow_frame[j];
  const uint32_t span = r.narrow_span[j];
  const uint32_t d = threadIdx.x;
  OpAcc<V, MM> st;
  st.init();
  const bool in = span < BR_TE && d <= span;
  if (in) st.add(mtab[r.narrow_mbase[j] + d]);
  const bool has = in && st.n != 0u;
  s_p[d] = has ? 1u : 0u;
  __syncthreads();
  block_excl_scan_lds<1024, 1024>(s_p, s_tmp);
  if (!has) return;
  const uint64_t g = (uint64_t)base[f] + s_p[d];
  const uint64_t low = r.narrow_omin[j] + d - omin_g;
  const uint64_t fk = key_bits >= 64 ? 0ull : ((uint64_t)f << key_bits);
  out_keys[g] = (CK)(fk | low);
  const T sum = st.s;
  if (out.sum) out.sum[g] = (V)sum;
  if (out.count) out.count[g] = (int64_t)st.n;
  if constexpr (MM) {
    if (out.mn) out.mn[g] = MmCode<V>::decode(st.mn);
    if (out.mx) out.mx[g] = MmCode<V>::decode(st.mx);
  }
  if (out.mean) {
    if constexpr (ValTraits<V>::is_float) out.mean[g] = (MT)(sum / (double)st.n);
    else out.mean[g] = (MT)((double)(V)sum / (double)st.n);
  }
}

// The reduction's *gw wide groups, g-th at base[f] + g - wstart[f] (f the
// frame in its composite key); columns moved as raw bits (VB value bytes, MB
// mean bytes).
template <typename CK, typename VU, typename MU>
__global__ void __launch_bounds__(256)
k_bw_copy(const uint32_t* gw_ptr, uint32_t key_bits, const uint32_t* __restrict__ base,
          const uint32_t* __restrict__ wstart, const CK* __restrict__ tkeys,
          const VU* __restrict__ tsum, const int64_t* __restrict__ tcount,
          const VU* __restrict__ tmin, const VU* __restrict__ tmax, const MU* __restrict__ tmean,
          CK* __restrict__ okeys, VU* __restrict__ osum, int64_t* __restrict__ ocount,
          VU* __restrict__ omin, VU* __restrict__ omax, MU* __restrict__ omean) {
  const uint32_t gw = *gw_ptr;
  for (uint32_t g = blockIdx.x * 256u + threadIdx.x; g < gw; g += gridDim.x * 256u) {
    const CK k = tkeys[g];
    const uint32_t f = key_bits >= 64 ? 0u : (uint32_t)((uint64_t)k >> key_bits);
    const uint64_t p = (uint64_t)base[f] + (g - wstart[f]);
    okeys[p] = k;
    if (osum) osum[p] = tsum[g];
    if (ocount) ocount[p] = tcount[g];
    if (omin) omin[p] = tmin[g];
    if (omax) omax[p] = tmax[g];
    if (omean) omean[p] = tmean[g];
  }
}

}  // namespace jfx
