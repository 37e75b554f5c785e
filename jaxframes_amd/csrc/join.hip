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
//   k_join_count    per tile: each row's bounds by a galloping search inside
//                   the window staged in LDS; the tile's pair count
//   k_join_scan     exclusive scan of tile counts (M read back by the host)
//   k_join_fill     per chunk of at most JFCH outputs of a tile: bounds
//                   again, an in-tile scan of counts, each row marking its
//                   outputs in LDS, then the outputs written in order, so a
//                   tile (or a row) with many matches is spread over chunks
//   k_join_first    left/right joins: first match or -1 (:1000-1034)
//   k_gather        payload gather by uint32 row, fill for row == ~0
//   k_pf_onepass (k_pf_count / k_pf_scan / k_pf_scatter above 2^30 rows)
//                   inner-join range pre-filter of one side before its sort:
//                   a row whose key lies outside [first, last] of the other
//                   side's sorted keys has no match (:989-994 pairs only
//                   equal keys), so it is dropped; the kept rows keep their
//                   order, and their payload columns are packed into the
//                   sort's records in the same pass
// Float keys: NaN never matches (== in the match matrix is false); -0.0
// matches +0.0. Comparisons run on the order-preserving unsigned image.
#include <algorithm>
#include <cstdlib>
#include <cstring>

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

// First index i in [p, n) of sorted LDS array s with s[i] > x (upper) or
// s[i] >= x (lower), n if none: probes p, p + 1, p + 3, p + 7, ... and
// bisects the last bracket, so a bound d entries past p takes about
// 2 log2(d) + 1 reads instead of log2(n - p).
template <typename K>
__device__ __forceinline__ uint32_t gallop(const K* s, uint32_t p, uint32_t n, K x, bool upper) {
  uint32_t lo = p, hi = n, step = 1u;
  while (true) {
    const uint32_t q = lo + step - 1u;
    if (q >= n) break;
    const K v = s[q];
    if (upper ? (v > x) : (v >= x)) {
      hi = q;
      break;
    }
    lo = q + 1u;
    step <<= 1;
  }
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const K v = s[mid];
    if (upper ? (v <= x) : (v < x)) lo = mid + 1u;
    else hi = mid;
  }
  return lo;
}

// The raw keys of this thread's JIPT consecutive rows of the tile starting
// at tb (rows past L read as 0 and are not used): 16-byte loads when the
// tile is whole and aligned.
template <typename K>
__device__ __forceinline__ void tile_rows(const K* sl, uint64_t L, uint64_t tb, K (&raw)[JIPT]) {
  const uint64_t r0 = tb + (uint64_t)threadIdx.x * JIPT;
  constexpr int PER = 16 / (int)sizeof(K);
  static_assert(JIPT % PER == 0, "rows per thread fill whole 16-byte loads");
  if (tb + JT <= L && (reinterpret_cast<uintptr_t>(sl + r0) & 15u) == 0) {
    const uint4* q = reinterpret_cast<const uint4*>(sl + r0);
#pragma unroll
    for (int v = 0; v < JIPT / PER; ++v) {
      const uint4 x = q[v];
      __builtin_memcpy(&raw[v * PER], &x, 16);
    }
  } else {
#pragma unroll
    for (int j = 0; j < JIPT; ++j) raw[j] = r0 + j < L ? sl[r0 + j] : K(0);
  }
}

// Bounds of the tile's rows: lo[j], cnt[j] for j in this thread's rows.
// A thread's rows are consecutive and sorted, so with the window staged in
// LDS a row equal to the previous one reuses its bounds and a larger one
// gallops from the previous upper bound (the 1e8 x 1e7 join: about 10 left
// rows per right key, so most bounds lie 0-2 entries on).
// The caller loads the rows (tile_rows) before staging the window, so the
// two reads are in flight together.
template <typename K>
__device__ __forceinline__ void tile_bounds(const K (&raw)[JIPT], uint64_t L, const K* sr, int oc,
                                            uint64_t tb, uint64_t w0, uint64_t w1, const K* s_w,
                                            bool staged, uint64_t* lo, uint32_t* cnt) {
  const uint32_t wn = (uint32_t)(w1 - w0);
  bool have = false;
  K px = K(0);
  uint32_t pa = 0u, pb = 0u;
#pragma unroll
  for (int j = 0; j < JIPT; ++j) {
    const uint64_t i = tb + (uint64_t)threadIdx.x * JIPT + j;
    lo[j] = 0;
    cnt[j] = 0;
    if (i < L) {
      if (is_nan_key(raw[j], oc)) continue;
      const K x = ordered(raw[j], oc);
      if (staged) {
        if (!(have && x == px)) {
          pa = gallop(s_w, have ? pb : 0u, wn, x, false);
          pb = gallop(s_w, pa, wn, x, true);
          px = x;
          have = true;
        }
        lo[j] = w0 + pa;
        cnt[j] = pb - pa;
      } else {
        const uint64_t a = gbound(sr, w0, w1, x, oc, false);
        const uint64_t b = gbound(sr, a, w1, x, oc, true);
        lo[j] = a;
        cnt[j] = (uint32_t)(b - a);
      }
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
  K raw[JIPT];
  tile_rows(sl, L, tb, raw);
  const bool staged = stage_window(sr, oc, w0, w1, s_w);
  __syncthreads();
  uint64_t lo[JIPT];
  uint32_t cnt[JIPT];
  tile_bounds(raw, L, sr, oc, tb, w0, w1, s_w, staged, lo, cnt);
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

// Fill chunks: a tile's pairs are written in chunks of at most JFCH outputs,
// each chunk by one workgroup, so a tile whose rows match long right runs (a
// key hot on both sides) is spread over many workgroups.
constexpr uint32_t JFCH = 4096;
// chunk map entries beyond one per tile: a tile of c pairs has
// ceil(c / JFCH) chunks, so fewer than 2^32 pairs need at most
// tiles + 2^32 / JFCH entries
constexpr uint64_t JCMAP_EXTRA = (1ull << 32) / JFCH;
constexpr uint32_t JSCAN = 8192;  // tile counts staged per round of k_join_scan

// Tile counts are scanned in rounds of JSCAN, one workgroup per round, in two
// kernels: k_join_scan_sums takes every round's pair and chunk totals, then
// k_join_scan, one workgroup per round again, starts from the totals of the
// rounds before it. (One workgroup walking the rounds in order took 84-88 us
// at 48,828 tiles.)
__device__ __forceinline__ void block_sum2(uint64_t& a, uint64_t& b, uint64_t (*s_w)[17]) {
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if (lane == 0) {
    s_w[0][wid] = a;
    s_w[1][wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t ta = 0, tb = 0;
    for (int w = 0; w < 16; ++w) {
      ta += s_w[0][w];
      tb += s_w[1][w];
    }
    a = ta;
    b = tb;
  }
}

// bsum[2 r], bsum[2 r + 1]: the pair and fill-chunk totals of round r.
__global__ void __launch_bounds__(1024)
k_join_scan_sums(const uint64_t* c, uint32_t n, uint64_t* bsum) {
  __shared__ uint64_t s_w[2][17];
  const uint32_t base = blockIdx.x * JSCAN;
  const uint32_t m = (n - base) < JSCAN ? (n - base) : JSCAN;
  uint64_t sp = 0, sc = 0;
  for (uint32_t i = threadIdx.x; i < m; i += 1024) {
    const uint64_t v = c[base + i];
    sp += v;
    sc += (v + JFCH - 1) / JFCH;
  }
  block_sum2(sp, sc, s_w);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = sp;
    bsum[2 * blockIdx.x + 1] = sc;
  }
}

// Round blockIdx.x of the exclusive scan of the n tile pair counts in place
// (c), the first chunk of every tile (cfirst), the tile of every chunk (cmap,
// at most cmap_cap entries), and, in the last round, the totals: total[0] =
// pairs, total[1] = chunks. Each thread scans 8 consecutive counts.
__global__ void __launch_bounds__(1024)
k_join_scan(uint64_t* c, uint32_t n, const uint64_t* bsum, uint64_t* total, uint32_t* cfirst,
            uint32_t* cmap, uint64_t cmap_cap) {
  __shared__ uint64_t s_c[JSCAN];
  __shared__ uint64_t s_w[2][17];
  __shared__ uint64_t s_carry[2];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
  const uint32_t base = blockIdx.x * JSCAN;
  const uint32_t m = (n - base) < JSCAN ? (n - base) : JSCAN;
  for (uint32_t i = tid; i < m; i += 1024) s_c[i] = c[base + i];
  if (wid == 0) {
    // the totals of the rounds before this one
    uint64_t cp = 0, cc = 0;
    for (uint32_t r = lane; r < blockIdx.x; r += 64) {
      cp += bsum[2 * r];
      cc += bsum[2 * r + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      cp += __shfl_xor(cp, off, 64);
      cc += __shfl_xor(cc, off, 64);
    }
    if (lane == 0) {
      s_carry[0] = cp;
      s_carry[1] = cc;
    }
  }
  __syncthreads();
  uint64_t v[8], sp = 0, sc = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t i = tid * 8 + j;
    v[j] = i < m ? s_c[i] : 0ull;
    sp += v[j];
    sc += (v[j] + JFCH - 1) / JFCH;
  }
  const uint64_t ip = wave_incl_scan64(sp), ic = wave_incl_scan64(sc);
  if (lane == 63) {
    s_w[0][wid] = ip;
    s_w[1][wid] = ic;
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t ap = 0, ac = 0;
    for (int w = 0; w < 16; ++w) {
      const uint64_t tp = s_w[0][w], tc = s_w[1][w];
      s_w[0][w] = ap;
      s_w[1][w] = ac;
      ap += tp;
      ac += tc;
    }
    s_w[0][16] = ap;
    s_w[1][16] = ac;
  }
  __syncthreads();
  const uint64_t carry_p = s_carry[0], carry_c = s_carry[1];
  uint64_t ep = carry_p + s_w[0][wid] + ip - sp, ec = carry_c + s_w[1][wid] + ic - sc;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t i = tid * 8 + j;
    if (i < m) {
      c[base + i] = ep;
      cfirst[base + i] = (uint32_t)ec;
      const uint64_t nch = (v[j] + JFCH - 1) / JFCH;
      for (uint64_t q = 0; q < nch; ++q)
        if (ec + q < cmap_cap) cmap[ec + q] = base + i;
      ep += v[j];
      ec += nch;
    }
  }
  if (tid == 0 && blockIdx.x == gridDim.x - 1) {
    total[0] = carry_p + s_w[0][16];
    total[1] = carry_c + s_w[1][16];
  }
}

// What the fill writes for one side of every pair: the source row, and
// payload columns read at the source row or at the sorted position, either
// from separate columns or as fields of one packed record per sorted
// position (a payload the sort carried).
struct JoinSideOut {
  const uint32_t* perm;   // source row of each sorted position (null: the position)
  uint32_t* out_rows;
  int ncols;
  int by_row;
  int rec_bytes;          // 0, 4 or 8
  const void* rec;
  const void* src[JFX_JOIN_MAXCOLS];
  int bytes[JFX_JOIN_MAXCOLS];
  int off[JFX_JOIN_MAXCOLS];
  void* out[JFX_JOIN_MAXCOLS];
};

// Layout of a side's payload, chosen once per launch: the cases the inner
// join produces get their own straight-line stores.
enum JoinPayMode : int { JP_NONE = 0, JP_REC8_2X4 = 1, JP_REC8_1X8 = 2, JP_REC4 = 3, JP_COLS = 4 };

__host__ inline int join_pay_mode(const JoinSideOut& s) {
  if (s.ncols == 0) return JP_NONE;
  if (s.rec_bytes == 8 && s.ncols == 2 && s.bytes[0] == 4 && s.bytes[1] == 4 && s.off[0] == 0 &&
      s.off[1] == 4)
    return JP_REC8_2X4;
  if (s.rec_bytes == 8 && s.ncols == 1 && s.bytes[0] == 8) return JP_REC8_1X8;
  if (s.rec_bytes == 4 && s.ncols == 1) return JP_REC4;
  return JP_COLS;
}

// Global-memory pointers: the fill's loads and stores through them are
// global (not flat) instructions, with a scalar base and a 32-bit offset
// when the base is wave-uniform.
#define JFX_GLOBAL __attribute__((address_space(1)))

// A wave-uniform pointer held in scalar registers, as a global pointer.
template <typename T>
__device__ __forceinline__ JFX_GLOBAL T* uniform_ptr(const void* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (JFX_GLOBAL T*)(((uint64_t)hi << 32) | lo);
}

// Element idx of a global array, the byte offset computed in 32 bits (the
// caller's arrays are below 4 GB from the base): a scalar base plus a 32-bit
// vector offset, one address form of the global instructions.
template <typename T>
__device__ __forceinline__ JFX_GLOBAL T& gat(JFX_GLOBAL T* base, uint32_t idx) {
  return *(JFX_GLOBAL T*)((JFX_GLOBAL char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const JFX_GLOBAL T& gat(JFX_GLOBAL const T* base, uint32_t idx) {
  return *(JFX_GLOBAL const T*)((JFX_GLOBAL const char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}

// One side's pointers for one chunk: sources offset so that a side's
// sorted-position index and the chunk's output index are 32-bit offsets from
// wave-uniform bases.
struct JoinChunkSide {
  JFX_GLOBAL const uint32_t* perm;
  JFX_GLOBAL uint32_t* out_rows;
  JFX_GLOBAL const char* rec;
  JFX_GLOBAL char* out0;
  JFX_GLOBAL char* out1;
  uint64_t out_base;
};

__device__ __forceinline__ JoinChunkSide join_chunk_side(const JoinSideOut& s, uint64_t pos0,
                                                         uint64_t out0) {
  JoinChunkSide c;
  c.perm = uniform_ptr<const uint32_t>(s.perm ? s.perm + pos0 : nullptr);
  c.out_rows = uniform_ptr<uint32_t>(s.out_rows ? s.out_rows + out0 : nullptr);
  c.rec = uniform_ptr<const char>(
      s.rec_bytes ? static_cast<const char*>(s.rec) + pos0 * s.rec_bytes : nullptr);
  c.out0 = uniform_ptr<char>(s.ncols > 0 ? static_cast<char*>(s.out[0]) + out0 * s.bytes[0] : nullptr);
  c.out1 = uniform_ptr<char>(s.ncols > 1 ? static_cast<char*>(s.out[1]) + out0 * s.bytes[1] : nullptr);
  c.out_base = out0;
  return c;
}

// Writes output q of the chunk for sorted position pos0 + p of this side.
__device__ __forceinline__ void join_emit_side(const JoinSideOut& s, const JoinChunkSide& c,
                                               int mode, uint32_t q, uint32_t p, uint64_t pos0) {
  const bool need_row = c.out_rows || (mode == JP_COLS && s.by_row);
  const uint32_t row = (need_row && c.perm) ? gat(c.perm, p) : (uint32_t)(pos0 + p);
  if (c.out_rows) gat(c.out_rows, q) = row;
  switch (mode) {
    case JP_NONE:
      return;
    case JP_REC8_2X4: {
      const uint64_t r = gat(reinterpret_cast<JFX_GLOBAL const uint64_t*>(c.rec), p);
      gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out0), q) = (uint32_t)r;
      gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out1), q) = (uint32_t)(r >> 32);
      return;
    }
    case JP_REC8_1X8:
      gat(reinterpret_cast<JFX_GLOBAL uint64_t*>(c.out0), q) =
          gat(reinterpret_cast<JFX_GLOBAL const uint64_t*>(c.rec), p);
      return;
    case JP_REC4:
      gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out0), q) =
          gat(reinterpret_cast<JFX_GLOBAL const uint32_t*>(c.rec), p);
      return;
    default:
      break;
  }
  // any other layout: per column, from the kernel arguments
  const uint64_t k = c.out_base + q;
  if (s.rec_bytes == 8) {
    const uint64_t r = reinterpret_cast<JFX_GLOBAL const uint64_t*>(c.rec)[p];
    for (int j = 0; j < s.ncols; ++j) {
      if (s.bytes[j] == 8) static_cast<uint64_t*>(s.out[j])[k] = r;
      else static_cast<uint32_t*>(s.out[j])[k] = (uint32_t)(r >> (8 * s.off[j]));
    }
    return;
  }
  const uint64_t idx = s.by_row ? (uint64_t)row : pos0 + p;
  for (int j = 0; j < s.ncols; ++j) {
    if (s.bytes[j] == 8)
      static_cast<uint64_t*>(s.out[j])[k] = static_cast<const uint64_t*>(s.src[j])[idx];
    else
      static_cast<uint32_t*>(s.out[j])[k] = static_cast<const uint32_t*>(s.src[j])[idx];
  }
}

// The record of sorted position p of a side in a record mode (JP_NONE: 0).
__device__ __forceinline__ uint64_t join_load_rec(const JoinChunkSide& c, int mode, uint32_t p) {
  if (mode == JP_REC8_2X4 || mode == JP_REC8_1X8)
    return gat(reinterpret_cast<JFX_GLOBAL const uint64_t*>(c.rec), p);
  if (mode == JP_REC4) return gat(reinterpret_cast<JFX_GLOBAL const uint32_t*>(c.rec), p);
  return 0;
}

// Non-temporal stores: the outputs are written once and not read again here
// (1.58-1.60 ms against 1.72 ms for the fill at 1.6e8 pairs; non-temporal
// record loads on top gave nothing more).
__device__ __forceinline__ void join_store_rec(const JoinChunkSide& c, int mode, uint32_t q,
                                               uint64_t r) {
  if (mode == JP_REC8_2X4) {
    __builtin_nontemporal_store((uint32_t)r, &gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out0), q));
    __builtin_nontemporal_store((uint32_t)(r >> 32), &gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out1), q));
  } else if (mode == JP_REC8_1X8) {
    __builtin_nontemporal_store(r, &gat(reinterpret_cast<JFX_GLOBAL uint64_t*>(c.out0), q));
  } else if (mode == JP_REC4) {
    __builtin_nontemporal_store((uint32_t)r, &gat(reinterpret_cast<JFX_GLOBAL uint32_t*>(c.out0), q));
  }
}

constexpr int JFU = 8;  // outputs per thread per round of the batched fill loop (4: 1.41 ms, 8: 1.38 ms, 16: 1.53 ms at 1.6e8 pairs, one workgroup per left tile)

// Persistent over the chunks of k_join_scan: chunk c of tile t covers the
// tile's outputs [(c - cfirst[t]) * JFCH, +JFCH). Per chunk the tile's
// bounds and in-tile scan are rebuilt (window staged in LDS), then output
// position k of the chunk finds its left row by binary search over the scan,
// so a row with many matches is spread over the workgroup.
template <typename K>
__global__ void __launch_bounds__(JB)
k_join_fill(const K* sl, uint64_t L, const K* sr, int oc, const uint64_t* wlo, const uint64_t* whi,
            const uint64_t* tile_off, const uint32_t* cfirst, const uint32_t* cmap,
            uint64_t cmap_cap, const uint64_t* total, K* out_key, JoinSideOut left,
            JoinSideOut right, int lmode, int rmode) {
  __shared__ K s_w[JWMAX];
  __shared__ uint32_t s_incl[JT];   // exclusive scan of the tile's counts
  __shared__ uint32_t s_lo[JT];
  __shared__ uint32_t s_tmp[JB / 64 + 1];
  // tile row of each output of the chunk, in the window's LDS (the window is
  // read only by tile_bounds, before the scan's barriers)
  static_assert(JWMAX * sizeof(K) >= JFCH * sizeof(uint16_t), "row map fits the window");
  uint16_t* const s_row = reinterpret_cast<uint16_t*>(s_w);
  const uint64_t nchunks = total[1] < cmap_cap ? total[1] : cmap_cap, npairs = total[0];
  // record payloads (or none) on both sides and no row outputs: the batched
  // output loop
  const bool fast = lmode != JP_COLS && rmode != JP_COLS && !left.out_rows && !right.out_rows;
  const uint32_t ntiles = (uint32_t)((L + JT - 1) / JT);
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const uint32_t t = cmap[c];
    const uint64_t tb = (uint64_t)t * JT;
    const uint64_t w0 = wlo[t], w1 = whi[t];
    K raw[JIPT];
    tile_rows(sl, L, tb, raw);
    const bool staged = stage_window(sr, oc, w0, w1, s_w);
    __syncthreads();
    uint64_t lo[JIPT];
    uint32_t cnt[JIPT];
    tile_bounds(raw, L, sr, oc, tb, w0, w1, s_w, staged, lo, cnt);
#pragma unroll
    for (int j = 0; j < JIPT; ++j) {
      s_incl[threadIdx.x * JIPT + j] = cnt[j];
      s_lo[threadIdx.x * JIPT + j] = (uint32_t)lo[j];
    }
    __syncthreads();
    const uint32_t tot = block_excl_scan_lds<JT, JB>(s_incl, s_tmp);  // now exclusive
    const uint64_t base = tile_off[t];
    const uint64_t tend = (t + 1 < ntiles ? tile_off[t + 1] : npairs) - base;
    const uint32_t k0 = (uint32_t)((c - cfirst[t]) * JFCH);
    const uint32_t k1 = (uint64_t)k0 + JFCH < tend ? k0 + JFCH : (uint32_t)tend;
    (void)tot;
    // each row marks its outputs inside [k0, k1) with its index, so an
    // output finds its row with one LDS read
#pragma unroll
    for (int j = 0; j < JIPT; ++j) {
      const uint32_t r = threadIdx.x * JIPT + j;
      const uint32_t a0 = s_incl[r], a1 = a0 + cnt[j];
      const uint32_t b0 = a0 > k0 ? a0 : k0, b1 = a1 < k1 ? a1 : k1;
      for (uint32_t q = b0; q < b1; ++q) s_row[q - k0] = (uint16_t)r;
    }
    __syncthreads();
    // wave-uniform bases: the chunk's outputs, the tile's left rows and the
    // tile's right window; the indices below are 32-bit offsets from them
    const uint64_t out0 = base + k0;
    const JoinChunkSide lc = join_chunk_side(left, tb, out0);
    const JoinChunkSide rc = join_chunk_side(right, w0, out0);
    JFX_GLOBAL const K* slt = uniform_ptr<const K>(sl + tb);
    JFX_GLOBAL K* okey = uniform_ptr<K>(out_key ? out_key + out0 : nullptr);
    const uint32_t w0l = (uint32_t)w0;
    const uint32_t nout = k1 - k0;
    if (fast) {
      // JFU outputs per thread per round: every load of the round is issued
      // before its stores (one memory round trip per JFU outputs)
      for (uint32_t q0 = threadIdx.x; q0 < nout; q0 += JB * JFU) {
        uint32_t a[JFU], rp[JFU];
        K key[JFU];
        uint64_t lr[JFU], rr[JFU];
#pragma unroll
        for (int u = 0; u < JFU; ++u) {
          const uint32_t q = q0 + u * JB;
          const uint32_t qc = q < nout ? q : nout - 1;
          a[u] = s_row[qc];
          rp[u] = s_lo[a[u]] - w0l + (k0 + qc - s_incl[a[u]]);
        }
#pragma unroll
        for (int u = 0; u < JFU; ++u) {
          key[u] = okey ? gat(slt, a[u]) : K(0);
          lr[u] = join_load_rec(lc, lmode, a[u]);
          rr[u] = join_load_rec(rc, rmode, rp[u]);
        }
#pragma unroll
        for (int u = 0; u < JFU; ++u) {
          const uint32_t q = q0 + u * JB;
          if (q < nout) {
            if (okey) __builtin_nontemporal_store(key[u], &gat(okey, q));
            join_store_rec(lc, lmode, q, lr[u]);
            join_store_rec(rc, rmode, q, rr[u]);
          }
        }
      }
    } else {
      for (uint32_t q = threadIdx.x; q < nout; q += JB) {
        const uint32_t a = s_row[q];
        const uint32_t rp = s_lo[a] - w0l + (k0 + q - s_incl[a]);
        if (okey) gat(okey, q) = gat(slt, a);
        join_emit_side(left, lc, lmode, q, a, tb);
        join_emit_side(right, rc, rmode, q, rp, w0);
      }
    }
    __syncthreads();   // LDS is rebuilt for the next chunk
  }
}

// Two 4-byte columns into 8-byte records, four rows per thread with 16-byte
// loads and stores (n4 = rows / 4; the rows beyond 4 * n4 go to k_pack_cols).
__global__ void __launch_bounds__(256)
k_pack_2x4(const uint4* lo, const uint4* hi, uint64_t n4, uint4* out) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride) {
    const uint4 a = lo[i], b = hi[i];
    out[2 * i] = make_uint4(a.x, b.x, a.y, b.y);
    out[2 * i + 1] = make_uint4(a.z, b.z, a.w, b.w);
  }
}

// Packs ncols columns (4 or 8 bytes each, at most 8 bytes together) into one
// record per row: column j at byte offset off[j]; rows [first, n).
__global__ void __launch_bounds__(256)
k_pack_cols(JoinSideOut p, uint64_t first, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = first + (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < JFX_JOIN_MAXCOLS; ++j) {
      if (j >= p.ncols) break;
      const uint64_t v = p.bytes[j] == 8 ? static_cast<const uint64_t*>(p.src[j])[i]
                                         : (uint64_t)static_cast<const uint32_t*>(p.src[j])[i];
      r |= v << (8 * p.off[j]);
    }
    if (p.rec_bytes == 8) static_cast<uint64_t*>(p.out[0])[i] = r;
    else static_cast<uint32_t*>(p.out[0])[i] = (uint32_t)r;
  }
}

// ---- inner-join range pre-filter ------------------------------------------
// Chunks of whole 4096-row tiles (at most PF_MAXCHUNKS); within a tile each
// wave owns 1024 consecutive rows, walked 64 at a time, so ballot order is
// row order.
constexpr int PF_BLOCK = 256;
constexpr int PF_W = PF_BLOCK / 64;
constexpr int PF_IPT = 16;
constexpr uint32_t PF_TILE = PF_BLOCK * PF_IPT;
constexpr uint32_t PF_MAXCHUNKS = 1024;

__device__ __forceinline__ uint64_t pf_row(uint64_t t, int w, int j, uint32_t lane) {
  return t + (uint64_t)w * 64u * PF_IPT + (uint64_t)j * 64u + lane;
}

// Row r's payload columns packed into one record (column c at byte offset
// off[c], as k_pack_cols).
__device__ __forceinline__ uint64_t pf_record(const JoinSideOut& p, bool two4, uint64_t r) {
  if (two4)
    return (uint64_t)static_cast<const uint32_t*>(p.src[0])[r] |
           ((uint64_t)static_cast<const uint32_t*>(p.src[1])[r] << 32);
  uint64_t rec = 0;
#pragma unroll
  for (int c = 0; c < JFX_JOIN_MAXCOLS; ++c) {
    if (c >= p.ncols) break;
    const uint64_t v = p.bytes[c] == 8 ? static_cast<const uint64_t*>(p.src[c])[r]
                                       : (uint64_t)static_cast<const uint32_t*>(p.src[c])[r];
    rec |= v << (8 * p.off[c]);
  }
  return rec;
}

__device__ __forceinline__ void pf_store_record(const JoinSideOut& p, uint32_t q, uint64_t rec) {
  if (p.rec_bytes == 8) static_cast<uint64_t*>(p.out[0])[q] = rec;
  else static_cast<uint32_t*>(p.out[0])[q] = (uint32_t)rec;
}

// Keys compare in their own type (signed or unsigned integers): the other
// side's first and last sorted keys bound every key that can match.
template <typename K>
__global__ void __launch_bounds__(PF_BLOCK)
k_pf_count(const K* keys, uint64_t n, uint64_t chunk, const K* other, uint64_t n_other,
           uint32_t* counts) {
  __shared__ uint32_t s_w[PF_W];
  const K lo = other[0], hi = other[n_other - 1];
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  uint32_t cnt = 0;
  for (uint64_t t = cbeg; t < cend; t += PF_TILE) {
    K kv[PF_IPT];
#pragma unroll
    for (int j = 0; j < PF_IPT; ++j) {
      const uint64_t r = pf_row(t, wid, j, lane);
      kv[j] = r < cend ? keys[r] : K(0);
    }
#pragma unroll
    for (int j = 0; j < PF_IPT; ++j) {
      const uint64_t r = pf_row(t, wid, j, lane);
      cnt += (uint32_t)__popcll(ballot64(r < cend && lo <= kv[j] && kv[j] <= hi));
    }
  }
  if (lane == 0) s_w[wid] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (int w = 0; w < PF_W; ++w) c += s_w[w];
    counts[blockIdx.x] = c;
  }
}

// Exclusive scan of the chunk counts in place; the kept total to *total.
__global__ void __launch_bounds__(1024) k_pf_scan(uint32_t* counts, uint32_t nchunks,
                                                  uint64_t* total) {
  __shared__ uint32_t s[PF_MAXCHUNKS];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  for (uint32_t i = threadIdx.x; i < PF_MAXCHUNKS; i += 1024) s[i] = i < nchunks ? counts[i] : 0u;
  __syncthreads();
  const uint32_t tot = block_excl_scan_lds<PF_MAXCHUNKS, 1024>(s, s_tmp);
  for (uint32_t i = threadIdx.x; i < nchunks; i += 1024) counts[i] = s[i];
  if (threadIdx.x == 0) *total = tot;
}

// Kept row i at output q: its key, its payload columns packed into one
// record (column j at byte offset off[j], as k_pack_cols) and, when
// out_rows is set, i itself.
template <typename K>
__global__ void __launch_bounds__(PF_BLOCK)
k_pf_scatter(const K* keys, uint64_t n, uint64_t chunk, const K* other, uint64_t n_other,
             const uint32_t* offsets, K* out_keys, JoinSideOut p, uint32_t* out_rows) {
  __shared__ uint32_t s_w[PF_W + 1];
  const K lo = other[0], hi = other[n_other - 1];
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  const bool two4 = p.rec_bytes == 8 && p.ncols == 2 && p.bytes[0] == 4 && p.bytes[1] == 4;
  uint32_t run = offsets[blockIdx.x];
  for (uint64_t t = cbeg; t < cend; t += PF_TILE) {
    K kv[PF_IPT];
    uint64_t m[PF_IPT];
    uint32_t wcnt = 0;
#pragma unroll
    for (int j = 0; j < PF_IPT; ++j) {
      const uint64_t r = pf_row(t, wid, j, lane);
      kv[j] = r < cend ? keys[r] : K(0);
    }
#pragma unroll
    for (int j = 0; j < PF_IPT; ++j) {
      const uint64_t r = pf_row(t, wid, j, lane);
      m[j] = ballot64(r < cend && lo <= kv[j] && kv[j] <= hi);
      wcnt += (uint32_t)__popcll(m[j]);
    }
    if (lane == 0) s_w[wid] = wcnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t c = 0;
      for (int w = 0; w < PF_W; ++w) {
        const uint32_t x = s_w[w];
        s_w[w] = c;
        c += x;
      }
      s_w[PF_W] = c;
    }
    __syncthreads();
    uint32_t pos = run + s_w[wid];
#pragma unroll
    for (int j = 0; j < PF_IPT; ++j) {
      if ((m[j] >> lane) & 1ull) {
        const uint64_t r = pf_row(t, wid, j, lane);
        const uint32_t q = pos + mbcnt(m[j]);
        out_keys[q] = kv[j];
        if (out_rows) out_rows[q] = (uint32_t)r;
        if (p.ncols) pf_store_record(p, q, pf_record(p, two4, r));
      }
      pos += (uint32_t)__popcll(m[j]);
    }
    run += s_w[PF_W];
    __syncthreads();
  }
}

// One pass of the pre-filter for n < 2^30 (the lookback's 30-bit counts):
// tile ids from an atomic counter, so every earlier tile is running; each
// tile counts its kept rows, takes the kept rows of the tiles before it by a
// decoupled lookback over one counter per tile, and writes its kept rows
// there. The keys are read once (the three-kernel form reads them twice).
// The last tile writes the kept total. Tiles of 16384 rows (16 waves): at
// 1e8 int64 keys 4096-row tiles took 779 us, 8192-row tiles 779 us and
// 16384-row tiles 686 us. (Loading the kept rows' payloads before the
// lookback instead of after it measured 816 against 779 us with 4096-row
// tiles.)
constexpr int PF1_BLOCK = 1024;
constexpr int PF1_W = PF1_BLOCK / 64;
constexpr uint32_t PF1_TILE = PF1_BLOCK * PF_IPT;

template <typename K>
__global__ void __launch_bounds__(PF1_BLOCK)
k_pf_onepass(const K* keys, uint64_t n, const K* other, uint64_t n_other, uint32_t* desc,
             uint32_t* counter, uint32_t ntiles, uint64_t* total, uint32_t* err, K* out_keys,
             JoinSideOut p, uint32_t* out_rows) {
  constexpr int PF_W = PF1_W;
  constexpr uint32_t PF_TILE = PF1_TILE;
  __shared__ uint32_t s_w[PF_W + 1];
  __shared__ uint32_t s_tile, s_excl;
  if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
  const K lo = other[0], hi = other[n_other - 1];
  const int wid = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const bool two4 = p.rec_bytes == 8 && p.ncols == 2 && p.bytes[0] == 4 && p.bytes[1] == 4;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t t = (uint64_t)tile * PF_TILE;
  const uint64_t tend = (t + PF_TILE) < n ? (t + PF_TILE) : n;
  K kv[PF_IPT];
  uint64_t m[PF_IPT];
  uint32_t wcnt = 0;
#pragma unroll
  for (int j = 0; j < PF_IPT; ++j) {
    const uint64_t r = pf_row(t, wid, j, lane);
    kv[j] = r < tend ? keys[r] : K(0);
  }
#pragma unroll
  for (int j = 0; j < PF_IPT; ++j) {
    const uint64_t r = pf_row(t, wid, j, lane);
    m[j] = ballot64(r < tend && lo <= kv[j] && kv[j] <= hi);
    wcnt += (uint32_t)__popcll(m[j]);
  }
  if (lane == 0) s_w[wid] = wcnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (int w = 0; w < PF_W; ++w) {
      const uint32_t x = s_w[w];
      s_w[w] = c;
      c += x;
    }
    s_w[PF_W] = c;
    lookback_publish(desc, tile, 1u, 0u, c);
  }
  __syncthreads();
  if (wid == 0) {
    const uint32_t ex = wave_lookback_u32(desc, tile, s_w[PF_W], err);
    if (lane == 0) {
      s_excl = ex;
      if (tile == ntiles - 1) *total = (uint64_t)ex + s_w[PF_W];
    }
  }
  __syncthreads();
  uint32_t pos = s_excl + s_w[wid];
#pragma unroll
  for (int j = 0; j < PF_IPT; ++j) {
    if ((m[j] >> lane) & 1ull) {
      const uint64_t r = pf_row(t, wid, j, lane);
      const uint32_t q = pos + mbcnt(m[j]);
      out_keys[q] = kv[j];
      if (out_rows) out_rows[q] = (uint32_t)r;
      if (p.ncols) pf_store_record(p, q, pf_record(p, two4, r));
    }
    pos += (uint32_t)__popcll(m[j]);
  }
}

// Clears the pre-filter's lookback descriptors and tile counter.
__global__ void __launch_bounds__(256) k_pf_zero(uint32_t* desc, uint32_t ntiles, uint32_t* counter) {
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < ntiles; i += stride) desc[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
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
  uint64_t* tcnt;    // tile pair counts, then their exclusive offsets
  uint64_t* total;   // [0] pairs, [1] fill chunks
  uint32_t* cfirst;  // first fill chunk of each tile
  uint32_t* cmap;    // tile of each fill chunk
  uint64_t cmap_cap;
  uint64_t* bsum;    // [2 x rounds] pair and chunk totals per scan round
};

JoinWs carve_join(Carver& c, uint64_t L) {
  JoinWs w{};
  const uint64_t t = jtiles(L);
  w.wlo = c.take<uint64_t>(t);
  w.whi = c.take<uint64_t>(t);
  w.tcnt = c.take<uint64_t>(t);
  w.total = c.take<uint64_t>(2);
  w.cfirst = c.take<uint32_t>(t);
  w.cmap_cap = t + JCMAP_EXTRA;
  w.cmap = c.take<uint32_t>(w.cmap_cap);
  w.bsum = c.take<uint64_t>(2 * ((t + JSCAN - 1) / JSCAN + 1));
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
  const uint32_t rounds = (nt + JSCAN - 1) / JSCAN;
  JFX_LAUNCH("join_scan_sums", s, (k_join_scan_sums<<<rounds, 1024, 0, s>>>(w.tcnt, nt, w.bsum)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("join_scan", s, (k_join_scan<<<rounds, 1024, 0, s>>>(w.tcnt, nt, w.bsum, w.total,
                                                                  w.cfirst, w.cmap, w.cmap_cap)));
  JFX_HIP(hipGetLastError());
  const int rc = read_back(s, w.total, sizeof(uint64_t), out_m);
  if (rc != JFX_OK) return rc;
  JFX_CHECK(*out_m < (1ull << 32), JFX_ERR_UNSUPPORTED,
            "the join produces 2^32 pairs or more; at most 2^32 - 1 are supported");
  return JFX_OK;
}

// Workgroups of the persistent fill: enough to fill every CU several times
// over, never more than the tiles' worst-case chunk count.
// Workgroups of the persistent fill: about one per left tile, between 2048
// and 32768. At 1e8 x 1e7 (39,063 tiles, 1.6e8 pairs) a grid of 2048 took
// 1448 us, 8192 1394 us, 32768 1374 us and 131072 1398 us: the output
// stores, most of the fill's bytes, stream best from many short-lived
// workgroups (a plain write stream runs at 4.6 TB/s from 2048 workgroups and
// 5.9 TB/s from 16384).
uint32_t fill_grid(uint64_t L) {
  const uint64_t bound = jtiles(L) + JCMAP_EXTRA;
  const uint64_t want = std::min<uint64_t>(32768u, std::max<uint64_t>(2048u, jtiles(L)));
  return (uint32_t)std::min<uint64_t>(bound, want);
}

template <typename K>
int join_fill_typed(int oc, const void* sl, uint64_t L, const void* sr, const JoinWs& w,
                    void* out_key, const JoinSideOut& left, const JoinSideOut& right,
                    hipStream_t s) {
  JFX_LAUNCH("join_fill", s, (k_join_fill<K><<<fill_grid(L), JB, 0, s>>>(
                                 (const K*)sl, L, (const K*)sr, oc, w.wlo, w.whi, w.tcnt, w.cfirst,
                                 w.cmap, w.cmap_cap, w.total, (K*)out_key, left, right,
                                 join_pay_mode(left), join_pay_mode(right))));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int join_fill_dispatch(int key_type, const void* sl, uint64_t L, const void* sr,
                       const JoinWs& w, void* out_key, const JoinSideOut& left,
                       const JoinSideOut& right, hipStream_t s) {
  const int oc = order_code(key_type);
  if (key_bytes(key_type) == 4)
    return join_fill_typed<uint32_t>(oc, sl, L, sr, w, out_key, left, right, s);
  return join_fill_typed<uint64_t>(oc, sl, L, sr, w, out_key, left, right, s);
}

// One side of jfx_join_fill_columns in the kernel's form, checked.
int side_from_abi(const jfx_join_side* a, JoinSideOut& o) {
  o = JoinSideOut{};
  if (!a) return JFX_OK;
  JFX_CHECK(a->ncols >= 0 && a->ncols <= JFX_JOIN_MAXCOLS, JFX_ERR_INVALID,
            "a join side has 0 to 8 payload columns");
  JFX_CHECK(a->record_bytes == 0 || a->record_bytes == 4 || a->record_bytes == 8,
            JFX_ERR_UNSUPPORTED, "record_bytes must be 0, 4 or 8");
  o.perm = a->rows;
  o.out_rows = a->out_rows;
  o.ncols = a->ncols;
  o.by_row = a->by_row;
  o.rec_bytes = a->record_bytes;
  o.rec = a->records;
  JFX_CHECK(!o.rec_bytes || o.rec, JFX_ERR_INVALID, "record_bytes without records");
  JFX_CHECK(!o.rec_bytes || !o.by_row, JFX_ERR_INVALID, "records are read by sorted position");
  JFX_CHECK(o.rec_bytes != 4 || o.ncols <= 1, JFX_ERR_INVALID, "a 4-byte record holds one column");
  for (int j = 0; j < o.ncols; ++j) {
    JFX_CHECK(a->col_bytes[j] == 4 || a->col_bytes[j] == 8, JFX_ERR_UNSUPPORTED,
              "payload columns must be 4 or 8 bytes");
    JFX_CHECK(a->out_cols[j] != nullptr, JFX_ERR_INVALID, "null payload output column");
    JFX_CHECK(o.rec_bytes || a->cols[j] != nullptr, JFX_ERR_INVALID, "null payload column");
    JFX_CHECK(!o.rec_bytes || (a->col_offset[j] >= 0 &&
                               a->col_offset[j] + a->col_bytes[j] <= o.rec_bytes),
              JFX_ERR_INVALID, "a column lies outside its record");
    o.src[j] = a->cols[j];
    o.bytes[j] = a->col_bytes[j];
    o.off[j] = a->col_offset[j];
    o.out[j] = a->out_cols[j];
  }
  return JFX_OK;
}

// Below 2^30 rows the one-pass kernel (lookback descriptors, one per
// tile); above it, or with JFX_PREFILTER=chunked, the three-kernel form
// (chunk counts, their scan, scatter).
constexpr uint64_t PF_ONEPASS_MAX = 1ull << 30;

struct PfWs {
  uint32_t* counts;   // [PF_MAXCHUNKS] chunk counts (three-kernel form)
  uint64_t* total;
  uint32_t* counter;
  uint32_t* desc;     // [tiles] lookback descriptors (one-pass form)
};

PfWs pf_carve(Carver& cv, uint64_t n) {
  PfWs w;
  w.counts = cv.take<uint32_t>(PF_MAXCHUNKS);
  w.total = cv.take<uint64_t>(1);
  w.counter = cv.take<uint32_t>(64);
  w.desc = cv.take<uint32_t>(n < PF_ONEPASS_MAX ? (n + PF1_TILE - 1) / PF1_TILE : 0);
  return w;
}

template <typename K>
int prefilter_typed(const void* keys, uint64_t n, const void* other, uint64_t n_other,
                    const JoinSideOut& p, void* out_keys, uint32_t* out_rows, uint64_t* out_count,
                    void* ws, uint64_t ws_bytes, hipStream_t s) {
  Carver cv(ws, ws_bytes);
  const PfWs w = pf_carve(cv, n);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "pre-filter workspace too small");
  uint32_t* counts = w.counts;
  uint64_t* total = w.total;
  const uint64_t tiles = (n + PF_TILE - 1) / PF_TILE;
  const char* form = getenv("JFX_PREFILTER");
  if (n < PF_ONEPASS_MAX && !(form && strcmp(form, "chunked") == 0)) {
    uint32_t* err = device_error_word();
    JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
    const uint32_t nt = (uint32_t)((n + PF1_TILE - 1) / PF1_TILE);
    JFX_LAUNCH("pf_zero", s, (k_pf_zero<<<std::min<uint32_t>(256, (nt + 255) / 256), 256, 0, s>>>(
                                 w.desc, nt, w.counter)));
    JFX_LAUNCH("pf_onepass", s, (k_pf_onepass<K><<<nt, PF1_BLOCK, 0, s>>>(
                                     static_cast<const K*>(keys), n, static_cast<const K*>(other),
                                     n_other, w.desc, w.counter, nt, total, err,
                                     static_cast<K*>(out_keys), p, out_rows)));
    JFX_HIP(hipGetLastError());
    return read_back(s, total, sizeof(uint64_t), out_count);
  }
  const uint64_t nch = std::max<uint64_t>(1, std::min<uint64_t>(PF_MAXCHUNKS, tiles));
  const uint64_t chunk = (tiles + nch - 1) / nch * PF_TILE;
  const uint32_t nchunks = (uint32_t)((n + chunk - 1) / chunk);
  const K* k = static_cast<const K*>(keys);
  const K* o = static_cast<const K*>(other);
  JFX_LAUNCH("pf_count", s, (k_pf_count<K><<<nchunks, PF_BLOCK, 0, s>>>(k, n, chunk, o, n_other,
                                                                          counts)));
  JFX_LAUNCH("pf_scan", s, (k_pf_scan<<<1, 1024, 0, s>>>(counts, nchunks, total)));
  JFX_LAUNCH("pf_scatter", s, (k_pf_scatter<K><<<nchunks, PF_BLOCK, 0, s>>>(
                                  k, n, chunk, o, n_other, counts, static_cast<K*>(out_keys), p,
                                  out_rows)));
  JFX_HIP(hipGetLastError());
  return read_back(s, total, sizeof(uint64_t), out_count);
}

}  // namespace

}  // namespace jfx

extern "C" {

int jfx_join_prefilter_workspace_size(uint64_t n, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  jfx::Carver c(nullptr, 0);
  jfx::pf_carve(c, n);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_join_prefilter(int key_type, const void* keys, uint64_t n, const void* sorted_other,
                       uint64_t n_other, int ncols, const void* const* cols, const int* col_bytes,
                       void* out_keys, void* out_records, uint32_t* out_rows, uint64_t* out_count,
                       void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(out_count != nullptr, JFX_ERR_INVALID, "out_count is null");
  JFX_CHECK(key_type == JFX_U32 || key_type == JFX_I32 || key_type == JFX_U64 ||
                key_type == JFX_I64,
            JFX_ERR_UNSUPPORTED, "the range pre-filter takes integer keys");
  JFX_CHECK(n < (1ull << 32) && n_other < (1ull << 32), JFX_ERR_UNSUPPORTED,
            "each side must have fewer than 2^32 rows");
  JFX_CHECK(ncols >= 0 && ncols <= JFX_JOIN_MAXCOLS && (ncols == 0 || (cols && col_bytes)),
            JFX_ERR_INVALID, "0 to 8 payload columns");
  jfx::JoinSideOut p{};
  p.ncols = ncols;
  int off = 0;
  for (int j = 0; j < ncols; ++j) {
    JFX_CHECK(col_bytes[j] == 4 || col_bytes[j] == 8, JFX_ERR_UNSUPPORTED,
              "columns must be 4 or 8 bytes");
    JFX_CHECK(n == 0 || cols[j] != nullptr, JFX_ERR_INVALID, "null column");
    p.src[j] = cols[j];
    p.bytes[j] = col_bytes[j];
    p.off[j] = off;
    off += col_bytes[j];
  }
  JFX_CHECK(off <= 8, JFX_ERR_UNSUPPORTED, "packed records hold at most 8 bytes");
  p.rec_bytes = off;
  p.out[0] = out_records;
  if (n == 0 || n_other == 0) {
    *out_count = 0;
    return JFX_OK;
  }
  JFX_CHECK(keys && sorted_other && out_keys && workspace && (ncols == 0 || out_records),
            JFX_ERR_INVALID, "null argument");
  hipStream_t s = jfx::as_stream(stream);
  if (jfx::key_bytes(key_type) == 4) {
    if (key_type == JFX_I32)
      return jfx::prefilter_typed<int32_t>(keys, n, sorted_other, n_other, p, out_keys, out_rows,
                                           out_count, workspace, ws_bytes, s);
    return jfx::prefilter_typed<uint32_t>(keys, n, sorted_other, n_other, p, out_keys, out_rows,
                                          out_count, workspace, ws_bytes, s);
  }
  if (key_type == JFX_I64)
    return jfx::prefilter_typed<int64_t>(keys, n, sorted_other, n_other, p, out_keys, out_rows,
                                         out_count, workspace, ws_bytes, s);
  return jfx::prefilter_typed<uint64_t>(keys, n, sorted_other, n_other, p, out_keys, out_rows,
                                        out_count, workspace, ws_bytes, s);
}

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
  jfx::JoinSideOut l{}, r{};
  l.perm = left_rows;
  l.out_rows = out_left_rows;
  r.perm = right_rows;
  r.out_rows = out_right_rows;
  return jfx::join_fill_dispatch(key_type, sorted_left, n_left, sorted_right, w, out_keys, l, r,
                                 jfx::as_stream(stream));
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
  jfx::JoinSideOut l{}, r{};
  l.perm = left_rows;
  l.out_rows = out_left_rows;
  l.by_row = left_by_row;
  if (out_left_values) {
    l.ncols = 1;
    l.src[0] = left_values;
    l.bytes[0] = left_elem_bytes;
    l.out[0] = out_left_values;
  }
  r.perm = right_rows;
  r.out_rows = out_right_rows;
  r.by_row = right_by_row;
  if (out_right_values) {
    r.ncols = 1;
    r.src[0] = right_values;
    r.bytes[0] = right_elem_bytes;
    r.out[0] = out_right_values;
  }
  return jfx::join_fill_dispatch(key_type, sorted_left, n_left, sorted_right, w, out_keys, l, r,
                                 jfx::as_stream(stream));
}

int jfx_join_fill_columns(int key_type, const void* sorted_left, uint64_t n_left,
                          const jfx_join_side* left, const void* sorted_right, uint64_t n_right,
                          const jfx_join_side* right, void* out_keys, void* workspace,
                          uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  if (n_left == 0 || n_right == 0) return JFX_OK;
  JFX_CHECK(sorted_left && sorted_right && workspace, JFX_ERR_INVALID, "null argument");
  jfx::JoinSideOut l{}, r{};
  int rc = jfx::side_from_abi(left, l);
  if (rc != JFX_OK) return rc;
  rc = jfx::side_from_abi(right, r);
  if (rc != JFX_OK) return rc;
  jfx::Carver c(workspace, ws_bytes);
  jfx::JoinWs w = jfx::carve_join(c, n_left);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "join workspace too small");
  return jfx::join_fill_dispatch(key_type, sorted_left, n_left, sorted_right, w, out_keys, l, r,
                                 jfx::as_stream(stream));
}

int jfx_pack_columns(int ncols, const void* const* cols, const int* col_bytes, uint64_t n,
                     void* out, int* out_record_bytes, void* stream) {
  JFX_CHECK(ncols >= 1 && ncols <= JFX_JOIN_MAXCOLS && cols && col_bytes, JFX_ERR_INVALID,
            "1 to 8 columns");
  jfx::JoinSideOut p{};
  p.ncols = ncols;
  int off = 0;
  for (int j = 0; j < ncols; ++j) {
    JFX_CHECK(col_bytes[j] == 4 || col_bytes[j] == 8, JFX_ERR_UNSUPPORTED,
              "columns must be 4 or 8 bytes");
    JFX_CHECK(n == 0 || cols[j] != nullptr, JFX_ERR_INVALID, "null column");
    p.src[j] = cols[j];
    p.bytes[j] = col_bytes[j];
    p.off[j] = off;
    off += col_bytes[j];
  }
  JFX_CHECK(off <= 8, JFX_ERR_UNSUPPORTED, "packed records hold at most 8 bytes");
  p.rec_bytes = off;
  p.out[0] = out;
  if (out_record_bytes) *out_record_bytes = off;
  if (n == 0) return JFX_OK;
  JFX_CHECK(out != nullptr, JFX_ERR_INVALID, "null output");
  hipStream_t s = jfx::as_stream(stream);
  uint64_t first = 0;
  if (ncols == 2 && col_bytes[0] == 4 && col_bytes[1] == 4 && (uintptr_t)cols[0] % 16 == 0 &&
      (uintptr_t)cols[1] % 16 == 0 && (uintptr_t)out % 16 == 0) {
    const uint64_t n4 = n / 4;
    if (n4) {
      const uint32_t b4 = (uint32_t)std::min<uint64_t>(4096, (n4 + 255) / 256);
      JFX_LAUNCH("pack_cols", s, (jfx::k_pack_2x4<<<b4, 256, 0, s>>>(
                                     (const uint4*)cols[0], (const uint4*)cols[1], n4, (uint4*)out)));
      JFX_HIP(hipGetLastError());
    }
    first = n4 * 4;
  }
  if (first < n) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(8192, (n - first + 255) / 256);
    JFX_LAUNCH("pack_cols", s, (jfx::k_pack_cols<<<blocks, 256, 0, s>>>(p, first, n)));
    JFX_HIP(hipGetLastError());
  }
  return JFX_OK;
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
