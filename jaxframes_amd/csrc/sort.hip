// Stable LSD radix sort for 32/64-bit keys with 0/4/8-byte payloads.
//
// Restates ParallelRadixSort.sort (parallel_algorithms.py:463-581): the
// reference runs one stable argsort per 8-bit digit over all digits; here
// each varying digit is one scatter pass over HBM. Pipeline per call:
//   k_sort_stats     one read of the keys, one workgroup per chunk: min, max,
//                    AND and OR of the ordered keys, and each chunk's
//                    histogram of the low 11 bits; for jfx_sort with 8-bit
//                    digits also every layout digit's histogram
//   host             reads the four statistics (one stream sync)
//   max-min < 2048:  one chunked pass with digit = ordered(key) - min, its
//                    per-chunk counts being the chunk histograms rotated by min
//   otherwise:       k_sort_hist (histograms of the varying layout digits,
//                    unless the statistics pass made them), k_sort_plan
//                    (digit offsets), then one onesweep pass with decoupled
//                    lookback per varying digit. Digits on which all keys agree
//                    are skipped (the reference's skip test at :512 only
//                    skips leading zero digits and never fires for signed keys).
// Descending order is the reversed ascending order (:552-555), written by the
// last pass at n-1-pos.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "chunked.hpp"
#include "onesweep.hpp"
#include "sortctx.hpp"

namespace jfx {

constexpr int OS_TILE = 8192;  // onesweep: keys per workgroup (shape per key width: Passes)
constexpr int CH_W = 8;      // chunked scatter: waves per workgroup
constexpr int STATS_BLOCK = 1024;
constexpr int HIST_BLOCK = 1024;

// Chunks are whole multiples of this many keys; every scatter variant's tile
// divides it.
template <typename K> constexpr int ch_granule() { return sizeof(K) == 4 ? 8192 : 4096; }

// Chunk-scatter variant: 0 = plain, 1 = software-pipelined loads (JFX_CHUNK_CFG
// overrides the default for tuning runs).
int chunk_cfg() {
  static int cfg = [] {
    const char* e = getenv("JFX_CHUNK_CFG");
    return e ? atoi(e) : 0;
  }();
  return cfg;
}

__global__ void k_sort_init(unsigned long long* stats, uint32_t* counters, uint32_t ncounters,
                            uint32_t* hist, uint32_t nhist) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nhist; i += stride) hist[i] = 0u;
  if (blockIdx.x == 0) {
    if (threadIdx.x < ncounters) counters[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
      stats[0] = ~0ull;  // min
      stats[1] = 0ull;   // max
      stats[2] = ~0ull;  // and
      stats[3] = 0ull;   // or
    }
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename K>
struct Stats4 {
  K mn, mx, an, orr;
  __device__ __forceinline__ void init() {
    mn = ~K(0);
    mx = K(0);
    an = ~K(0);
    orr = K(0);
  }
  __device__ __forceinline__ void add(K o) {
    mn = o < mn ? o : mn;
    mx = o > mx ? o : mx;
    an &= o;
    orr |= o;
  }
};

// One workgroup per chunk: statistics and the chunk's low-11-bit histogram.
// With HIST it also adds the chunk's 8-bit digit histograms of every layout
// pass into g_hist[pass][256], so a multi-pass sort reads its keys once
// before the scatter passes instead of twice: digit 0 is folded from the
// low-11-bit histogram, the higher digits take one LDS atomic each (one per
// wave when the whole wave agrees on the digit). These atomics bound the
// pass, so each higher digit is counted in HREP interleaved copies: lane l
// adds bin b at b * HREP + l % HREP. The lanes of one copy (l, l + 16, l +
// 32, l + 48) then share one group of 4 LDS banks and the copies never share
// a bank, so a wave instruction meets at most 4-way bank conflicts (two on
// average) where one table per digit met 4-5 on uniform digits and up to
// 64-way same-address conflicts on skewed ones (keys of a narrow range, or
// Zipf keys, give a digit two or three values in a wave).
constexpr int HREP = 16;

template <typename K, bool HIST>
__global__ void __launch_bounds__(STATS_BLOCK)
k_sort_stats(const K* __restrict__ keys, uint64_t n, uint64_t chunk, int oc, uint32_t* counts,
             unsigned long long* g_stats, uint32_t* g_maxbin, uint32_t* g_hist) {
  constexpr int NP = (int)sizeof(K);  // 8-bit layout digits
  constexpr int NH = HIST ? (NP - 1) * 256 * HREP : 1;  // digits 1 .. NP-1, HREP copies
  __shared__ uint32_t s_low[2048];
  __shared__ uint32_t s_h[NH];
  for (int i = threadIdx.x; i < 2048; i += STATS_BLOCK) s_low[i] = 0u;
  if (HIST) {
    for (int i = threadIdx.x; i < NH; i += STATS_BLOCK) s_h[i] = 0u;
  }
  __syncthreads();
  const uint32_t lane = lane_id();
  uint32_t* const my_h = s_h + (HIST ? (lane % HREP) : 0u);
  auto digits = [&](K o) {
    if (!HIST) return;
    const uint64_t act = ballot64(true);
#pragma unroll
    for (int p = 1; p < NP; ++p) {
      const uint32_t bin = (uint32_t)((o >> (p * 8)) & (K)255);
      const uint32_t b0 = __builtin_amdgcn_readfirstlane(bin);
      uint32_t* const tab = s_h + (p - 1) * 256 * HREP;
      if (ballot64(bin == b0) == act) {
        if (lane == (uint32_t)__builtin_ctzll(act)) atomicAdd(&tab[b0 * HREP], (uint32_t)__popcll(act));
      } else {
        atomicAdd(&my_h[(p - 1) * 256 * HREP + bin * HREP], 1u);
      }
    }
  };
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  Stats4<K> st;
  st.init();
  constexpr int VEC = 16 / sizeof(K);
  const bool aligned = ((reinterpret_cast<uintptr_t>(keys + cbeg) & 15u) == 0);
  const uint64_t nv = aligned ? (cend - cbeg) / VEC : 0;
  const uint4* kv = reinterpret_cast<const uint4*>(keys + cbeg);
  uint64_t i = threadIdx.x;
  for (; i + 3 * STATS_BLOCK < nv; i += 4 * STATS_BLOCK) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = kv[i + u * STATS_BLOCK];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const K* e = reinterpret_cast<const K*>(&v[u]);
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const K o = ordered(e[j], oc);
        st.add(o);
        atomicAdd(&s_low[(uint32_t)o & 2047u], 1u);
        digits(o);
      }
    }
  }
  for (; i < nv; i += STATS_BLOCK) {
    uint4 v = kv[i];
    const K* e = reinterpret_cast<const K*>(&v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const K o = ordered(e[j], oc);
      st.add(o);
      atomicAdd(&s_low[(uint32_t)o & 2047u], 1u);
      digits(o);
    }
  }
  for (uint64_t t = cbeg + nv * VEC + threadIdx.x; t < cend; t += STATS_BLOCK) {
    const K o = ordered(keys[t], oc);
    st.add(o);
    atomicAdd(&s_low[(uint32_t)o & 2047u], 1u);
    digits(o);
  }
  unsigned long long a0 = st.mn, a1 = st.mx, a2 = st.an, a3 = st.orr;
  if (sizeof(K) == 4) a2 |= 0xFFFFFFFF00000000ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long b0 = __shfl_xor(a0, off, 64), b1 = __shfl_xor(a1, off, 64);
    const unsigned long long b2 = __shfl_xor(a2, off, 64), b3 = __shfl_xor(a3, off, 64);
    a0 = b0 < a0 ? b0 : a0;
    a1 = b1 > a1 ? b1 : a1;
    a2 &= b2;
    a3 |= b3;
  }
  // Same-address global atomics serialise at the coherence point (~10 ns
  // each), so the workgroup reduces in LDS and stores one partial per chunk.
  __shared__ unsigned long long s_red[4][STATS_BLOCK / 64];
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_red[0][wid] = a0;
    s_red[1][wid] = a1;
    s_red[2][wid] = a2;
    s_red[3][wid] = a3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < STATS_BLOCK / 64; ++w) {
      a0 = s_red[0][w] < a0 ? s_red[0][w] : a0;
      a1 = s_red[1][w] > a1 ? s_red[1][w] : a1;
      a2 &= s_red[2][w];
      a3 |= s_red[3][w];
    }
    g_stats[(size_t)blockIdx.x * 4 + 0] = a0;
    g_stats[(size_t)blockIdx.x * 4 + 1] = a1;
    g_stats[(size_t)blockIdx.x * 4 + 2] = a2;
    g_stats[(size_t)blockIdx.x * 4 + 3] = a3;
  }
  // the chunk's largest bin: summed over chunks it bounds the largest digit
  // count from above (the small-range groupby uses it to rule out skew)
  uint32_t mb = 0u;
  for (int t = threadIdx.x; t < 2048; t += STATS_BLOCK) {
    const uint32_t c = s_low[t];
    counts[(size_t)blockIdx.x * 2048 + t] = c;
    mb = c > mb ? c : mb;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(mb, off, 64);
    mb = o > mb ? o : mb;
  }
  __shared__ uint32_t s_mb[STATS_BLOCK / 64];
  if (lane_id() == 0) s_mb[wid] = mb;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < STATS_BLOCK / 64; ++w) mb = s_mb[w] > mb ? s_mb[w] : mb;
    g_maxbin[blockIdx.x] = mb;
  }
  if (HIST) {
    // s_low and s_h are complete: the barrier above follows every atomic
    if (threadIdx.x < 256) {
      uint32_t d0 = 0u;
#pragma unroll
      for (int j = 0; j < 8; ++j) d0 += s_low[threadIdx.x + 256 * j];
      if (d0) atomicAdd(&g_hist[threadIdx.x], d0);
    }
    for (int t = threadIdx.x; t < (NP - 1) * 256; t += STATS_BLOCK) {
      uint32_t v = 0u;
#pragma unroll
      for (int c = 0; c < HREP; ++c) v += s_h[t * HREP + c];
      if (v) atomicAdd(&g_hist[256 + t], v);
    }
  }
}

// Reduce the per-chunk statistics (min, max, and, or) into out[0..3] and the
// sum of the chunks' largest low-11-bit bins into out[4].
__global__ void __launch_bounds__(256)
k_stats_reduce(const unsigned long long* part, const uint32_t* maxbin, uint32_t nchunks,
               unsigned long long* out) {
  __shared__ unsigned long long s[5][256];
  unsigned long long a0 = ~0ull, a1 = 0ull, a2 = ~0ull, a3 = 0ull, a4 = 0ull;
  for (uint32_t b = threadIdx.x; b < nchunks; b += 256) {
    const unsigned long long* p = part + (size_t)b * 4;
    a0 = p[0] < a0 ? p[0] : a0;
    a1 = p[1] > a1 ? p[1] : a1;
    a2 &= p[2];
    a3 |= p[3];
    a4 += maxbin[b];
  }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  s[2][threadIdx.x] = a2;
  s[3][threadIdx.x] = a3;
  s[4][threadIdx.x] = a4;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const int o = threadIdx.x + off;
      s[0][threadIdx.x] = s[0][o] < s[0][threadIdx.x] ? s[0][o] : s[0][threadIdx.x];
      s[1][threadIdx.x] = s[1][o] > s[1][threadIdx.x] ? s[1][o] : s[1][threadIdx.x];
      s[2][threadIdx.x] &= s[2][o];
      s[3][threadIdx.x] |= s[3][o];
      s[4][threadIdx.x] += s[4][o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s[0][0];
    out[1] = s[1][0];
    out[2] = s[2][0];
    out[3] = s[3][0];
    out[4] = s[4][0];
  }
}

// Global histograms of the varying layout digits (bit p of vary set).
template <typename K, int D>
__global__ void __launch_bounds__(HIST_BLOCK)
k_sort_hist(const K* __restrict__ keys, uint64_t n, int oc, uint32_t vary, uint32_t* g_hist) {
  constexpr int RADIX = 1 << D;
  constexpr int NP = (int)(sizeof(K) * 8 + D - 1) / D;
  // interleaved copies against LDS bank conflicts, as in k_sort_stats (two
  // for the 11-bit radix, whose tables are 8x larger)
  constexpr int REP = D == 8 ? HREP : 2;
  __shared__ uint32_t s_h[NP * RADIX * REP];
  for (int i = threadIdx.x; i < NP * RADIX * REP; i += HIST_BLOCK) s_h[i] = 0u;
  __syncthreads();
  constexpr int VEC = 16 / sizeof(K);
  const uint64_t nvec = ((reinterpret_cast<uintptr_t>(keys) & 15u) == 0) ? n / VEC : 0;
  const uint64_t stride = (uint64_t)gridDim.x * HIST_BLOCK;
  const uint4* kv = reinterpret_cast<const uint4*>(keys);
  const uint32_t lane = lane_id();
  uint32_t* const my_h = s_h + lane % REP;
  // A digit on which the whole wave agrees (high digits of clustered or
  // frame-major keys) is added once with the wave's count instead of 64
  // colliding LDS atomics.
  auto one = [&](K raw) {
    const K o = ordered(raw, oc);
    const uint64_t act = ballot64(true);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      if (!((vary >> p) & 1u)) continue;
      const uint32_t bin = (uint32_t)((o >> (p * D)) & (K)(RADIX - 1));
      const uint32_t b0 = __builtin_amdgcn_readfirstlane(bin);
      if (ballot64(bin == b0) == act) {
        if (lane == (uint32_t)__builtin_ctzll(act)) atomicAdd(&s_h[(p * RADIX + b0) * REP], (uint32_t)__popcll(act));
      } else {
        atomicAdd(&my_h[(p * RADIX + bin) * REP], 1u);
      }
    }
  };
  uint64_t i = (uint64_t)blockIdx.x * HIST_BLOCK + threadIdx.x;
  for (; i + 3 * stride < nvec; i += 4 * stride) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = kv[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const K* e = reinterpret_cast<const K*>(&v[u]);
#pragma unroll
      for (int j = 0; j < VEC; ++j) one(e[j]);
    }
  }
  for (; i < nvec; i += stride) {
    uint4 v = kv[i];
    const K* e = reinterpret_cast<const K*>(&v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) one(e[j]);
  }
  for (uint64_t t = nvec * VEC + (uint64_t)blockIdx.x * HIST_BLOCK + threadIdx.x; t < n; t += stride)
    one(keys[t]);
  __syncthreads();
  for (int t = threadIdx.x; t < NP * RADIX; t += HIST_BLOCK) {
    uint32_t v = 0u;
#pragma unroll
    for (int c = 0; c < REP; ++c) v += s_h[t * REP + c];
    if (v) atomicAdd(&g_hist[t], v);
  }
}

// Exclusive digit offsets of every varying layout pass.
__global__ void __launch_bounds__(1024)
k_sort_plan(int radix, int npass, uint32_t vary, const uint32_t* g_hist, uint32_t* g_base) {
  __shared__ uint32_t s[2048];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  const int tid = threadIdx.x;
  for (int p = 0; p < npass; ++p) {
    if (!((vary >> p) & 1u)) continue;
    for (int i = tid; i < 2048; i += 1024) s[i] = i < radix ? g_hist[p * radix + i] : 0u;
    __syncthreads();
    block_excl_scan_lds<2048, 1024>(s, s_tmp);
    for (int i = tid; i < radix; i += 1024) g_base[p * radix + i] = s[i];
    __syncthreads();
  }
}

namespace {

void geometry(SortCtx& c) {
  const uint64_t tile = c.kb == 4 ? (uint64_t)ch_granule<uint32_t>() : (uint64_t)ch_granule<uint64_t>();
  const uint64_t tiles = (c.n + tile - 1) / tile;
  const uint64_t nch = std::max<uint64_t>(1, std::min<uint64_t>(SORT_MAXCHUNKS, tiles));
  const uint64_t tpc = (tiles + nch - 1) / nch;
  c.chunk = tpc * tile;
  c.nchunks = (uint32_t)((c.n + c.chunk - 1) / c.chunk);
}

uint64_t os_tiles(uint64_t n, uint32_t tile = OS_TILE) { return (n + tile - 1) / tile; }

// keys per tile of a shape
uint32_t shape_tile(int shape) { return shape == 4 ? 16384u : shape >= 2 ? 4096u : 8192u; }

void carve(SortCtx& c, Carver& cv) {
  c.alt_keys = cv.take<char>(c.n * c.kb);
  c.alt_pay = cv.take<char>(c.n * (uint64_t)c.pb);
  c.counts = cv.take<uint32_t>((uint64_t)SORT_MAXCHUNKS * 2048);
  c.colex = cv.take<uint32_t>((uint64_t)SORT_MAXCHUNKS * 2048);
  c.total = cv.take<uint32_t>(2048);
  c.hist = cv.take<uint32_t>(SORT_MAXP * SORT_MAXR);
  c.base = cv.take<uint32_t>(SORT_MAXP * SORT_MAXR);
  c.stats = cv.take<unsigned long long>(5);
  c.stats_part = cv.take<unsigned long long>((uint64_t)SORT_MAXCHUNKS * 4);
  c.maxbin = cv.take<uint32_t>(SORT_MAXCHUNKS);
  c.counters = cv.take<uint32_t>(64);
  c.hdr = cv.take<uint32_t>(16);
  c.desc[0] = cv.take<uint32_t>(os_tiles(c.n, c.desc_tile) * SORT_MAXR);
  c.desc[1] = cv.take<uint32_t>(os_tiles(c.n, c.desc_tile) * SORT_MAXR);
}

template <typename K, typename PT>
struct Passes {
  // The same OS_TILE keys per workgroup for both key widths: 32-bit keys as
  // 16 waves x 8 keys, 64-bit keys as 8 waves x 16 (measured at 100M keys +
  // u32 rows: 2.13 vs 2.22 ms for 32-bit keys; 8 x 16 best for 64-bit keys
  // among 8x8, 4x16, 8x12, 16x8, 4x24)
  static constexpr int W = sizeof(K) == 4 ? 16 : 8;
  static constexpr int KPT = OS_TILE / (W * 64);
  static_assert(W * 64 * KPT == OS_TILE, "onesweep tile shape");
  template <int W_, int KPT_>
  static void onesweep_wk(int D, const ScatterArgs<K, PT>& a, RadixDigit<K> dig, hipStream_t s) {
    const uint32_t tiles = (uint32_t)os_tiles(a.n, W_ * 64 * KPT_);
    if (D == 8) {
      JFX_LAUNCH("onesweep_d8", s, (k_onesweep<K, PT, 8, W_, KPT_, RadixDigit<K>><<<tiles, W_ * 64, 0, s>>>(a, dig)));
    } else {
      JFX_LAUNCH("onesweep_d11", s, (k_onesweep<K, PT, 11, W_, KPT_, RadixDigit<K>><<<tiles, W_ * 64, 0, s>>>(a, dig)));
    }
  }
  static hipError_t onesweep(int D, int shape, const ScatterArgs<K, PT>& a, RadixDigit<K> dig,
                             hipStream_t s) {
    constexpr bool K4 = sizeof(K) == 4;
    if constexpr (!std::is_same<PT, NoPay>::value) {
      switch (shape) {
        case 1: onesweep_wk<K4 ? 8 : 16, K4 ? 16 : 8>(D, a, dig, s); return hipGetLastError();
        case 2: onesweep_wk<K4 ? 8 : 4, K4 ? 8 : 16>(D, a, dig, s); return hipGetLastError();
        case 3: onesweep_wk<K4 ? 4 : 8, K4 ? 16 : 8>(D, a, dig, s); return hipGetLastError();
        case 4:
          if (D == 8) {
            onesweep_d8_wk<16, 16>(a, dig, s);
            return hipGetLastError();
          }
          break;
        default: break;
      }
    }
    // 64-bit keys with at most a 4-byte payload, 8-bit digits: 16384-key
    // tiles (16 waves x 16 keys), half the tiles and lookbacks per pass
    // (100M uint64 keys + uint32 rows: 5.09 against 5.39 ms of passes;
    // with 8-byte payloads 8192-key tiles stay faster, 6.50 against 6.78 ms
    // in the 1e8 x 1e7 join)
    if constexpr (sizeof(K) == 8 && sizeof(PT) <= 4) {
      if (D == 8 && shape == 0) {
        onesweep_d8_wk<16, 16>(a, dig, s);
        return hipGetLastError();
      }
    }
    onesweep_wk<W, KPT>(D, a, dig, s);
    return hipGetLastError();
  }
  template <int W_, int KPT_>
  static void onesweep_d8_wk(const ScatterArgs<K, PT>& a, RadixDigit<K> dig, hipStream_t s) {
    const uint32_t tiles = (uint32_t)os_tiles(a.n, W_ * 64 * KPT_);
    JFX_LAUNCH("onesweep_d8", s, (k_onesweep<K, PT, 8, W_, KPT_, RadixDigit<K>><<<tiles, W_ * 64, 0, s>>>(a, dig)));
  }
  template <int D, int W, int KPT, bool PF, int WPE>
  static void chunk_launch(const char* name, const ScatterArgs<K, PT>& a, const ChunkGeom& g,
                           const uint32_t* colex, const uint32_t* total, RadixDigit<K> dig,
                           hipStream_t s) {
    JFX_LAUNCH(name, s, (k_chunk_scatter<K, PT, D, W, KPT, PF, WPE, RadixDigit<K>><<<g.nchunks, W * 64, 0, s>>>(a, g, colex, total, dig)));
  }
  template <int D>
  static void chunk_cfgs(const char* name, const ScatterArgs<K, PT>& a, const ChunkGeom& g,
                         const uint32_t* colex, const uint32_t* total, RadixDigit<K> dig,
                         hipStream_t s) {
    switch (chunk_cfg()) {
      case 1: chunk_launch<D, 8, 16, false, 1>(name, a, g, colex, total, dig, s); break;
      case 2: chunk_launch<D, 8, 8, true, 4>(name, a, g, colex, total, dig, s); break;
      case 3: chunk_launch<D, 4, 16, false, 1>(name, a, g, colex, total, dig, s); break;
      case 4: chunk_launch<D, 4, 16, true, 1>(name, a, g, colex, total, dig, s); break;
      default:
        chunk_launch<D, 8, 8, false, 4>(name, a, g, colex, total, dig, s);
    }
  }
  static hipError_t chunked(int D, const ScatterArgs<K, PT>& a, const ChunkGeom& g,
                            const uint32_t* colex, const uint32_t* total, RadixDigit<K> dig,
                            hipStream_t s) {
    if (D == 8) chunk_cfgs<8>("chunk_scatter_d8", a, g, colex, total, dig, s);
    else if (D == 10) chunk_cfgs<10>("chunk_scatter_d10", a, g, colex, total, dig, s);
    else chunk_cfgs<11>("chunk_scatter_d11", a, g, colex, total, dig, s);
    return hipGetLastError();
  }
};

template <typename K>
int stats_launch(SortCtx& c, const void* keys, bool with_hist) {
  if (with_hist) {
    JFX_LAUNCH("sort_stats_hist", c.s, (k_sort_stats<K, true><<<c.nchunks, STATS_BLOCK, 0, c.s>>>(
                                          static_cast<const K*>(keys), c.n, c.chunk, c.oc, c.counts,
                                          c.stats_part, c.maxbin, c.hist)));
  } else {
    JFX_LAUNCH("sort_stats", c.s, (k_sort_stats<K, false><<<c.nchunks, STATS_BLOCK, 0, c.s>>>(
                                     static_cast<const K*>(keys), c.n, c.chunk, c.oc, c.counts,
                                     c.stats_part, c.maxbin, nullptr)));
  }
  JFX_HIP(hipGetLastError());
  c.hist_done = with_hist;
  JFX_LAUNCH("stats_reduce", c.s, (k_stats_reduce<<<1, 256, 0, c.s>>>(c.stats_part, c.maxbin,
                                                                      c.nchunks, c.stats)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <typename K, typename PT>
int single_pass_typed(SortCtx& c, const void* keys_in, void* keys_out, const void* pay_in,
                      void* pay_out, int iota, int descending) {
  const uint32_t R = c.range_radix();
  const int D = R <= 256 ? 8 : (R <= 1024 ? 10 : 11);
  JFX_LAUNCH("chunk_colscan", c.s, (k_chunk_colscan<<<(R + 3) / 4, 256, 0, c.s>>>(
                                        c.counts, 2048u, CM_ROTATE, (uint32_t)(c.mn & 2047u), R,
                                        c.nchunks, c.colex, c.total)));
  JFX_HIP(hipGetLastError());
  ScatterArgs<K, PT> a{};
  a.keys_in = static_cast<const K*>(keys_in);
  a.keys_out = static_cast<K*>(keys_out);
  a.pay_in = static_cast<const PT*>(pay_in);
  a.pay_out = static_cast<PT*>(pay_out);
  a.n = c.n;
  a.err = c.err;
  a.radix_eff = R;
  a.write_keys = keys_out != nullptr;
  a.reverse = descending;
  a.iota = iota;
  ChunkGeom g{c.n, c.chunk, c.nchunks};
  RadixDigit<K> dig{c.oc, 0u, (K)c.mn};
  JFX_HIP(Passes<K, PT>::chunked(D, a, g, c.colex, c.total, dig, c.s));
  return JFX_OK;
}

template <typename K, typename PT>
int multi_pass_typed(SortCtx& c, const void* keys_in, void* keys_out, int want_keys,
                     const void* pay_in, void* pay_out, int iota, int descending) {
  const int D = c.digit_bits;
  const int radix = 1 << D;
  const int npass = (c.kb * 8 + D - 1) / D;
  const uint64_t dmask = (uint64_t)radix - 1;
  const uint64_t diff = c.band ^ c.bor;
  uint32_t vary = 0;
  for (int p = 0; p < npass; ++p)
    if ((diff >> (p * D)) & dmask) vary |= 1u << p;
  // histograms of the varying digits (unless the statistics pass already
  // made all of them) and their offsets
  const uint32_t ub = 512;
  if (c.hist_done && D == 8) {
  } else if (D == 8) {
    JFX_LAUNCH("sort_hist", c.s, (k_sort_hist<K, 8><<<ub, HIST_BLOCK, 0, c.s>>>(
                                     static_cast<const K*>(keys_in), c.n, c.oc, vary, c.hist)));
  } else {
    JFX_LAUNCH("sort_hist", c.s, (k_sort_hist<K, 11><<<ub, HIST_BLOCK, 0, c.s>>>(
                                     static_cast<const K*>(keys_in), c.n, c.oc, vary, c.hist)));
  }
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("sort_plan", c.s, (k_sort_plan<<<1, 1024, 0, c.s>>>(radix, npass, vary, c.hist, c.base)));
  JFX_HIP(hipGetLastError());

  std::vector<int> passes;
  for (int p = 0; p < npass; ++p)
    if ((vary >> p) & 1u) passes.push_back(p);
  const int A = (int)passes.size();
  const uint64_t tiles = os_tiles(c.n, shape_tile(c.shape));
  // every layout pass is a onesweep pass: measured faster than a chunked
  // first pass fed by the statistics pass's chunk histograms (692 vs 814 us
  // per pass at 100M u64 keys + u32 rows, which also wrote 27% more)
  JFX_LAUNCH("fill_desc", c.s, (k_fill_u32<<<512, 256, 0, c.s>>>(c.desc[0], tiles * radix, 0u)));
  JFX_HIP(hipGetLastError());
  const K* src_k = static_cast<const K*>(keys_in);
  const PT* src_p = static_cast<const PT*>(pay_in);
  for (int a = 0; a < A; ++a) {
    const bool last = (a == A - 1);
    const bool to_out = ((A - 1 - a) % 2) == 0;
    K* dst_k = to_out ? static_cast<K*>(keys_out) : static_cast<K*>(c.alt_keys);
    PT* dst_p = to_out ? static_cast<PT*>(pay_out) : static_cast<PT*>(c.alt_pay);
    ScatterArgs<K, PT> args{};
    args.keys_in = src_k;
    args.keys_out = dst_k;
    args.pay_in = src_p;
    args.pay_out = dst_p;
    args.n = c.n;
    args.base = c.base + passes[a] * radix;
    args.desc = c.desc[a & 1];
    args.desc_next = last ? nullptr : c.desc[(a + 1) & 1];
    args.tile_counter = c.counters + a;
    args.err = c.err;
    args.radix_eff = (uint32_t)radix;
    args.write_keys = (!last || want_keys) ? 1 : 0;
    args.reverse = (last && descending) ? 1 : 0;
    args.iota = (a == 0 && iota) ? 1 : 0;
    RadixDigit<K> dig{c.oc, (uint32_t)(passes[a] * D), (K)0};
    JFX_HIP(Passes<K, PT>::onesweep(D, c.shape, args, dig, c.s));
    src_k = dst_k;
    src_p = dst_p;
  }
  return JFX_OK;
}

template <typename K>
int execute_k(SortCtx& c, const void* ki, void* ko, int want_keys, const void* pi, void* po,
              int iota, int desc) {
  if (c.single_pass()) {
    void* kout = want_keys ? ko : nullptr;
    if (c.pb == 0) return single_pass_typed<K, NoPay>(c, ki, kout, nullptr, nullptr, 0, desc);
    if (c.pb == 4) return single_pass_typed<K, uint32_t>(c, ki, kout, pi, po, iota, desc);
    return single_pass_typed<K, uint64_t>(c, ki, kout, pi, po, 0, desc);
  }
  if (c.pb == 0) return multi_pass_typed<K, NoPay>(c, ki, ko, want_keys, nullptr, nullptr, 0, desc);
  if (c.pb == 4) return multi_pass_typed<K, uint32_t>(c, ki, ko, want_keys, pi, po, iota, desc);
  return multi_pass_typed<K, uint64_t>(c, ki, ko, want_keys, pi, po, 0, desc);
}

}  // namespace

uint64_t sort_ws_bytes(int kt, int pay_bytes, uint64_t n, bool small_tiles) {
  SortCtx c;
  c.kb = key_bytes(kt);
  c.pb = pay_bytes;
  c.n = n;
  c.desc_tile = small_tiles ? 4096u : 8192u;
  Carver cv(nullptr, 0);
  carve(c, cv);
  return cv.used();
}

int default_digit_bits(int kt, uint64_t n) {
  (void)kt;
  (void)n;
  return 8;
}

int sort_prepare(SortCtx& c, int kt, const void* keys_in, uint64_t n, int pay_bytes,
                 int digit_bits, void* ws, uint64_t ws_bytes, hipStream_t s, bool with_hist) {
  c.kt = kt;
  c.kb = key_bytes(kt);
  c.oc = order_code(kt);
  c.pb = pay_bytes;
  c.n = n;
  c.s = s;
  c.digit_bits = digit_bits ? digit_bits : default_digit_bits(kt, n);
  JFX_CHECK(c.digit_bits == 8 || c.digit_bits == 11, JFX_ERR_INVALID, "digit_bits must be 8 or 11");
  Carver cv(ws, ws_bytes);
  carve(c, cv);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "sort workspace too small");
  geometry(c);
  c.err = device_error_word();
  JFX_CHECK(c.err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  JFX_LAUNCH("sort_init", s, (k_sort_init<<<64, 256, 0, s>>>(c.stats, c.counters, 64, c.hist,
                                                             SORT_MAXP * SORT_MAXR)));
  JFX_HIP(hipGetLastError());
  const bool hist = with_hist && c.digit_bits == 8;
  int rc = c.kb == 4 ? stats_launch<uint32_t>(c, keys_in, hist) : stats_launch<uint64_t>(c, keys_in, hist);
  if (rc != JFX_OK) return rc;
  unsigned long long h[5];
  rc = read_back(s, c.stats, sizeof(h), h);
  if (rc != JFX_OK) return rc;
  c.maxbin_sum = h[4];
  c.mn = h[0];
  c.mx = h[1];
  c.band = h[2];
  c.bor = h[3];
  if (c.kb == 4) c.band &= 0xFFFFFFFFull;
  return JFX_OK;
}

int sort_prepare_known(SortCtx& c, int kt, const void* keys_in, uint64_t n, int pay_bytes,
                       int digit_bits, uint32_t key_bits, void* ws, uint64_t ws_bytes,
                       hipStream_t s, int shape, bool small_tiles) {
  JFX_CHECK(kt == JFX_U32 || kt == JFX_U64, JFX_ERR_INVALID, "known-range sorts take unsigned keys");
  JFX_CHECK(shape >= 0 && shape < SORT_SHAPES, JFX_ERR_INVALID, "bad sort shape");
  JFX_CHECK(small_tiles || shape_tile(shape) >= 8192u, JFX_ERR_INVALID,
            "4096-key tiles need a workspace carved for them");
  c.shape = shape;
  c.desc_tile = small_tiles ? 4096u : 8192u;
  JFX_CHECK(key_bits >= 1 && key_bits <= (uint32_t)key_bytes(kt) * 8, JFX_ERR_INVALID, "bad key_bits");
  c.kt = kt;
  c.kb = key_bytes(kt);
  c.oc = order_code(kt);
  c.pb = pay_bytes;
  c.n = n;
  c.s = s;
  c.digit_bits = digit_bits ? digit_bits : default_digit_bits(kt, n);
  JFX_CHECK(c.digit_bits == 8 || c.digit_bits == 11, JFX_ERR_INVALID, "digit_bits must be 8 or 11");
  Carver cv(ws, ws_bytes);
  carve(c, cv);
  JFX_CHECK(cv.fits(), JFX_ERR_WORKSPACE, "sort workspace too small");
  geometry(c);
  c.err = device_error_word();
  JFX_CHECK(c.err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  JFX_LAUNCH("sort_init", s, (k_sort_init<<<64, 256, 0, s>>>(c.stats, c.counters, 64, c.hist,
                                                             SORT_MAXP * SORT_MAXR)));
  JFX_HIP(hipGetLastError());
  const uint64_t top = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
  // a single counting pass (keys below 2048) takes its chunk histograms from
  // the statistics kernel; layout passes need only the digit histograms
  if (top < 2048ull) {
    const int rc = c.kb == 4 ? stats_launch<uint32_t>(c, keys_in, false)
                             : stats_launch<uint64_t>(c, keys_in, false);
    if (rc != JFX_OK) return rc;
  }
  c.mn = 0;
  c.mx = top;
  c.band = 0;
  c.bor = top;
  return JFX_OK;
}

int sort_execute(SortCtx& c, const void* keys_in, void* keys_out, int want_keys, const void* pay_in,
                 void* pay_out, int iota, int descending) {
  return c.kb == 4 ? execute_k<uint32_t>(c, keys_in, keys_out, want_keys, pay_in, pay_out, iota, descending)
                   : execute_k<uint64_t>(c, keys_in, keys_out, want_keys, pay_in, pay_out, iota, descending);
}

int sort_single_pass(SortCtx& c, const void* keys_in, void* keys_out, const void* pay_in,
                     void* pay_out, int iota, int descending) {
  JFX_CHECK(c.single_pass(), JFX_ERR_INVALID, "single pass needs max-min < 2048");
  if (c.kb == 4) {
    if (c.pb == 0) return single_pass_typed<uint32_t, NoPay>(c, keys_in, keys_out, nullptr, nullptr, 0, descending);
    if (c.pb == 4) return single_pass_typed<uint32_t, uint32_t>(c, keys_in, keys_out, pay_in, pay_out, iota, descending);
    return single_pass_typed<uint32_t, uint64_t>(c, keys_in, keys_out, pay_in, pay_out, 0, descending);
  }
  if (c.pb == 0) return single_pass_typed<uint64_t, NoPay>(c, keys_in, keys_out, nullptr, nullptr, 0, descending);
  if (c.pb == 4) return single_pass_typed<uint64_t, uint32_t>(c, keys_in, keys_out, pay_in, pay_out, iota, descending);
  return single_pass_typed<uint64_t, uint64_t>(c, keys_in, keys_out, pay_in, pay_out, 0, descending);
}

int sort_digit_totals(SortCtx& c) {
  JFX_CHECK(c.single_pass(), JFX_ERR_INVALID, "digit totals need max-min < 2048");
  const uint32_t R = c.range_radix();
  JFX_LAUNCH("chunk_colscan", c.s, (k_chunk_colscan<<<(R + 3) / 4, 256, 0, c.s>>>(
                                        c.counts, 2048u, CM_ROTATE, (uint32_t)(c.mn & 2047u), R,
                                        c.nchunks, c.colex, c.total)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int sort_dispatch(int kt, const void* keys_in, void* keys_out, int want_keys, int pay_bytes,
                  const void* pay_in, void* pay_out, uint64_t n, int descending, int digit_bits,
                  void* ws, uint64_t ws_bytes, hipStream_t s) {
  SortCtx c;
  int rc = sort_prepare(c, kt, keys_in, n, pay_bytes, digit_bits, ws, ws_bytes, s, true);
  if (rc != JFX_OK) return rc;
  const int iota = (pay_bytes == 4 && pay_in == nullptr) ? 1 : 0;
  return sort_execute(c, keys_in, keys_out, want_keys, pay_in, pay_out, iota, descending);
}

}  // namespace jfx

extern "C" {

int jfx_sort_workspace_size(int key_type, int pay_bytes, uint64_t n, uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(pay_bytes == 0 || pay_bytes == 4 || pay_bytes == 8, JFX_ERR_INVALID,
            "pay_bytes must be 0, 4 or 8");
  *out_bytes = jfx::sort_ws_bytes(key_type, pay_bytes, n);
  return JFX_OK;
}

int jfx_sort(int key_type, const void* keys_in, void* keys_out, int want_keys, int pay_bytes,
             const void* pay_in, void* pay_out, uint64_t n, int descending, int digit_bits,
             void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(pay_bytes == 0 || pay_bytes == 4 || pay_bytes == 8, JFX_ERR_INVALID,
            "pay_bytes must be 0, 4 or 8");
  JFX_CHECK(n < (1ull << 30), JFX_ERR_UNSUPPORTED, "n must be below 2^30 rows");
  if (n == 0) return JFX_OK;
  JFX_CHECK(keys_in != nullptr && keys_out != nullptr, JFX_ERR_INVALID, "null key buffer");
  JFX_CHECK(keys_in != keys_out, JFX_ERR_INVALID, "keys_out must not alias keys_in");
  JFX_CHECK(pay_bytes == 0 || pay_out != nullptr, JFX_ERR_INVALID, "null payload output");
  JFX_CHECK(pay_bytes != 8 || pay_in != nullptr, JFX_ERR_INVALID, "8-byte payload needs pay_in");
  JFX_CHECK(pay_in == nullptr || pay_in != pay_out, JFX_ERR_INVALID,
            "pay_out must not alias pay_in");
  JFX_CHECK(workspace != nullptr, JFX_ERR_INVALID, "null workspace");
  JFX_CHECK(digit_bits == 0 || digit_bits == 8 || digit_bits == 11, JFX_ERR_INVALID,
            "digit_bits must be 0, 8 or 11");
  return jfx::sort_dispatch(key_type, keys_in, keys_out, want_keys, pay_bytes, pay_in, pay_out, n,
                            descending, digit_bits, workspace, ws_bytes, jfx::as_stream(stream));
}

}  // extern "C"
