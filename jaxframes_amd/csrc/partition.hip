// Row hashing and the local hash partition that feeds a repartition exchange.
//
// Restates hash_multi_columns (parallel_algorithms.py:69-136) and the local
// half of collective_hash_repartition: target = hash % P (:214) and the
// stable bucketing of _bucketize_local_column (:163-182). The reference
// builds a padded [P, local_n] buffer per column; here the output is compact:
// rows grouped by bucket (original order inside a bucket) plus P+1 offsets,
// which is what an all-to-all with exact counts consumes.
//   k_part_hist     one read of the keys; each chunk's bucket histogram
//   k_chunk_colscan the bucket totals (and each chunk's starting offsets)
//   k_part_offsets  bucket offsets (int64, P+1), the scatter's bucket starts,
//                   its lookback descriptors cleared
//   k_onesweep      the sort's stable scatter pass with the bucket as digit:
//                   keys (+ row index) by bucket, input order kept inside a
//                   bucket (k_chunk_scatter, the chunked scatter, with
//                   JFX_PART_SCATTER=chunked)
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "chunked.hpp"
#include "runtime.hpp"
#include "sortctx.hpp"

namespace jfx {

__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// uint32 image of a key as the reference hashes it: integers keep their low
// 32 bits (astype(uint32)), float32 keys their bit pattern, float64 keys the
// bits of their float32 rounding (the multi-key rule at :111-114).
template <typename K>
__device__ __forceinline__ uint32_t key_u32(K k, int kt) {
  if constexpr (sizeof(K) == 8) {
    if (kt == KT_F64) {
      const float f = (float)__longlong_as_double((long long)k);
      return __float_as_uint(f);
    }
    return (uint32_t)k;
  } else {
    return (uint32_t)k;
  }
}

// Bucket of a key: fmix32(u32(key)) % nb, or u32(key) % nb when the keys
// are already row hashes (raw).
template <typename K>
struct HashDigit {
  int kt;
  uint32_t nb;
  uint32_t pow2_mask;  // nb - 1 when nb is a power of two, else 0
  int raw;
  __device__ __forceinline__ uint32_t hash(K k) const {
    const uint32_t u = key_u32(k, kt);
    return raw ? u : fmix32(u);
  }
  __device__ __forceinline__ uint32_t operator()(K k) const {
    const uint32_t h = hash(k);
    return pow2_mask ? (h & pow2_mask) : (h % nb);
  }
};

// Each chunk's bucket histogram. Keys are read with 16-byte loads, four in
// flight per thread. Up to 256 buckets are counted in 16 lane-interleaved
// copies (bucket b of copy c at b * 16 + c, lane l adding to copy l % 16):
// the lanes of one copy share a group of 4 LDS banks and copies never share
// a bank, so an atomic instruction meets at most 4-way bank conflicts where
// one table of uniform buckets met 4-5.
template <typename K>
__global__ void __launch_bounds__(1024)
k_part_hist(const K* __restrict__ keys, uint64_t n, uint64_t chunk, HashDigit<K> hd,
            uint32_t* counts, uint32_t* hashes) {
  constexpr uint32_t REP = 16;
  __shared__ uint32_t s_h[2048 > 256 * REP ? 2048 : 256 * REP];
  const bool rep = hd.nb <= 256u;
  const uint32_t ents = rep ? hd.nb * REP : hd.nb;
  for (uint32_t i = threadIdx.x; i < ents; i += 1024) s_h[i] = 0u;
  __syncthreads();
  uint32_t* const my = s_h + (rep ? (lane_id() % REP) : 0u);
  const uint32_t sh = rep ? 4u : 0u;  // log2(REP)
  auto add = [&](uint64_t i, K k) {
    const uint32_t h = hd.hash(k);
    if (hashes) hashes[i] = h;
    const uint32_t b = hd.pow2_mask ? (h & hd.pow2_mask) : (h % hd.nb);
    atomicAdd(&my[b << sh], 1u);
  };
  const uint64_t cbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t cend = (cbeg + chunk) < n ? (cbeg + chunk) : n;
  constexpr int PER = 16 / (int)sizeof(K);
  uint64_t tail = cbeg;
  if ((reinterpret_cast<uintptr_t>(keys + cbeg) & 15u) == 0) {
    const uint64_t nv = (cend - cbeg) / PER;
    const uint4* kv = reinterpret_cast<const uint4*>(keys + cbeg);
    for (uint64_t v = threadIdx.x; v < nv; v += 4096) {
      uint4 x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = (v + u * 1024 < nv) ? kv[v + u * 1024] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (v + u * 1024 < nv) {
          K e[PER];
          __builtin_memcpy(e, &x[u], 16);
#pragma unroll
          for (int j = 0; j < PER; ++j) add(cbeg + (v + u * 1024) * PER + j, e[j]);
        }
      }
    }
    tail = cbeg + nv * PER;
  }
  for (uint64_t i = tail + threadIdx.x; i < cend; i += 1024) add(i, keys[i]);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < hd.nb; i += 1024) {
    uint32_t c = 0u;
    if (rep) {
#pragma unroll
      for (uint32_t r = 0; r < REP; ++r) c += s_h[i * REP + r];
    } else {
      c = s_h[i];
    }
    counts[(size_t)blockIdx.x * hd.nb + i] = c;
  }
}

// Exclusive bucket offsets (int64, nb + 1 entries) from the bucket totals.
__global__ void k_fill_i64(int64_t* p, uint32_t n, int64_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Bucket offsets from the bucket totals: int64 offsets[nb + 1] for the
// caller and, when base is non-null, the uint32 starts the onesweep scatter
// adds its lookback offsets to; also clears the scatter's lookback
// descriptors (ndesc words) and tile counter.
__global__ void __launch_bounds__(1024)
k_part_offsets(const uint32_t* total, uint32_t nb, int64_t* offsets, uint32_t* base,
               uint32_t* desc, uint64_t ndesc, uint32_t* tile_counter) {
  if (blockIdx.x > 0) {
    const uint64_t stride = (uint64_t)(gridDim.x - 1) * 1024;
    for (uint64_t i = (uint64_t)(blockIdx.x - 1) * 1024 + threadIdx.x; i < ndesc; i += stride)
      desc[i] = 0u;
    return;
  }
  __shared__ uint32_t s[2048];
  __shared__ uint32_t s_tmp[17];
  for (uint32_t i = threadIdx.x; i < 2048; i += 1024) s[i] = i < nb ? total[i] : 0u;
  __syncthreads();
  const uint32_t tot = block_excl_scan_lds<2048, 1024>(s, s_tmp);
  for (uint32_t i = threadIdx.x; i < nb; i += 1024) {
    offsets[i] = (int64_t)s[i];
    if (base) base[i] = s[i];
  }
  if (threadIdx.x == 0) {
    offsets[nb] = (int64_t)tot;
    if (tile_counter) *tile_counter = 0u;
  }
}

// Multi-key row hash (the K > 1 branch of hash_multi_columns): seed 0x1337;
// per key i: h ^= u32(k_i) * (0x9E3779B1 + i); h = rotl(h, 13); h *= 0x85EBCA6B;
// then fmix32. Single key: fmix32(u32(k)).
struct HashCols {
  const void* keys[8];
  int kt[8];
  int nkeys;
};

__global__ void __launch_bounds__(256) k_hash_rows(HashCols c, uint64_t n, uint32_t* out) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    uint32_t u[8];
    for (int j = 0; j < c.nkeys; ++j) {
      if (key_bytes(c.kt[j]) == 8) u[j] = key_u32(static_cast<const uint64_t*>(c.keys[j])[i], c.kt[j]);
      else u[j] = key_u32(static_cast<const uint32_t*>(c.keys[j])[i], c.kt[j]);
    }
    uint32_t h;
    if (c.nkeys == 1) {
      h = fmix32(u[0]);
    } else {
      h = 0x1337u;
      for (int j = 0; j < c.nkeys; ++j) {
        h ^= u[j] * (0x9E3779B1u + (uint32_t)j);
        h = (h << 13) | (h >> 19);
        h *= 0x85EBCA6Bu;
      }
      h = fmix32(h);
    }
    out[i] = h;
  }
}

namespace {

struct PartWs {
  uint32_t* counts;
  uint32_t* colex;
  uint32_t* total;
  uint32_t* base;     // [2048] bucket starts (onesweep scatter)
  uint32_t* counter;  // onesweep tile counter
  uint32_t* desc;     // [tiles][nb] onesweep lookback descriptors
  uint64_t chunk;
  uint32_t nchunks;
};

// onesweep scatter of the partition: the sort's tile of 8192 keys, 32-bit
// keys as 16 waves x 8 keys and 64-bit keys as 8 x 16
constexpr uint32_t PART_TILE = 8192;  // the smallest tile (sizes the descriptors)
uint64_t part_tiles(uint64_t n) { return (n + PART_TILE - 1) / PART_TILE; }
// Keys per onesweep tile: 16384 (16 waves x 16 keys, one workgroup per CU,
// keys ranked in pairs) for 64-bit keys into at most 256 buckets, measured
// at 2e8 int64 keys; 8192 otherwise.
constexpr uint32_t part_tile(int kb, uint32_t nb) { return (kb == 8 && nb <= 256) ? 16384u : 8192u; }

constexpr uint32_t PART_MAXCHUNKS = 512;

void part_geometry(PartWs& w, int kb, uint64_t n) {
  const uint64_t gran = kb == 4 ? 8192 : 4096;
  const uint64_t tiles = (n + gran - 1) / gran;
  const uint64_t nch = tiles < PART_MAXCHUNKS ? (tiles ? tiles : 1) : PART_MAXCHUNKS;
  const uint64_t tpc = (tiles + nch - 1) / nch;
  w.chunk = (tpc ? tpc : 1) * gran;
  w.nchunks = (uint32_t)((n + w.chunk - 1) / w.chunk);
}

void part_carve(PartWs& w, Carver& c, uint32_t nb, uint64_t n) {
  w.counts = c.take<uint32_t>((uint64_t)PART_MAXCHUNKS * nb);
  w.colex = c.take<uint32_t>((uint64_t)PART_MAXCHUNKS * nb);
  w.total = c.take<uint32_t>(2048);
  w.base = c.take<uint32_t>(2048);
  w.counter = c.take<uint32_t>(64);
  w.desc = c.take<uint32_t>(part_tiles(n) * nb);
}

template <typename K, typename PT, int D>
void part_onesweep_launch(const ScatterArgs<K, PT>& a, HashDigit<K> hd, hipStream_t s) {
  constexpr uint32_t TILE = part_tile((int)sizeof(K), D == 8 ? 256u : 2048u);
  constexpr int W = (sizeof(K) == 4 || TILE == 16384u) ? 16 : 8;
  constexpr int KPT = TILE / (W * 64);
  static_assert(W * 64 * KPT == (int)TILE, "partition tile shape");
  const uint32_t tiles = (uint32_t)((a.n + TILE - 1) / TILE);
  JFX_LAUNCH(D == 8 ? "part_onesweep_d8" : "part_onesweep_d11", s,
             (k_onesweep<K, PT, D, W, KPT, HashDigit<K>><<<tiles, W * 64, 0, s>>>(a, hd)));
}

template <typename K, typename PT, int D>
void part_scatter_launch(const ScatterArgs<K, PT>& a, const ChunkGeom& g, const uint32_t* colex,
                         const uint32_t* total, HashDigit<K> hd, hipStream_t s) {
  constexpr int KPT = sizeof(K) == 4 ? 16 : 8;
  JFX_LAUNCH(D == 8 ? "part_scatter_d8" : "part_scatter_d11", s,
             (k_chunk_scatter<K, PT, D, 8, KPT, false, 1, HashDigit<K>><<<g.nchunks, 512, 0, s>>>(
                 a, g, colex, total, hd)));
}

// Scatter of the partition: the sort's onesweep pass (tiles ranked in LDS,
// digit offsets by decoupled lookback), or with JFX_PART_SCATTER=chunked the
// chunked scatter fed by the histogram pass's chunk counts. Measured at 2e8
// int64 keys, 256 buckets: see the partition profile.
bool part_chunked() {
  const char* e = getenv("JFX_PART_SCATTER");
  return e && strcmp(e, "chunked") == 0;
}

template <typename K, typename PT>
int partition_typed(int kt, const void* keys, uint64_t n, uint32_t nb, int raw, void* out_keys,
                    void* out_rows, int64_t* out_offsets, uint32_t* out_hashes, PartWs& w,
                    hipStream_t s) {
  HashDigit<K> hd{kt, nb, (nb & (nb - 1)) == 0 ? nb - 1 : 0u, raw};
  uint32_t* err = device_error_word();
  JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  JFX_LAUNCH("part_hist", s, (k_part_hist<K><<<w.nchunks, 1024, 0, s>>>(
                                 static_cast<const K*>(keys), n, w.chunk, hd, w.counts, out_hashes)));
  JFX_HIP(hipGetLastError());
  JFX_LAUNCH("chunk_colscan", s, (k_chunk_colscan<<<(nb + 3) / 4, 256, 0, s>>>(
                                     w.counts, nb, CM_DIRECT, 0u, nb, w.nchunks, w.colex, w.total)));
  JFX_HIP(hipGetLastError());
  const bool chunked = part_chunked();
  const uint32_t ptile = part_tile((int)sizeof(K), nb);
  const uint64_t ndesc = chunked ? 0 : (n + ptile - 1) / ptile * nb;
  const uint32_t zero_blocks = ndesc ? (uint32_t)std::min<uint64_t>(256, (ndesc + 1023) / 1024) : 0u;
  JFX_LAUNCH("part_offsets", s, (k_part_offsets<<<1 + zero_blocks, 1024, 0, s>>>(
                                     w.total, nb, out_offsets, chunked ? nullptr : w.base, w.desc,
                                     ndesc, chunked ? nullptr : w.counter)));
  JFX_HIP(hipGetLastError());
  ScatterArgs<K, PT> a{};
  a.keys_in = static_cast<const K*>(keys);
  a.keys_out = static_cast<K*>(out_keys);
  a.pay_in = nullptr;
  a.pay_out = static_cast<PT*>(out_rows);
  a.n = n;
  a.err = err;
  a.radix_eff = nb;
  a.write_keys = out_keys != nullptr;
  a.reverse = 0;
  a.iota = 1;
  if (chunked) {
    ChunkGeom g{n, w.chunk, w.nchunks};
    if (nb <= 256) part_scatter_launch<K, PT, 8>(a, g, w.colex, w.total, hd, s);
    else part_scatter_launch<K, PT, 11>(a, g, w.colex, w.total, hd, s);
  } else {
    a.base = w.base;
    a.desc = w.desc;
    a.desc_next = nullptr;
    a.tile_counter = w.counter;
    if (nb <= 256) part_onesweep_launch<K, PT, 8>(a, hd, s);
    else part_onesweep_launch<K, PT, 11>(a, hd, s);
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

}  // namespace

}  // namespace jfx

extern "C" {

int jfx_hash_rows(const void* const* keys, const int* key_types, int nkeys, uint64_t n,
                  uint32_t* out_hashes, void* stream) {
  JFX_CHECK(nkeys >= 1 && nkeys <= 8, JFX_ERR_UNSUPPORTED, "1 to 8 key columns are supported");
  JFX_CHECK(out_hashes != nullptr || n == 0, JFX_ERR_INVALID, "null output");
  if (n == 0) return JFX_OK;
  jfx::HashCols c{};
  for (int j = 0; j < nkeys; ++j) {
    JFX_CHECK(keys[j] != nullptr, JFX_ERR_INVALID, "null key column");
    JFX_CHECK(key_types[j] >= JFX_U32 && key_types[j] <= JFX_F64, JFX_ERR_INVALID, "bad key type");
    c.keys[j] = keys[j];
    c.kt[j] = key_types[j];
  }
  c.nkeys = nkeys;
  hipStream_t s = jfx::as_stream(stream);
  const uint64_t blocks = std::min<uint64_t>(4096, (n + 255) / 256);
  JFX_LAUNCH("hash_rows", s, (jfx::k_hash_rows<<<(uint32_t)blocks, 256, 0, s>>>(c, n, out_hashes)));
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

int jfx_hash_partition_workspace_size(int key_type, uint64_t n, uint32_t nbuckets,
                                      uint64_t* out_bytes) {
  JFX_CHECK(out_bytes != nullptr, JFX_ERR_INVALID, "out_bytes is null");
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(nbuckets >= 1 && nbuckets <= 2048, JFX_ERR_UNSUPPORTED, "1 to 2048 buckets");
  jfx::Carver c(nullptr, 0);
  jfx::PartWs w{};
  jfx::part_carve(w, c, nbuckets, n);
  *out_bytes = c.used();
  return JFX_OK;
}

int jfx_hash_partition(int key_type, const void* keys, uint64_t n, uint32_t nbuckets,
                       int prehashed, void* out_keys, uint32_t* out_rows, int64_t* out_offsets,
                       uint32_t* out_hashes, void* workspace, uint64_t ws_bytes, void* stream) {
  JFX_CHECK(key_type >= JFX_U32 && key_type <= JFX_F64, JFX_ERR_INVALID, "bad key type");
  JFX_CHECK(nbuckets >= 1 && nbuckets <= 2048, JFX_ERR_UNSUPPORTED, "1 to 2048 buckets");
  JFX_CHECK(n < (1ull << 32), JFX_ERR_UNSUPPORTED, "n must be below 2^32 rows");
  JFX_CHECK(out_offsets != nullptr, JFX_ERR_INVALID, "null offsets");
  hipStream_t s = jfx::as_stream(stream);
  if (n == 0) {
    JFX_LAUNCH("fill_offsets", s, (jfx::k_fill_i64<<<1, 256, 0, s>>>(out_offsets, nbuckets + 1, 0)));
    JFX_HIP(hipGetLastError());
    return JFX_OK;
  }
  JFX_CHECK(!prehashed || key_type == JFX_U32, JFX_ERR_INVALID, "prehashed keys must be uint32");
  JFX_CHECK(keys != nullptr && workspace != nullptr, JFX_ERR_INVALID, "null keys or workspace");
  JFX_CHECK(out_keys == nullptr || out_keys != keys, JFX_ERR_INVALID, "out_keys must not alias keys");
  jfx::Carver c(workspace, ws_bytes);
  jfx::PartWs w{};
  jfx::part_carve(w, c, nbuckets, n);
  JFX_CHECK(c.fits(), JFX_ERR_WORKSPACE, "partition workspace too small");
  jfx::part_geometry(w, jfx::key_bytes(key_type), n);
  if (jfx::key_bytes(key_type) == 4) {
    if (out_rows)
      return jfx::partition_typed<uint32_t, uint32_t>(key_type, keys, n, nbuckets, prehashed, out_keys,
                                                      out_rows, out_offsets, out_hashes, w, s);
    return jfx::partition_typed<uint32_t, jfx::NoPay>(key_type, keys, n, nbuckets, prehashed, out_keys,
                                                      nullptr, out_offsets, out_hashes, w, s);
  }
  if (out_rows)
    return jfx::partition_typed<uint64_t, uint32_t>(key_type, keys, n, nbuckets, 0, out_keys,
                                                    out_rows, out_offsets, out_hashes, w, s);
  return jfx::partition_typed<uint64_t, jfx::NoPay>(key_type, keys, n, nbuckets, 0, out_keys, nullptr,
                                                    out_offsets, out_hashes, w, s);
}

}  // extern "C"
