// HBM bandwidth probes: a read-only stream and a copy, hand-written with
// 16-byte loads, against which the kernels' achieved bandwidth is read.
//
// k_stream_read   every thread walks the buffer with a grid stride, U
//                 16-byte loads in flight per step, and folds them into an
//                 XOR that each workgroup stores once (so no load is dead);
// k_stream_copy   the same walk storing what it loads.
// k_stream_write  the same walk storing a value derived from the index (no
//                 loads).
// k_scatter_copy  a radix pass's memory pattern without its ranking: tiles
//                 of 16384 uint64 keys (and uint32 payloads) read in order
//                 and written as 16384 / runlen runs per tile, run d of tile
//                 t to [d * T * runlen + t * runlen, +runlen) of T tiles, so
//                 that every run lands in a different region of the output
//                 as a digit run of an ideally uniform radix pass does.
// NT selects non-temporal loads (streamed past the caches' reuse policy); for
// the scatter copy it selects non-temporal stores.
#include "runtime.hpp"

namespace jfx {
namespace {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__device__ __forceinline__ v4u ld16(const v4u* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_stream_read(const v4u* src, uint64_t n16, uint32_t* sink) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  v4u acc = {0u, 0u, 0u, 0u};
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    v4u v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = ld16<U, NT>(src + i + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) acc ^= v[k];
  }
  for (; i < n16; i += stride) acc ^= ld16<U, NT>(src + i);
  __shared__ uint32_t s_x[256];
  s_x[threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t x = 0u;
    for (int t = 0; t < 256; ++t) x ^= s_x[t];
    sink[blockIdx.x] = x;
  }
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const v4u* src, v4u* dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    v4u v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = ld16<U, NT>(src + i + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) dst[i + k * stride] = v[k];
  }
  for (; i < n16; i += stride) dst[i] = ld16<U, NT>(src + i);
}

constexpr uint32_t SC_TILE = 16384;  // keys per tile of the scatter copy
constexpr int SC_KPT = 16;             // keys per lane (16 waves of 64 lanes)

// One tile per workgroup; lane l of wave w holds keys w*1024 + i*64 + l
// (the onesweep load order), so one wave store writes 64 consecutive keys of
// a run (runlen >= 64) or 64 / runlen whole runs.
template <bool PAY, bool NT>
__global__ void __launch_bounds__(1024)
k_scatter_copy(const uint64_t* kin, const uint32_t* pin, uint64_t* kout, uint32_t* pout,
               uint32_t tiles, uint32_t runlen) {
  const uint32_t t = blockIdx.x;
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
  const uint64_t tb = (uint64_t)t * SC_TILE;
  const uint64_t region = (uint64_t)tiles * runlen;
  uint64_t k[SC_KPT];
  uint32_t p[SC_KPT];
#pragma unroll
  for (int i = 0; i < SC_KPT; ++i) {
    const uint32_t j = w * 64u * SC_KPT + (uint32_t)i * 64u + l;
    k[i] = kin[tb + j];
    if constexpr (PAY) p[i] = pin[tb + j];
  }
#pragma unroll
  for (int i = 0; i < SC_KPT; ++i) {
    const uint32_t j = w * 64u * SC_KPT + (uint32_t)i * 64u + l;
    const uint64_t pos = (uint64_t)(j / runlen) * region + (uint64_t)t * runlen + (j % runlen);
    if constexpr (NT) __builtin_nontemporal_store(k[i], kout + pos);
    else kout[pos] = k[i];
  }
  if constexpr (PAY) {
#pragma unroll
    for (int i = 0; i < SC_KPT; ++i) {
      const uint32_t j = w * 64u * SC_KPT + (uint32_t)i * 64u + l;
      const uint64_t pos = (uint64_t)(j / runlen) * region + (uint64_t)t * runlen + (j % runlen);
      if constexpr (NT) __builtin_nontemporal_store(p[i], pout + pos);
      else pout[pos] = p[i];
    }
  }
}

template <bool NT>
int scatter_launch(bool pay, const void* src, void* dst, uint64_t n, uint32_t runlen,
                   hipStream_t s) {
  const uint32_t tiles = (uint32_t)(n / SC_TILE);
  const uint64_t* kin = static_cast<const uint64_t*>(src);
  uint64_t* kout = static_cast<uint64_t*>(dst);
  const uint32_t* pin = reinterpret_cast<const uint32_t*>(kin + n);
  uint32_t* pout = reinterpret_cast<uint32_t*>(kout + n);
  if (pay) {
    JFX_LAUNCH("scatter_copy", s, (k_scatter_copy<true, NT><<<tiles, 1024, 0, s>>>(
                                       kin, pin, kout, pout, tiles, runlen)));
  } else {
    JFX_LAUNCH("scatter_copy", s, (k_scatter_copy<false, NT><<<tiles, 1024, 0, s>>>(
                                       kin, nullptr, kout, nullptr, tiles, runlen)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_stream_write(v4u* dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t x = (uint32_t)(i + k * stride);
      const v4u v = {x, x + 1u, x + 2u, x + 3u};
      if constexpr (NT) __builtin_nontemporal_store(v, dst + i + k * stride);
      else dst[i + k * stride] = v;
    }
  }
  for (; i < n16; i += stride) {
    const uint32_t x = (uint32_t)i;
    const v4u v = {x, x + 1u, x + 2u, x + 3u};
    if constexpr (NT) __builtin_nontemporal_store(v, dst + i);
    else dst[i] = v;
  }
}

template <int U, bool NT>
int probe_launch(int mode, const void* src, void* dst, uint64_t n16, uint32_t blocks, hipStream_t s) {
  if (mode == 4) {
    JFX_LAUNCH("stream_write", s, (k_stream_write<U, NT><<<blocks, 256, 0, s>>>(
                                       static_cast<v4u*>(dst), n16)));
  } else if (mode == 0) {
    JFX_LAUNCH("stream_read", s, (k_stream_read<U, NT><<<blocks, 256, 0, s>>>(
                                      static_cast<const v4u*>(src), n16, static_cast<uint32_t*>(dst))));
  } else {
    JFX_LAUNCH("stream_copy", s, (k_stream_copy<U, NT><<<blocks, 256, 0, s>>>(
                                      static_cast<const v4u*>(src), static_cast<v4u*>(dst), n16)));
  }
  JFX_HIP(hipGetLastError());
  return JFX_OK;
}

template <bool NT>
int probe_unroll(int mode, const void* src, void* dst, uint64_t n16, uint32_t blocks, uint32_t unroll,
                 hipStream_t s) {
  switch (unroll) {
    case 1: return probe_launch<1, NT>(mode, src, dst, n16, blocks, s);
    case 2: return probe_launch<2, NT>(mode, src, dst, n16, blocks, s);
    case 4: return probe_launch<4, NT>(mode, src, dst, n16, blocks, s);
    case 8: return probe_launch<8, NT>(mode, src, dst, n16, blocks, s);
  }
  return set_error(JFX_ERR_INVALID, "unroll must be 1, 2, 4 or 8");
}

}  // namespace
}  // namespace jfx

extern "C" int jfx_bandwidth_probe(int mode, const void* src, void* dst, uint64_t bytes,
                                   uint32_t blocks, uint32_t unroll, int nontemporal, void* stream) {
  JFX_CHECK(mode >= 0 && mode <= 4, JFX_ERR_INVALID,
            "mode must be 0 (read), 1 (copy), 2 (scatter copy of keys and payloads), 3 "
            "(scatter copy of keys) or 4 (write)");
  JFX_CHECK(src != nullptr && dst != nullptr, JFX_ERR_INVALID, "null buffer");
  if (mode == 2 || mode == 3) {
    // unroll is the run length in keys; blocks is unused (one tile each)
    const uint64_t rec = mode == 2 ? 12u : 8u;
    JFX_CHECK(bytes % (rec * jfx::SC_TILE) == 0 && bytes / rec / jfx::SC_TILE <= (1u << 20),
              JFX_ERR_INVALID, "scatter copy: bytes must be a multiple of 16384 records, "
              "at most 2^20 tiles");
    JFX_CHECK(unroll >= 1 && unroll <= jfx::SC_TILE && jfx::SC_TILE % unroll == 0,
              JFX_ERR_INVALID, "scatter copy: run length must divide 16384");
    JFX_CHECK((uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0, JFX_ERR_INVALID,
              "buffers must be 16-byte aligned");
    const uint64_t n = bytes / rec;
    hipStream_t s = jfx::as_stream(stream);
    return nontemporal ? jfx::scatter_launch<true>(mode == 2, src, dst, n, unroll, s)
                       : jfx::scatter_launch<false>(mode == 2, src, dst, n, unroll, s);
  }
  JFX_CHECK(bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0,
            JFX_ERR_INVALID, "buffers and size must be 16-byte aligned");
  JFX_CHECK(blocks >= 1 && blocks <= (1u << 20), JFX_ERR_INVALID, "blocks must be 1..2^20");
  hipStream_t s = jfx::as_stream(stream);
  const uint64_t n16 = bytes / 16;
  return nontemporal ? jfx::probe_unroll<true>(mode, src, dst, n16, blocks, unroll, s)
                     : jfx::probe_unroll<false>(mode, src, dst, n16, blocks, unroll, s);
}
