// HBM bandwidth probes: a read-only stream and a copy, hand-written with
// 16-byte loads, against which the kernels' achieved bandwidth is read.
//
// k_stream_read   every thread walks the buffer with a grid stride, U
//                 16-byte loads in flight per step, and folds them into an
//                 XOR that each workgroup stores once (so no load is dead);
// k_stream_copy   the same walk storing what it loads.
// NT selects non-temporal loads (streamed past the caches' reuse policy).
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

template <int U, bool NT>
int probe_launch(int mode, const void* src, void* dst, uint64_t n16, uint32_t blocks, hipStream_t s) {
  if (mode == 0) {
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
  JFX_CHECK(mode == 0 || mode == 1, JFX_ERR_INVALID, "mode must be 0 (read) or 1 (copy)");
  JFX_CHECK(src != nullptr && dst != nullptr, JFX_ERR_INVALID, "null buffer");
  JFX_CHECK(bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0,
            JFX_ERR_INVALID, "buffers and size must be 16-byte aligned");
  JFX_CHECK(blocks >= 1 && blocks <= (1u << 20), JFX_ERR_INVALID, "blocks must be 1..2^20");
  hipStream_t s = jfx::as_stream(stream);
  const uint64_t n16 = bytes / 16;
  return nontemporal ? jfx::probe_unroll<true>(mode, src, dst, n16, blocks, unroll, s)
                     : jfx::probe_unroll<false>(mode, src, dst, n16, blocks, unroll, s);
}
