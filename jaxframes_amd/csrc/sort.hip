// Stable LSD radix sort (onesweep) for 32/64-bit keys with 0/4/8-byte payloads.
//
// Restates ParallelRadixSort.sort (parallel_algorithms.py:463-581): the
// reference runs one stable argsort per 8-bit digit; here each digit is one
// scatter pass over HBM. Pipeline per call:
//   k_sort_init     zero histograms / counters, seed min/max
//   k_sort_upfront  one read of the keys: min/max of the ordered keys, the
//                   digit histograms of every pass of the layout, and a
//                   histogram of the low 11 bits
//   k_sort_plan     per-pass skip flags (a digit shared by all keys) and
//                   exclusive digit offsets; offsets for a single range-reduced
//                   pass when max-min < 2048 (digit = key - min)
//   host            reads the 5-word plan header (one stream sync)
//   k_onesweep x P  one scatter pass per varying digit
#include <algorithm>
#include <vector>

#include "onesweep.hpp"
#include "runtime.hpp"

namespace jfx {

constexpr int MAXP = 8;          // most layout passes (64-bit keys, 8-bit digits)
constexpr int MAXR = 2048;       // largest radix
constexpr int SORT_W = 8;        // waves per scatter workgroup
constexpr int SORT_KPT = 16;     // keys per thread
constexpr int SORT_TILE = SORT_W * 64 * SORT_KPT;
constexpr int UPFRONT_BLOCK = 1024;

__global__ void k_sort_init(uint32_t* hist, uint32_t nhist, unsigned long long* minmax,
                            uint32_t* counters, uint32_t ncounters) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nhist; i += stride) hist[i] = 0u;
  if (blockIdx.x == 0) {
    if (threadIdx.x < ncounters) counters[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
      minmax[0] = ~0ull;
      minmax[1] = 0ull;
    }
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename K, int D>
__device__ __forceinline__ void upfront_one(K raw, int oc, uint32_t* s_h, uint32_t* s_low, K& mn,
                                            K& mx) {
  constexpr int RADIX = 1 << D;
  constexpr int NP = (int)(sizeof(K) * 8 + D - 1) / D;
  const K o = ordered(raw, oc);
  mn = o < mn ? o : mn;
  mx = o > mx ? o : mx;
  atomicAdd(&s_low[(uint32_t)o & 2047u], 1u);
#pragma unroll
  for (int p = 0; p < NP; ++p)
    atomicAdd(&s_h[p * RADIX + (uint32_t)((o >> (p * D)) & (K)(RADIX - 1))], 1u);
}

template <typename K, int D>
__global__ void __launch_bounds__(UPFRONT_BLOCK)
k_sort_upfront(const K* __restrict__ keys, uint64_t n, int oc, uint32_t* g_hist, uint32_t* g_low,
               unsigned long long* g_minmax) {
  constexpr int RADIX = 1 << D;
  constexpr int NP = (int)(sizeof(K) * 8 + D - 1) / D;
  __shared__ uint32_t s_h[NP * RADIX];
  __shared__ uint32_t s_low[2048];
  for (int i = threadIdx.x; i < NP * RADIX; i += blockDim.x) s_h[i] = 0u;
  for (int i = threadIdx.x; i < 2048; i += blockDim.x) s_low[i] = 0u;
  __syncthreads();
  K mn = ~K(0), mx = K(0);
  constexpr int VEC = 16 / sizeof(K);
  const uint64_t nvec = ((reinterpret_cast<uintptr_t>(keys) & 15u) == 0) ? n / VEC : 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint4* kv = reinterpret_cast<const uint4*>(keys);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
    uint4 v = kv[i];
    const K* e = reinterpret_cast<const K*>(&v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) upfront_one<K, D>(e[j], oc, s_h, s_low, mn, mx);
  }
  for (uint64_t i = nvec * VEC + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    upfront_one<K, D>(keys[i], oc, s_h, s_low, mn, mx);
  // wave min/max
  unsigned long long wmn = (unsigned long long)mn, wmx = (unsigned long long)mx;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long a = __shfl_xor(wmn, off, 64);
    unsigned long long b = __shfl_xor(wmx, off, 64);
    wmn = a < wmn ? a : wmn;
    wmx = b > wmx ? b : wmx;
  }
  if (lane_id() == 0) {
    atomicMin(&g_minmax[0], wmn);
    atomicMax(&g_minmax[1], wmx);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NP * RADIX; i += blockDim.x) {
    const uint32_t v = s_h[i];
    if (v) atomicAdd(&g_hist[i], v);
  }
  for (int i = threadIdx.x; i < 2048; i += blockDim.x) {
    const uint32_t v = s_low[i];
    if (v) atomicAdd(&g_low[i], v);
  }
}

// Plan header words: [0,1] min (lo,hi), [2,3] max (lo,hi), [4] skip mask.
__global__ void __launch_bounds__(1024)
k_sort_plan(int radix, int npass, uint64_t n, const uint32_t* g_hist, const uint32_t* g_low,
            const unsigned long long* g_minmax, uint32_t* g_base, uint32_t* g_base11,
            uint32_t* g_base8, uint32_t* hdr) {
  __shared__ uint32_t s[2048];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  __shared__ uint32_t s_skip;
  const int tid = threadIdx.x;
  const unsigned long long mn = g_minmax[0];
  if (tid == 0) s_skip = 0u;
  for (int p = 0; p < npass; ++p) {
    for (int i = tid; i < 2048; i += 1024) {
      const uint32_t c = i < radix ? g_hist[p * radix + i] : 0u;
      s[i] = c;
      if ((uint64_t)c == n) atomicOr(&s_skip, 1u << p);
    }
    __syncthreads();
    block_excl_scan_lds<2048, 1024>(s, s_tmp);
    for (int i = tid; i < radix; i += 1024) g_base[p * radix + i] = s[i];
    __syncthreads();
  }
  // single range-reduced pass, 11-bit digit: rotate the low-11-bit histogram
  for (int i = tid; i < 2048; i += 1024) s[(uint32_t)(i - (uint32_t)mn) & 2047u] = g_low[i];
  __syncthreads();
  block_excl_scan_lds<2048, 1024>(s, s_tmp);
  for (int i = tid; i < 2048; i += 1024) g_base11[i] = s[i];
  __syncthreads();
  for (int i = tid; i < 2048; i += 1024) s[i] = 0u;
  __syncthreads();
  for (int i = tid; i < 2048; i += 1024) atomicAdd(&s[(uint32_t)(i - (uint32_t)mn) & 255u], g_low[i]);
  __syncthreads();
  block_excl_scan_lds<2048, 1024>(s, s_tmp);
  for (int i = tid; i < 256; i += 1024) g_base8[i] = s[i];
  if (tid == 0) {
    hdr[0] = (uint32_t)mn;
    hdr[1] = (uint32_t)(mn >> 32);
    hdr[2] = (uint32_t)g_minmax[1];
    hdr[3] = (uint32_t)(g_minmax[1] >> 32);
    hdr[4] = s_skip;
  }
}

namespace {

struct SortWs {
  void* alt_keys;
  void* alt_pay;
  uint32_t* hist;
  uint32_t* low;
  unsigned long long* minmax;
  uint32_t* base;
  uint32_t* base11;
  uint32_t* base8;
  uint32_t* counters;
  uint32_t* hdr;
  uint32_t* desc[2];
  uint64_t bytes;
};

uint64_t num_tiles(uint64_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

SortWs carve_sort_ws(void* ws, uint64_t cap, int kb, int pb, uint64_t n) {
  Carver c(ws, cap);
  SortWs w{};
  w.alt_keys = c.take<char>(n * kb);
  w.alt_pay = c.take<char>(n * (uint64_t)pb);
  w.hist = c.take<uint32_t>(MAXP * MAXR);
  w.low = c.take<uint32_t>(2048);
  w.minmax = c.take<unsigned long long>(2);
  w.base = c.take<uint32_t>(MAXP * MAXR);
  w.base11 = c.take<uint32_t>(2048);
  w.base8 = c.take<uint32_t>(256);
  w.counters = c.take<uint32_t>(64);
  w.hdr = c.take<uint32_t>(16);
  w.desc[0] = c.take<uint32_t>(num_tiles(n) * MAXR);
  w.desc[1] = c.take<uint32_t>(num_tiles(n) * MAXR);
  w.bytes = c.used();
  return w;
}

template <typename K, typename PT, int D>
hipError_t launch_pass(const ScatterArgs<K, PT>& a, RadixDigit<K> dig, hipStream_t s) {
  const uint32_t tiles = (uint32_t)num_tiles(a.n);
  k_onesweep<K, PT, D, SORT_W, SORT_KPT, RadixDigit<K>><<<tiles, SORT_W * 64, 0, s>>>(a, dig);
  return hipGetLastError();
}

template <typename K, typename PT>
int sort_impl(int kt, const K* keys_in, K* keys_out, int want_keys, const PT* pay_in, PT* pay_out,
              int iota, uint64_t n, int descending, int digit_bits, void* ws, uint64_t ws_bytes,
              hipStream_t s) {
  const int kb = sizeof(K);
  constexpr int pb = std::is_same<PT, NoPay>::value ? 0 : (int)sizeof(PT);
  SortWs w = carve_sort_ws(ws, ws_bytes, kb, pb, n);
  JFX_CHECK(w.bytes <= ws_bytes, JFX_ERR_WORKSPACE, "sort workspace too small");
  const int oc = order_code(kt);
  const int D = digit_bits;
  const int radix = 1 << D;
  const int npass = (kb * 8 + D - 1) / D;
  uint32_t* err = device_error_word();
  JFX_CHECK(err != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");

  // histograms: hist (MAXP*MAXR) and low (2048) are carved contiguously? zero both.
  k_sort_init<<<64, 256, 0, s>>>(w.hist, MAXP * MAXR, w.minmax, w.counters, 64);
  JFX_HIP(hipGetLastError());
  k_fill_u32<<<8, 256, 0, s>>>(w.low, 2048, 0u);
  JFX_HIP(hipGetLastError());
  int dev = 0, cus = 256;
  JFX_HIP(hipGetDevice(&dev));
  JFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  uint64_t want_blocks = (n + UPFRONT_BLOCK * 16 - 1) / (UPFRONT_BLOCK * 16);
  uint32_t ub = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus, want_blocks));
  if (D == 8)
    k_sort_upfront<K, 8><<<ub, UPFRONT_BLOCK, 0, s>>>(keys_in, n, oc, w.hist, w.low, w.minmax);
  else
    k_sort_upfront<K, 11><<<ub, UPFRONT_BLOCK, 0, s>>>(keys_in, n, oc, w.hist, w.low, w.minmax);
  JFX_HIP(hipGetLastError());
  k_sort_plan<<<1, 1024, 0, s>>>(radix, npass, n, w.hist, w.low, w.minmax, w.base, w.base11,
                                 w.base8, w.hdr);
  JFX_HIP(hipGetLastError());
  uint32_t* h = static_cast<uint32_t*>(pinned_scratch(16 * sizeof(uint32_t)));
  JFX_CHECK(h != nullptr, JFX_ERR_HIP, "pinned host allocation failed");
  JFX_HIP(hipMemcpyAsync(h, w.hdr, 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  JFX_HIP(hipStreamSynchronize(s));
  const uint64_t mn = (uint64_t)h[0] | ((uint64_t)h[1] << 32);
  const uint64_t mx = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
  const uint32_t skip = h[4];
  const uint64_t range = mx - mn;

  struct PassSpec {
    int d;
    uint32_t shift;
    K sub;
    const uint32_t* base;
    uint32_t radix_eff;
  };
  std::vector<PassSpec> passes;
  if (range < 256) {
    passes.push_back({8, 0u, (K)mn, w.base8, (uint32_t)range + 1u});
  } else if (range < 2048) {
    passes.push_back({11, 0u, (K)mn, w.base11, (uint32_t)range + 1u});
  } else {
    for (int p = 0; p < npass; ++p)
      if (!((skip >> p) & 1u)) passes.push_back({D, (uint32_t)(p * D), (K)0, w.base + p * radix, (uint32_t)radix});
  }
  const int A = (int)passes.size();
  const uint64_t tiles = num_tiles(n);
  k_fill_u32<<<512, 256, 0, s>>>(w.desc[0], tiles * passes[0].radix_eff, 0u);
  JFX_HIP(hipGetLastError());

  const K* src_k = keys_in;
  const PT* src_p = pay_in;
  for (int a = 0; a < A; ++a) {
    const bool last = (a == A - 1);
    // the final pass lands in the caller's buffers
    const bool to_out = ((A - 1 - a) % 2) == 0;
    K* dst_k = to_out ? keys_out : static_cast<K*>(w.alt_keys);
    PT* dst_p = to_out ? pay_out : static_cast<PT*>(w.alt_pay);
    ScatterArgs<K, PT> args{};
    args.keys_in = src_k;
    args.keys_out = dst_k;
    args.pay_in = src_p;
    args.pay_out = dst_p;
    args.n = n;
    args.base = passes[a].base;
    args.desc = w.desc[a & 1];
    args.desc_next = last ? nullptr : w.desc[(a + 1) & 1];
    args.tile_counter = w.counters + a;
    args.err = err;
    args.radix_eff = passes[a].radix_eff;
    args.write_keys = (!last || want_keys) ? 1 : 0;
    args.reverse = (last && descending) ? 1 : 0;
    args.iota = (a == 0 && iota) ? 1 : 0;
    RadixDigit<K> dig{oc, passes[a].shift, passes[a].sub};
    hipError_t e = passes[a].d == 8 ? launch_pass<K, PT, 8>(args, dig, s)
                                    : launch_pass<K, PT, 11>(args, dig, s);
    JFX_HIP(e);
    src_k = dst_k;
    src_p = dst_p;
  }
  return JFX_OK;
}

}  // namespace

int sort_dispatch(int kt, const void* keys_in, void* keys_out, int want_keys, int pay_bytes,
                  const void* pay_in, void* pay_out, uint64_t n, int descending, int digit_bits,
                  void* ws, uint64_t ws_bytes, hipStream_t s) {
  const int kb = key_bytes(kt);
  const bool iota = (pay_bytes == 4 && pay_in == nullptr);
  if (kb == 4) {
    using K = uint32_t;
    if (pay_bytes == 0)
      return sort_impl<K, NoPay>(kt, (const K*)keys_in, (K*)keys_out, want_keys, nullptr, nullptr,
                                 0, n, descending, digit_bits, ws, ws_bytes, s);
    if (pay_bytes == 4)
      return sort_impl<K, uint32_t>(kt, (const K*)keys_in, (K*)keys_out, want_keys,
                                    (const uint32_t*)pay_in, (uint32_t*)pay_out, iota, n,
                                    descending, digit_bits, ws, ws_bytes, s);
    return sort_impl<K, uint64_t>(kt, (const K*)keys_in, (K*)keys_out, want_keys,
                                  (const uint64_t*)pay_in, (uint64_t*)pay_out, 0, n, descending,
                                  digit_bits, ws, ws_bytes, s);
  }
  using K = uint64_t;
  if (pay_bytes == 0)
    return sort_impl<K, NoPay>(kt, (const K*)keys_in, (K*)keys_out, want_keys, nullptr, nullptr, 0,
                               n, descending, digit_bits, ws, ws_bytes, s);
  if (pay_bytes == 4)
    return sort_impl<K, uint32_t>(kt, (const K*)keys_in, (K*)keys_out, want_keys,
                                  (const uint32_t*)pay_in, (uint32_t*)pay_out, iota, n, descending,
                                  digit_bits, ws, ws_bytes, s);
  return sort_impl<K, uint64_t>(kt, (const K*)keys_in, (K*)keys_out, want_keys,
                                (const uint64_t*)pay_in, (uint64_t*)pay_out, 0, n, descending,
                                digit_bits, ws, ws_bytes, s);
}

uint64_t sort_ws_bytes(int kt, int pay_bytes, uint64_t n) {
  return carve_sort_ws(nullptr, 0, key_bytes(kt), pay_bytes, n).bytes;
}

int default_digit_bits(int kt, uint64_t n) {
  (void)kt;
  (void)n;
  return 8;
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
  if (digit_bits == 0) digit_bits = jfx::default_digit_bits(key_type, n);
  JFX_CHECK(digit_bits == 8 || digit_bits == 11, JFX_ERR_INVALID, "digit_bits must be 0, 8 or 11");
  return jfx::sort_dispatch(key_type, keys_in, keys_out, want_keys, pay_bytes, pay_in, pay_out, n,
                            descending, digit_bits, workspace, ws_bytes, jfx::as_stream(stream));
}

}  // extern "C"
