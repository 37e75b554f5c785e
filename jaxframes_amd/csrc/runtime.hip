// Host runtime: error strings, pinned staging, device error word, version.
#include "runtime.hpp"

#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

namespace jfx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return JFX_ERR_HIP;
}

namespace {
struct Pinned {
  void* p = nullptr;
  size_t cap = 0;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};
thread_local Pinned t_pinned;
}  // namespace

void* pinned_scratch(size_t bytes) {
  if (t_pinned.cap < bytes) {
    if (t_pinned.p) (void)hipHostFree(t_pinned.p);
    t_pinned.p = nullptr;
    t_pinned.cap = 0;
    size_t cap = bytes < 4096 ? 4096 : bytes;
    if (hipHostMalloc(&t_pinned.p, cap, hipHostMallocDefault) != hipSuccess) return nullptr;
    t_pinned.cap = cap;
  }
  return t_pinned.p;
}

namespace {
struct Coherent {
  uint32_t* h = nullptr;   // host view
  uint32_t* d = nullptr;   // device view
  uint64_t seq = 0;
  ~Coherent() {
    if (h) (void)hipHostFree(h);
  }
};
thread_local Coherent t_coh;
constexpr size_t RB_BYTES = READBACK_MAX + 64;
constexpr size_t RB_SEQ = READBACK_MAX / sizeof(uint64_t);  // index of the sequence word
}  // namespace

__global__ void __launch_bounds__(256)
k_read_back(const uint32_t* src, uint32_t nwords, uint32_t* dst, uint64_t* seq_slot, uint64_t seq) {
  for (uint32_t i = threadIdx.x; i < nwords; i += 256) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(seq_slot, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

int read_back_slot(ReadBackSlot& out) {
  if (!t_coh.h) {
    void* p = nullptr;
    JFX_HIP(hipHostMalloc(&p, RB_BYTES, hipHostMallocCoherent | hipHostMallocMapped));
    void* dp = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&dp, p, 0);
    if (e != hipSuccess) {
      (void)hipHostFree(p);
      return hip_fail(e, "hipHostGetDevicePointer");
    }
    t_coh.h = static_cast<uint32_t*>(p);
    t_coh.d = static_cast<uint32_t*>(dp);
  }
  out.dst = t_coh.d;
  out.seq_slot = reinterpret_cast<uint64_t*>(t_coh.d) + RB_SEQ;
  out.seq = ++t_coh.seq;
  return JFX_OK;
}

int read_back_wait(hipStream_t s, const ReadBackSlot& slot, size_t bytes, void* host_out) {
  JFX_CHECK(bytes <= READBACK_MAX, JFX_ERR_INVALID, "bad read-back size");
  uint64_t* hseq = reinterpret_cast<uint64_t*>(t_coh.h) + RB_SEQ;
  const uint64_t seq = slot.seq;
  const auto t0 = std::chrono::steady_clock::now();
  uint32_t polls = 0;
  while (__atomic_load_n(hseq, __ATOMIC_ACQUIRE) != seq) {
    if ((++polls & 255u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      JFX_HIP(hipStreamSynchronize(s));
      JFX_CHECK(__atomic_load_n(hseq, __ATOMIC_ACQUIRE) == seq, JFX_ERR_HIP,
                "read-back did not complete");
      break;
    }
  }
  std::memcpy(host_out, t_coh.h, bytes);
  return JFX_OK;
}

bool rb_decode(uint64_t word, uint64_t seq, uint32_t* res) {
  if ((uint32_t)(word >> 32) != (uint32_t)seq) return false;
  res[0] = (uint32_t)(word >> 30) & 3u;
  res[1] = (uint32_t)word & ((1u << 30) - 1u);
  return true;
}

int read_back_wait_word(hipStream_t s, const ReadBackSlot& slot, uint32_t* res) {
  uint64_t* hseq = reinterpret_cast<uint64_t*>(t_coh.h) + RB_SEQ;
  const auto t0 = std::chrono::steady_clock::now();
  uint32_t polls = 0;
  while (!rb_decode(__atomic_load_n(hseq, __ATOMIC_ACQUIRE), slot.seq, res)) {
    if ((++polls & 255u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      JFX_HIP(hipStreamSynchronize(s));
      JFX_CHECK(rb_decode(__atomic_load_n(hseq, __ATOMIC_ACQUIRE), slot.seq, res), JFX_ERR_HIP,
                "read-back did not complete");
      break;
    }
  }
  return JFX_OK;
}

int read_back(hipStream_t s, const void* dev_src, size_t bytes, void* host_out) {
  JFX_CHECK(bytes % 4 == 0 && bytes <= READBACK_MAX, JFX_ERR_INVALID, "bad read-back size");
  ReadBackSlot slot{};
  int rc = read_back_slot(slot);
  if (rc != JFX_OK) return rc;
  JFX_LAUNCH("read_back", s, (k_read_back<<<1, 256, 0, s>>>(static_cast<const uint32_t*>(dev_src),
                                                           (uint32_t)(bytes / 4), slot.dst,
                                                           slot.seq_slot, slot.seq)));
  JFX_HIP(hipGetLastError());
  return read_back_wait(s, slot, bytes, host_out);
}

__device__ uint32_t g_device_error;

uint32_t* device_error_word() {
  // One symbol address per device; resolved with the device current.
  static std::mutex mu;
  static std::vector<uint32_t*> addr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_device_error)) != hipSuccess) return nullptr;
    addr[dev] = static_cast<uint32_t*>(p);
  }
  return addr[dev];
}

namespace {
struct ProfState {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> pool;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::string filter;   // time only kernels whose name starts with this
  uint32_t every = 1;   // ... and of those every `every`-th launch
  uint64_t seen = 0;    // launches that passed the filter since the count restarted
  struct Acc {
    double ms = 0.0;
    uint64_t count = 0;
  };
  std::vector<std::pair<std::string, Acc>> acc;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // events are created up front, so that a timed launch does not pay for
  // hipEventCreate
  void reserve(size_t k) {
    for (size_t have = pool.size() + 2 * pending.size(); have < k; ++have) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return;
      pool.push_back(e);
    }
  }
  Acc& slot(const std::string& n) {
    for (auto& p : acc)
      if (p.first == n) return p.second;
    acc.emplace_back(n, Acc{});
    return acc.back().second;
  }
  int collect() {
    for (auto& p : pending) {
      if (hipEventSynchronize(p.b) != hipSuccess) return JFX_ERR_HIP;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) return JFX_ERR_HIP;
      Acc& a = slot(p.name);
      a.ms += ms;
      a.count += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
    return JFX_OK;
  }
};
ProfState g_prof;
thread_local hipEvent_t t_prof_start = nullptr;  // start event of this thread's open launch
}  // namespace

bool prof_wants(const char* name) {
  if (!g_prof.on) return false;
  std::lock_guard<std::mutex> lk(g_prof.mu);
  if (!g_prof.filter.empty() &&
      std::strncmp(name, g_prof.filter.c_str(), g_prof.filter.size()) != 0)
    return false;
  return g_prof.seen++ % g_prof.every == 0;
}

void prof_begin(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  t_prof_start = g_prof.get();
  (void)hipEventRecord(t_prof_start, s);
}

void prof_end(const char* name, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  hipEvent_t b = g_prof.get();
  (void)hipEventRecord(b, s);
  g_prof.pending.push_back({name, t_prof_start, b});
  t_prof_start = nullptr;
}

}  // namespace jfx

extern "C" {

int jfx_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  jfx::g_prof.on = on != 0;
  if (jfx::g_prof.on) jfx::g_prof.reserve(512);
  return JFX_OK;
}

int jfx_prof_filter(const char* name_prefix) {
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  jfx::g_prof.filter = name_prefix ? name_prefix : "";
  jfx::g_prof.seen = 0;
  return JFX_OK;
}

int jfx_prof_sample(uint32_t every) {
  JFX_CHECK(every >= 1, JFX_ERR_INVALID, "every must be at least 1");
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  jfx::g_prof.every = every;
  jfx::g_prof.seen = 0;
  return JFX_OK;
}

int jfx_prof_reset(void) {
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  int rc = jfx::g_prof.collect();
  jfx::g_prof.acc.clear();
  jfx::g_prof.seen = 0;
  return rc;
}

int jfx_prof_count(int* out_n) {
  JFX_CHECK(out_n != nullptr, JFX_ERR_INVALID, "out_n is null");
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  int rc = jfx::g_prof.collect();
  *out_n = (int)jfx::g_prof.acc.size();
  return rc;
}

int jfx_prof_get(int i, char* name_buf, uint64_t buf_len, double* total_ms, uint64_t* count) {
  std::lock_guard<std::mutex> lk(jfx::g_prof.mu);
  JFX_CHECK(i >= 0 && i < (int)jfx::g_prof.acc.size(), JFX_ERR_INVALID, "index out of range");
  const auto& p = jfx::g_prof.acc[i];
  if (name_buf && buf_len) {
    size_t m = p.first.size() < buf_len - 1 ? p.first.size() : buf_len - 1;
    memcpy(name_buf, p.first.data(), m);
    name_buf[m] = 0;
  }
  if (total_ms) *total_ms = p.second.ms;
  if (count) *count = p.second.count;
  return JFX_OK;
}

const char* jfx_last_error(void) { return jfx::g_last_error.c_str(); }

int jfx_version(void) { return 1; }

int jfx_device_error(void* stream, uint32_t* out_code) {
  JFX_CHECK(out_code != nullptr, JFX_ERR_INVALID, "out_code is null");
  uint32_t* w = jfx::device_error_word();
  JFX_CHECK(w != nullptr, JFX_ERR_HIP, "cannot resolve the device error word");
  hipStream_t s = jfx::as_stream(stream);
  uint32_t* h = static_cast<uint32_t*>(jfx::pinned_scratch(sizeof(uint32_t)));
  JFX_CHECK(h != nullptr, JFX_ERR_HIP, "pinned host allocation failed");
  JFX_HIP(hipMemcpyAsync(h, w, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  JFX_HIP(hipStreamSynchronize(s));
  *out_code = *h;
  if (*h != 0) {
    uint32_t zero = 0;
    JFX_HIP(hipMemcpyAsync(w, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    JFX_HIP(hipStreamSynchronize(s));
  }
  return JFX_OK;
}

}  // extern "C"
