// Host runtime: error strings, pinned staging, device error word, version.
#include "runtime.hpp"

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

}  // namespace jfx

extern "C" {

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
