// Host-side runtime shared by the C-ABI entry points: error reporting,
// workspace carving, a pinned staging buffer per thread, the device error word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "../../include/jfx.h"

namespace jfx {

int set_error(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define JFX_HIP(...)                                          \
  do {                                                        \
    hipError_t _e = (__VA_ARGS__);                            \
    if (_e != hipSuccess) return ::jfx::hip_fail(_e, #__VA_ARGS__); \
  } while (0)

#define JFX_CHECK(cond, code, msg)                      \
  do {                                                  \
    if (!(cond)) return ::jfx::set_error((code), (msg)); \
  } while (0)

// Bump allocator over a caller-provided workspace (256-byte aligned pieces).
struct Carver {
  char* base;
  uint64_t cap;
  uint64_t off = 0;
  Carver(void* b, uint64_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <typename T> T* take(uint64_t count) {
    off = (off + 255) & ~uint64_t(255);
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += count * sizeof(T);
    return p;
  }
  bool fits() const { return off <= cap; }
  uint64_t used() const { return (off + 255) & ~uint64_t(255); }
};

// A per-thread pinned host buffer for small device-to-host reads.
void* pinned_scratch(size_t bytes);

// Reads `bytes` (a multiple of 4, at most READBACK_MAX) of device memory
// written by earlier work on stream s into host_out: a one-workgroup kernel
// copies them into this thread's coherent pinned buffer and then publishes a
// sequence number; the host spins on it (about 6 us after the data is ready,
// against about 15 us for a copy plus hipStreamSynchronize), and after 2 ms
// falls back to hipStreamSynchronize so that long waits block and stream
// errors surface.
constexpr size_t READBACK_MAX = 16384 - 64;
int read_back(hipStream_t s, const void* dev_src, size_t bytes, void* host_out);

// The same exchange with the words written by a kernel of the caller: the
// kernel stores them to slot.dst, then __threadfence_system() and a
// system-scope release store of slot.seq to *slot.seq_slot; read_back_wait
// spins for that sequence number (falling back to a stream sync after 2 ms)
// and copies the words out. One slot per host thread, one exchange at a time.
struct ReadBackSlot {
  uint32_t* dst;        // device view of the coherent host buffer
  uint64_t* seq_slot;   // device view of its sequence word
  uint64_t seq;         // the value to publish
};
int read_back_slot(ReadBackSlot& out);
int read_back_wait(hipStream_t s, const ReadBackSlot& slot, size_t bytes, void* host_out);

// A result of a 2-bit status and a value below 2^30, published as one 64-bit
// word in the slot's sequence word: the low 32 bits of the sequence number
// above the status and value. One relaxed system-scope store and no fence:
// the host needs only this word (the kernel's other outputs are read later in
// stream order), where a separate flag would need a system-scope release
// ordering the words before it.
__host__ __device__ inline uint64_t rb_word(uint64_t seq, uint32_t status, uint32_t value) {
  return ((uint64_t)(uint32_t)seq << 32) | ((uint64_t)(status & 3u) << 30) |
         (uint64_t)(value < (1u << 30) ? value : (1u << 30) - 1u);
}
// Spins for slot.seq's word (falling back to a stream sync after 2 ms) and
// decodes it into res[0] = status, res[1] = value.
int read_back_wait_word(hipStream_t s, const ReadBackSlot& slot, uint32_t* res);
// Decodes a word copied back by the caller; false when it is not seq's.
bool rb_decode(uint64_t word, uint64_t seq, uint32_t* res);

// Device error word (one per device, in device memory owned by the library
// image as a __device__ variable).
uint32_t* device_error_word();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Per-kernel timing with HIP events on the launch stream (enabled by
// jfx_prof_enable). prof_begin/prof_end bracket one launch; when profiling
// is off they cost one branch.
bool prof_wants(const char* name);
void prof_begin(hipStream_t s);
void prof_end(const char* name, hipStream_t s);

// Launch helper: times the kernel launch statement given as the trailing
// arguments under `name`.
#define JFX_LAUNCH(name, stream, ...)                    \
  do {                                                   \
    const bool _jfx_prof = ::jfx::prof_wants(name);      \
    if (_jfx_prof) ::jfx::prof_begin(stream);            \
    __VA_ARGS__;                                         \
    if (_jfx_prof) ::jfx::prof_end(name, stream);        \
  } while (0)

}  // namespace jfx
