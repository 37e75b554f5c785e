// Vendor-library baseline for the benchmark: rocPRIM's radix_sort_pairs on
// uint64 keys carrying uint32 values, timed beside jfx_sort on the same
// device-resident inputs. Built into its own libjfx_cmp.so, which only the
// benchmark loads; nothing in libjfx.so calls it.
#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

extern "C" {

// temp == nullptr: *temp_bytes receives the scratch size rocPRIM asks for.
// Otherwise sorts n (key, value) pairs by the key's 64 bits, stable.
// Returns 0, or the hipError_t of the failing call.
int jfx_cmp_rocprim_sort_u64_u32(const uint64_t* keys_in, uint64_t* keys_out,
                                 const uint32_t* vals_in, uint32_t* vals_out, uint64_t n,
                                 void* temp, uint64_t* temp_bytes, hipStream_t stream) {
  size_t tb = temp ? (size_t)*temp_bytes : 0;
  const hipError_t e = rocprim::radix_sort_pairs(temp, tb, keys_in, keys_out, vals_in, vals_out,
                                                 (size_t)n, 0u, 64u, stream);
  if (!temp) *temp_bytes = (uint64_t)tb;
  return (int)e;
}

}  // extern "C"
