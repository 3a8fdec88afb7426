// ggks_sort.hip -- full device sort of a key vector (rocPRIM radix sort), the
// sorting step of the reference's sort-and-choose baseline
// (multiselectTimingFunctions.cu:41-95, a thrust::sort of a device copy
// followed by picks at n - k).  Kept in its own translation unit so the
// rocPRIM headers compile once.  Floats sort in rocPRIM's radix order, which
// is the order-preserving key order (-0.0 before +0.0), as the reference's
// b40c radix sort.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>

namespace ggks {

template <typename T>
static hipError_t sort_keys(void* temp, size_t& bytes, const void* in, void* out, uint64_t n,
                            hipStream_t s) {
  return rocprim::radix_sort_keys(temp, bytes, static_cast<const T*>(in), static_cast<T*>(out),
                                  (size_t)n, 0, (unsigned)(8 * sizeof(T)), s);
}

static hipError_t sort_dispatch(int dtype, void* temp, size_t& bytes, const void* in, void* out,
                                uint64_t n, hipStream_t s) {
  switch (dtype) {
    case 0: return sort_keys<float>(temp, bytes, in, out, n, s);
    case 1: return sort_keys<double>(temp, bytes, in, out, n, s);
    case 2: return sort_keys<uint32_t>(temp, bytes, in, out, n, s);
    default: return sort_keys<int32_t>(temp, bytes, in, out, n, s);
  }
}

// Temporary storage the sort of n keys of dtype needs (0 on error).
size_t device_sort_temp_bytes(int dtype, uint64_t n) {
  size_t bytes = 0;
  if (sort_dispatch(dtype, nullptr, bytes, nullptr, nullptr, n, nullptr) != hipSuccess) return 0;
  return bytes;
}

// Sort n keys from `in` into `out` (ascending), stream-ordered.
hipError_t device_sort(int dtype, const void* in, void* out, uint64_t n, void* temp,
                       size_t temp_bytes, hipStream_t s) {
  size_t bytes = temp_bytes;
  return sort_dispatch(dtype, temp, bytes, in, out, n, s);
}

}  // namespace ggks
