// ggks_compat.hpp -- the reference's C++ wrapper names over libggks.so.
//
// Header-only, host C++ only (no HIP or CUDA headers): code written against
// the reference's wrappers compiles against these templates unchanged and
// links with -lggks.  Each forwards to the C ABI in ggks.h.  Semantics are
// the reference's (k-th largest, 1-based, outputs in k order, top-K
// descending) with two deliberate differences: the input vector is never
// modified, and errors throw ggks_compat::error instead of printing and
// returning the success value.
#ifndef GGKS_COMPAT_HPP
#define GGKS_COMPAT_HPP

#include <stdint.h>

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "ggks.h"

namespace ggks_compat {

struct error : std::runtime_error {
  int code;
  error(int c, const char* what) : std::runtime_error(what ? what : "ggks error"), code(c) {}
};

inline void check(int rc) {
  if (rc != GGKS_OK) throw error(rc, ggks_last_error());
}

template <typename T> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = GGKS_F32; };
template <> struct dtype_of<double> { static constexpr int value = GGKS_F64; };
template <> struct dtype_of<unsigned int> { static constexpr int value = GGKS_U32; };
template <> struct dtype_of<int> { static constexpr int value = GGKS_I32; };

}  // namespace ggks_compat

namespace BucketMultiselect {

// bucketMultiselect.cu:837-871: outputs[i] = the kVals_ori[i]-th largest of
// the first `length` elements of the device vector d_vector; kVals_ori and
// outputs are host arrays of numKs entries.  blocks / threads are hints
// (0 = tuned), as the reference takes them from the device properties.
// Returns T(1), the reference's return value (its `return 1` from a function
// returning T, bucketMultiselect.cu:837-838, :870).
template <typename T>
T bucketMultiselectWrapper(T* d_vector, int length, unsigned int* kVals_ori, int numKs,
                           T* outputs, int blocks, int threads) {
  ggks_opts o;
  std::memset(&o, 0, sizeof o);
  o.blocks = blocks;
  o.threads = threads;
  ggks_compat::check(ggks_bucket_multiselect(nullptr, ggks_compat::dtype_of<T>::value, d_vector,
                                             length, kVals_ori, numKs, outputs, &o));
  return (T)1;
}

}  // namespace BucketMultiselect

namespace RadixSelect {

// merrillSelect.cu:210-259: the k-th largest of the first `size` elements of
// the device vector d_vec (the reference overwrites the vector; this does
// not).
template <typename T>
T RadixSelectWrapper(T* d_vec, unsigned int size, unsigned int k) {
  T out;
  ggks_compat::check(ggks_radix_multiselect(nullptr, ggks_compat::dtype_of<T>::value, d_vec,
                                            size, &k, 1, &out, nullptr));
  return out;
}

}  // namespace RadixSelect

namespace InplaceRadix {

// inplaceRadixSelect.cu:83-108: the k-th largest of the first `size` elements
// of the device vector (uint, float and double overloads, as in the
// reference; the reference converts the vector in place, this leaves it
// intact).
inline unsigned int inplaceRadixSelectWrapper(unsigned int* d_vec, unsigned int size,
                                              unsigned int k) {
  return RadixSelect::RadixSelectWrapper<unsigned int>(d_vec, size, k);
}
inline float inplaceRadixSelectWrapper(float* d_vec, unsigned int size, unsigned int k) {
  return RadixSelect::RadixSelectWrapper<float>(d_vec, size, k);
}
inline double inplaceRadixSelectWrapper(double* d_vec, unsigned int size, unsigned int k) {
  return RadixSelect::RadixSelectWrapper<double>(d_vec, size, k);
}

}  // namespace InplaceRadix

// randomizedTopkSelect.cu:210-213: the k largest of the first `size`
// elements of the device vector `list`, descending, in a malloc'd host array
// the caller frees (as with the reference's return value).
template <typename T>
T* randomizedTopkSelectWrapper(T* list, int size, int k) {
  T* out = static_cast<T*>(std::malloc(sizeof(T) * (size_t)(k > 0 ? k : 1)));
  if (!out) throw ggks_compat::error(GGKS_ENOMEM, "malloc failed");
  const int rc = ggks_topk(nullptr, ggks_compat::dtype_of<T>::value, list, size, k, out, nullptr);
  if (rc != GGKS_OK) {
    std::free(out);
    ggks_compat::check(rc);
  }
  return out;
}

#endif  // GGKS_COMPAT_HPP
