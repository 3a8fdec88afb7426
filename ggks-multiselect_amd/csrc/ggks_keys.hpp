// ggks_keys.hpp -- order-preserving key conversion, host and device.
//
// Every kernel works on raw bit patterns (uint32_t for float32/uint32/int32,
// uint64_t for float64) converted on load into unsigned keys whose integer
// order is the value order.  The maps are the reference's:
//   float/double  keyConversion.cu:10-24 (preProcess), :26-42 (postProcess);
//                 b40c radixsort_key_conversion.h:112-131, :142-160
//   int           radixsort_key_conversion.h:249-265 (flip the sign bit)
//   unsigned      identity
// They are written branch-free here: for floats the mask is all ones when the
// sign bit is set and the sign bit alone otherwise.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GGKS_HD __host__ __device__ __forceinline__
#else
#define GGKS_HD inline
#endif

namespace ggks {

enum Conv : int { CONV_U = 0, CONV_I = 1, CONV_F = 2 };

template <typename KT>
struct KeyBits {
  static constexpr int W = int(sizeof(KT) * 8);
  static constexpr KT TOP = KT(1) << (W - 1);
  static constexpr KT MAX = ~KT(0);
};

template <int C, typename KT>
GGKS_HD KT to_key(KT bits) {
  if (C == CONV_F) {
    KT mask = (KT(0) - (bits >> (KeyBits<KT>::W - 1))) | KeyBits<KT>::TOP;
    return bits ^ mask;
  } else if (C == CONV_I) {
    return bits ^ KeyBits<KT>::TOP;
  }
  return bits;
}

template <int C, typename KT>
GGKS_HD KT from_key(KT key) {
  if (C == CONV_F) {
    KT mask = ((key >> (KeyBits<KT>::W - 1)) - KT(1)) | KeyBits<KT>::TOP;
    return key ^ mask;
  } else if (C == CONV_I) {
    return key ^ KeyBits<KT>::TOP;
  }
  return key;
}

// dtype code (ggks.h) -> conversion kind / key width
GGKS_HD int dtype_conv(int dtype) {
  return dtype == 2 ? CONV_U : (dtype == 3 ? CONV_I : CONV_F);
}
GGKS_HD int dtype_bytes(int dtype) { return dtype == 1 ? 8 : 4; }

}  // namespace ggks
