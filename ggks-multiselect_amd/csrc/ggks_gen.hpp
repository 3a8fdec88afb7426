// ggks_gen.hpp -- seeded synthetic inputs generated on the device.
//
// The same counter-based SplitMix64 stream as the host generators of the
// Python package (workloads.py): element i of a vector generated from
// (seed, start) uses output start + i of the stream of `seed`, so the uniform
// and integer-range families are bit-identical to the host's.  The normal
// family is Box-Muller in double from two streams of the same seed, as on the
// host, but with the device's log / sqrt / cos, so a value can differ from
// the host's in its last bit.  The levels family gives floats with the
// integer values lo..hi (heavy duplicates), as the host's int_range cast to
// float; the Cauchy family is tan(pi (u - 1/2)) clipped to finite values; the
// plateau family is 0.5 with probability 1/2 and uniform otherwise.  -0.0 is canonicalised to +0.0.  Stands in for
// the reference's cuRAND problem generators (generateProblems.cu:25-428),
// whose streams are seeded by the clock and cannot be reproduced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ggks_types.hpp"

namespace ggks {

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

enum GenFamily : int { GEN_UNIFORM = 0, GEN_NORMAL = 1, GEN_LEVELS = 2, GEN_CAUCHY = 3, GEN_PLATEAU = 4 };

// dtype codes as in ggks.h: 0 f32, 1 f64, 2 u32, 3 i32.
__global__ __launch_bounds__(256) void k_generate(void* __restrict__ dst, int dtype, int family,
                                                  uint64_t seed, uint64_t start, uint64_t n,
                                                  int64_t lo, int64_t hi) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t z = splitmix_at(seed, start + i);
    if (dtype == 2 || dtype == 3) {
      const uint64_t span = (uint64_t)(hi - lo) + 1;  // <= 2^32
      const uint64_t v = (uint64_t)lo + (z >> 16) % span;
      static_cast<uint32_t*>(dst)[i] = (uint32_t)v;
      continue;
    }
    double x;
    if (family == GEN_LEVELS) {  // integer values lo..hi as floats: heavy duplicates
      x = (double)((uint64_t)lo + (z >> 16) % ((uint64_t)(hi - lo) + 1));
    } else if (family == GEN_CAUCHY) {  // tan(pi (u - 1/2)), clipped to finite values
      double u = (double)(z >> 11) * 0x1p-53;
      u = u < 0x1p-40 ? 0x1p-40 : (u > 1.0 - 0x1p-40 ? 1.0 - 0x1p-40 : u);
      x = tan(3.141592653589793 * (u - 0.5));
      const double fmax = dtype == 0 ? 3.4028234663852886e38 : 1.7976931348623157e308;
      x = x > fmax ? fmax : (x < -fmax ? -fmax : x);
    } else if (family == GEN_PLATEAU) {  // 0.5 with probability 1/2, else uniform
      x = (z >> 63) ? 0.5 : (dtype == 0 ? (double)((z >> 39) & 0xFFFFFFull) * 0x1p-24
                                        : (double)((z >> 10) & ((1ull << 53) - 1)) * 0x1p-53);
    } else if (family == GEN_NORMAL) {
      const double u1 = ((double)(z >> 11) + 1.0) * 0x1p-53;
      const double u2 = (double)(splitmix_at(seed ^ 0x5DEECE66Dull, start + i) >> 11) * 0x1p-53;
      x = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    } else {
      x = dtype == 0 ? (double)(z >> 40) * 0x1p-24 : (double)(z >> 11) * 0x1p-53;
    }
    if (dtype == 0) {
      float f = (float)x;
      static_cast<float*>(dst)[i] = f == 0.0f ? 0.0f : f;
    } else {
      static_cast<double*>(dst)[i] = x == 0.0 ? 0.0 : x;
    }
  }
}

}  // namespace ggks
