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

enum GenFamily : int {
  GEN_UNIFORM = 0,     // floats U[0,1); integers U[lo, hi]
  GEN_NORMAL = 1,      // N(0,1) (floats)
  GEN_LEVELS = 2,      // floats holding the integers lo..hi
  GEN_CAUCHY = 3,      // tan(pi (u - 1/2)), clipped to finite (floats)
  GEN_PLATEAU = 4,     // 0.5 with probability 1/2, else U[0,1) (floats)
  GEN_SORTED = 5,      // non-decreasing over positions 0..hi-1 (any dtype)
  GEN_KILLER = 6,      // bucket killer: tiny values plus the powers of two (floats)
  GEN_NOISY = 7,       // 1 (p = 0.2) or 0, plus N(0, 0.01) noise (floats)
  GEN_HALF_NORMAL = 8, // |N(0,1)| (floats)
  GEN_HUGE = 9,        // U[0,1) * 10^6 (floats)
  GEN_ONES = 10,       // every value 1
  GEN_ONES_TWOS = 11,  // 1 or 2 with probability 1/2 each
  GEN_NORMAL100 = 12,  // N(0, 100^2) (floats)
  GEN_FAMILIES = 13
};

// Box-Muller N(0,1) in double from two streams of the same seed (the host
// generator's pairing, so only the device libm can differ in the last bit).
__device__ __forceinline__ double gen_normal(uint64_t seed, uint64_t j, uint64_t z) {
  const double u1 = ((double)(z >> 11) + 1.0) * 0x1p-53;
  const double u2 = (double)(splitmix_at(seed ^ 0x5DEECE66Dull, j) >> 11) * 0x1p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// dtype codes as in ggks.h: 0 f32, 1 f64, 2 u32, 3 i32.
//
// Families and the reference shapes they stand in for (generateProblems.cu):
//   sorted     generateSortedArrayUints (:34-45): the reference sorts cuRAND
//              uints; here element j of a length-hi sequence is
//              floor((j * 2^32 + r_j) / hi) with r_j < 2^32 from the stream,
//              non-decreasing in j (i32: minus 2^31; f32 / f64: that times
//              2^-32, so a value is a pure function of (seed, j) and the
//              k-th largest of any window is one element, computable alone)
//   killer     generateBucketKillerFloats / Doubles (:344-367, :385-423):
//              uniform bits masked to tiny magnitudes (f32 0x80EFFFFF, f64
//              0x800FFFFFFFFFFFFF), and positions 0..252 (f64 0..2044) set to
//              the powers of two 2^-126..2^126 (2^-1022..2^1022); the
//              reference also sorts the vector, which no result depends on
//   noisy      generateNoisyVector (:249-277): 20 % ones, 80 % zeros, each
//              plus N(0, 0.01) noise (here a per-element coin, not an exact
//              20 % count)
//   half normal  generateHalfNormalFloats (:322-341)
//   huge       generateHugeUniformFloats (:287-298): U[0,1) * 10^6
//   ones       generateAllOnesFloats (:180-197)
//   ones-twos  generateOnesTwosFloats (:154-177) (per-element coin)
//   normal100  generateNormalFloats100 (:300-306)
__global__ __launch_bounds__(256) void k_generate(void* __restrict__ dst, int dtype, int family,
                                                  uint64_t seed, uint64_t start, uint64_t n,
                                                  int64_t lo, int64_t hi) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t j = start + i;
    const uint64_t z = splitmix_at(seed, j);
    if (family == GEN_SORTED) {
      // floor((j * 2^32 + r) / hi), r = top 32 bits of z: < 2^32 for j < hi
      const unsigned __int128 num = ((unsigned __int128)j << 32) + (z >> 32);
      const uint32_t v = (uint32_t)(num / (unsigned __int128)(uint64_t)hi);
      if (dtype == 2) static_cast<uint32_t*>(dst)[i] = v;
      else if (dtype == 3) static_cast<int32_t*>(dst)[i] = (int32_t)(v ^ 0x80000000u);
      else if (dtype == 0) static_cast<float*>(dst)[i] = (float)((double)v * 0x1p-32);
      else static_cast<double*>(dst)[i] = (double)v * 0x1p-32;
      continue;
    }
    if (family == GEN_ONES || family == GEN_ONES_TWOS) {
      const uint32_t v = family == GEN_ONES ? 1u : 1u + (uint32_t)(z >> 63);
      if (dtype == 2) static_cast<uint32_t*>(dst)[i] = v;
      else if (dtype == 3) static_cast<int32_t*>(dst)[i] = (int32_t)v;
      else if (dtype == 0) static_cast<float*>(dst)[i] = (float)v;
      else static_cast<double*>(dst)[i] = (double)v;
      continue;
    }
    if (dtype == 2 || dtype == 3) {
      const uint64_t span = (uint64_t)(hi - lo) + 1;  // <= 2^32
      const uint64_t v = (uint64_t)lo + (z >> 16) % span;
      static_cast<uint32_t*>(dst)[i] = (uint32_t)v;
      continue;
    }
    if (family == GEN_KILLER) {
      if (dtype == 0) {
        const uint32_t np = 253;  // 2^-126 .. 2^126: exponent field j + 1
        uint32_t b;
        if (j < np) {
          b = (uint32_t)(j + 1) << 23;
        } else {  // a uniform float's bits, masked
          const float u = (float)(z >> 40) * 0x1p-24f;
          b = __float_as_uint(u) & 0x80EFFFFFu;
        }
        static_cast<uint32_t*>(dst)[i] = b;
      } else {
        const uint64_t np = 2045;  // 2^-1022 .. 2^1022: exponent field j + 1
        uint64_t b;
        if (j < np) {
          b = (j + 1) << 52;
        } else {
          const double u = (double)(z >> 11) * 0x1p-53;
          b = (uint64_t)__double_as_longlong(u) & 0x800FFFFFFFFFFFFFull;
        }
        static_cast<uint64_t*>(dst)[i] = b;
      }
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
    } else if (family == GEN_NORMAL || family == GEN_HALF_NORMAL || family == GEN_NORMAL100) {
      x = gen_normal(seed, j, z);
      if (family == GEN_HALF_NORMAL) x = fabs(x);
      if (family == GEN_NORMAL100) x *= 100.0;
    } else if (family == GEN_NOISY) {  // 1 with probability 1/5, else 0, plus N(0, 0.01)
      const double base = (double)(splitmix_at(seed ^ 0x9E3779B9ull, j) >> 11) * 0x1p-53 < 0.2
                              ? 1.0 : 0.0;
      x = base + 0.01 * gen_normal(seed, j, z);
    } else {
      x = dtype == 0 ? (double)(z >> 40) * 0x1p-24 : (double)(z >> 11) * 0x1p-53;
      if (family == GEN_HUGE) x = dtype == 0 ? (double)((float)x * 1.0e6f) : x * 1.0e6;
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
