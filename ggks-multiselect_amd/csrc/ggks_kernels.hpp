// ggks_kernels.hpp -- the gfx950 kernels of the selection engine.
//
// Pipeline for one level over a segment table (all on the device, no host
// round trip):
//   k_sample  gather a seeded random sample of each segment's keys
//   k_lut     build each segment's bucket function: a prefix look-up table
//             (key - kb) >> s -> (first bucket, buckets in prefix); inside a
//             prefix buckets split the key range evenly, so a heavy prefix is
//             cut finely and an empty run of prefixes shares one bucket
//   k_count   pass 1 over HBM: bucket every key, LDS histogram per chunk
//   k_reduce  sum chunk rows per bucket; turn rows into per-chunk offsets
//   k_findk   per segment: scan the totals, find each rank's bucket, answer
//             single-key buckets at once, mark the rest active, lay out one
//             child per active bucket and group the ranks by child
//   k_compact pass 2 over HBM: copy only keys of active buckets, each chunk
//             into its private slice of the child's range
//   k_finish_small  per child with few ranks: top-digit histogram, then the
//             rank's digit gathered and ranked (keys in registers when they
//             fit); per child of up to 1024 keys with many ranks: LDS sort
//   k_finish  every other child: multi-rank MSD radix select on 8-bit digits
//             of (key - lo), or a sort, keys resident in LDS when they fit
// Plans of at most 4096 buckets per segment run k_findk_small and
// k_compact_small: the same bodies with smaller tables, two work-groups per
// CU.  Plans whose children are small run k_finish_small_lite (4-byte keys):
// the same body with fewer keys in registers, six work-groups per CU.
//
// Reference counterparts: bucketMultiselect.cu phaseOne (:521-833) --
// assignSmartBucket (:129-205) -> k_count, sumCounts (:214-221) -> k_reduce,
// findKBuckets + host reindex (:87-111, :714-730) + reindexCounts (:229-241)
// -> k_findk, copyElements (:249-308) -> k_compact, thrust::sort +
// copyValuesInChunk (:799-818) -> k_finish.  The reference's pivots
// (:414-508) are replaced by the sampled prefix LUT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ggks_keys.hpp"
#include "ggks_types.hpp"

namespace ggks {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Inclusive prefix sum across a wavefront with DPP moves (no LDS traffic):
// row_shr 1, 2, 4, 8 scan each row of 16 lanes, then row_bcast 15 and 31
// carry the row totals into rows 1, 3 and 2, 3.  Lanes without a source
// receive 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// LDS atomic add of 1 for every active lane on h[key], returning each
// lane's old value.  When all active lanes of the wavefront name the same
// key (all of top-K's ranks in one child), one lane adds their number and
// the lanes take consecutive values, instead of 64 atomics on one address.
__device__ __forceinline__ uint32_t lds_inc(uint32_t* h, uint32_t key) {
  const uint64_t act = __ballot(1);
  const uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)key);
  if (__ballot(key == k0) == act) {  // uniform
    const uint64_t below = act & ((1ull << (threadIdx.x & 63u)) - 1ull);
    uint32_t base = 0;
    if (below == 0) base = atomicAdd(&h[k0], (uint32_t)__popcll(act));
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    return base + (uint32_t)__popcll(below);
  }
  return atomicAdd(&h[key], 1u);
}
// LDS atomic OR of v into w[idx]; one atomic for the wavefront when every
// active lane names the same word and value.
__device__ __forceinline__ void lds_or(uint32_t* w, uint32_t idx, uint32_t v) {
  const uint64_t act = __ballot(1);
  const uint32_t i0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
  const uint32_t v0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
  if (__ballot(idx == i0 && v == v0) == act) {  // uniform
    if ((act & ((1ull << (threadIdx.x & 63u)) - 1ull)) == 0) atomicOr(&w[i0], v0);
    return;
  }
  atomicOr(&w[idx], v);
}

// Value of lane 63 (the wavefront total after an inclusive scan), uniform.
__device__ __forceinline__ uint32_t wave_last(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Block-wide exclusive scan of one value per thread (blockDim.x == 1024).
// wsum needs 17 words of LDS.  Ends with a barrier.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum,
                                                    uint32_t& total) {
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (w == 0) {
    uint32_t s = lane < 16 ? wsum[lane] : 0u;
    uint32_t si = wave_incl_scan(s);
    if (lane < 16) wsum[lane] = si - s;
    if (lane == 15) wsum[16] = si;
  }
  __syncthreads();
  uint32_t res = inc - v + wsum[w];
  total = wsum[16];
  __syncthreads();
  return res;
}

// In-place exclusive scan of a[0, L) in LDS by one 1024-thread block.  Each
// wavefront owns a contiguous segment and walks it 64 consecutive words at a
// time (no bank conflicts); wavefront totals are scanned once.  wsum needs 17
// words.  Returns the total.
__device__ uint32_t block_scan_array(uint32_t* a, uint32_t L, uint32_t* wsum) {
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  constexpr uint32_t NW = kThreads / 64;
  const uint32_t seg = ((L + NW - 1) / NW + 63) / 64 * 64;
  const uint32_t b0 = w * seg, b1 = min(L, b0 + seg);
  uint32_t carry = 0;
  constexpr uint32_t U = 8;  // independent wavefront scans in flight
  for (uint32_t base = b0; base < b1; base += 64 * U) {
    uint32_t v[U], inc[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = base + u * 64 + lane;
      v[u] = i < b1 ? a[i] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) inc[u] = wave_incl_scan(v[u]);
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = base + u * 64 + lane;
      if (i < b1) a[i] = carry + inc[u] - v[u];
      carry += wave_last(inc[u]);
    }
  }
  if (lane == 0) wsum[w] = carry;
  __syncthreads();
  if (w == 0) {
    uint32_t t = lane < NW ? wsum[lane] : 0u;
    uint32_t ti = wave_incl_scan(t);
    if (lane < NW) wsum[lane] = ti - t;
    if (lane == NW - 1) wsum[NW] = ti;
  }
  __syncthreads();
  const uint32_t add = wsum[w];
  for (uint32_t i = b0 + lane; i < b1; i += 64) a[i] += add;
  const uint32_t total = wsum[NW];
  __syncthreads();
  return total;
}

// f(i, p[i]) for i = threadIdx.x, +kThreads, ... < n with 16 loads in flight
// per thread (the loops that read another kernel's output are latency bound).
template <uint32_t D = 16, typename T, typename F>
__device__ __forceinline__ void strided_for(const T* __restrict__ p, uint32_t n, F&& f) {
  uint32_t i = threadIdx.x;
  for (; i + (D - 1) * kThreads < n; i += D * kThreads) {
    T v[D];
#pragma unroll
    for (uint32_t u = 0; u < D; ++u) v[u] = p[i + u * kThreads];
#pragma unroll
    for (uint32_t u = 0; u < D; ++u) f(i + u * kThreads, v[u]);
  }
  for (; i < n; i += kThreads) f(i, p[i]);
}

// Ascending bitonic sort of keys[0, np) in LDS (np a power of two) by a
// work-group of NT threads.  In every stage a thread owns up to MAXP of the
// np / 2 disjoint compare-exchange pairs: it loads all of them, then
// compares and stores, so its LDS round trips overlap.
template <uint32_t NT, uint32_t MAXP, typename KT>
__device__ __forceinline__ void lds_bitonic_sort(KT* keys, uint32_t np) {
  for (uint32_t kk = 2; kk <= np; kk <<= 1) {
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      KT a[MAXP], b[MAXP];
#pragma unroll
      for (uint32_t u = 0; u < MAXP; ++u) {
        const uint32_t pr = threadIdx.x + u * NT;
        if (pr < np / 2) {
          const uint32_t i = ((pr & ~(j - 1)) << 1) | (pr & (j - 1));
          a[u] = keys[i];
          b[u] = keys[i | j];
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < MAXP; ++u) {
        const uint32_t pr = threadIdx.x + u * NT;
        if (pr < np / 2) {
          const uint32_t i = ((pr & ~(j - 1)) << 1) | (pr & (j - 1));
          if ((a[u] > b[u]) == ((i & kk) == 0)) {
            keys[i] = b[u];
            keys[i | j] = a[u];
          }
        }
      }
      __syncthreads();
    }
  }
}

// Ascending LSD radix sort of keys[0, np) in LDS by a work-group of
// kThreads (np a multiple of kThreads, at most S * kThreads), on the low
// nbits bits (every key below 2^nbits, or padding equal to the all-ones key,
// which sorts last), 8-bit digits.  The keys live in registers between
// passes: wavefront w holds the tile [w * np / 16, (w + 1) * np / 16), lane
// l its keys w * np / 16 + 64 s + l; each pass counts digits per
// (digit, wavefront) in whist (256 * 16 words), scans the counts in that
// order (so equal digits keep tile order: the sort is stable) and scatters
// every key to its place in LDS, which the next pass reads back.  A key's
// place among its wavefront's equal digits in one step is the number of
// lower lanes with the same digit, from eight ballots over the digit's bits.
// Far fewer barriers than a bitonic sort (3 per pass against 91 stages for
// 8 Ki keys).
template <typename KT, uint32_t S>
__device__ void lds_radix_sort(KT* keys, uint32_t np, int nbits, uint32_t* whist,
                               uint32_t* wsum) {
  constexpr uint32_t NW = kThreads / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const uint32_t steps = np / kThreads;  // keys per lane
  const uint32_t tile = w * (np / NW);
  const uint64_t below = (1ull << lane) - 1ull;
  KT k[S];
#pragma unroll
  for (uint32_t t = 0; t < S; ++t)
    if (t < steps) k[t] = keys[tile + t * 64 + lane];
  for (int sh = 0; sh < nbits; sh += 8) {
    for (uint32_t i = tid; i < 256 * NW; i += kThreads) whist[i] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t t = 0; t < S; ++t)
      if (t < steps) atomicAdd(&whist[(uint32_t)((k[t] >> sh) & 0xFF) * NW + w], 1u);
    __syncthreads();
    block_scan_array(whist, 256 * NW, wsum);  // whist[d * NW + w] = first place (synchronises)
#pragma unroll
    for (uint32_t t = 0; t < S; ++t) {
      if (t < steps) {  // uniform
        const uint32_t d = (uint32_t)((k[t] >> sh) & 0xFF);
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const uint64_t m = __ballot((d >> b) & 1u);
          peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        uint32_t off = 0;
        if (rank == 0) {  // the lowest lane of the digit's peers claims their places
          off = whist[d * NW + w];
          whist[d * NW + w] = off + (uint32_t)__popcll(peers);
        }
        off = (uint32_t)__shfl((int)off, __ffsll((unsigned long long)peers) - 1, 64);
        keys[off + rank] = k[t];
      }
    }
    __syncthreads();
    if (sh + 8 < nbits) {
#pragma unroll
      for (uint32_t t = 0; t < S; ++t)
        if (t < steps) k[t] = keys[tile + t * 64 + lane];
    }
  }
  __syncthreads();
}

template <typename KT>
__device__ __forceinline__ uint32_t mulhi_nb(KT x, uint32_t nb);
template <>
__device__ __forceinline__ uint32_t mulhi_nb<uint32_t>(uint32_t x, uint32_t nb) {
  return __umulhi(x, nb);
}
template <>
__device__ __forceinline__ uint32_t mulhi_nb<uint64_t>(uint64_t x, uint32_t nb) {
  return (uint32_t)__umul64hi(x, (uint64_t)nb);
}

template <typename KT>
__device__ __forceinline__ int bitlen(KT v) {
  return v == 0 ? 0 : (sizeof(KT) == 8 ? 64 - __clzll((long long)v) : 32 - __clz((int)v));
}

// The bucket function of a segment (field layout in SegState).  Every step
// is a select, min/max, bit-field extract or a full-rate 24-bit multiply, so
// a key costs about a dozen VALU instructions and one LDS read.  It is
// monotone in the key: the prefix is, and inside a prefix the sub-bucket
// floor(t * nb / 2^24) is monotone in t.  Keys below the sampled minimum or
// above the last prefix land in the first or last bucket.
template <typename KT>
struct BucketFn {
  KT kb, lmax, hv;
  uint32_t poff, pwid, rsh, twid, lsh, hvon;
  __device__ __forceinline__ void init(const SegState& st) {
    kb = (KT)st.kb;
    lmax = (KT)st.lmax;
    poff = st.poff;
    pwid = st.pwid;
    rsh = st.rsh;
    twid = st.twid;
    lsh = st.lsh;
    hv = (KT)st.hv;
    hvon = st.hvon;
  }
  // F from L for a segment with a heavy value: + [key >= hv] + [key > hv]
  __device__ __forceinline__ uint32_t heavy(uint32_t b, KT key) const {
    return b + (key >= hv ? 1u : 0u) + (key > hv ? 1u : 0u);
  }
  // stage 1: LUT index p and the normalised sub-field t of a key
  __device__ __forceinline__ void split(KT key, uint32_t& p, uint32_t& t) const {
    KT low = (key > kb ? key : kb) - kb;
    low = low < lmax ? low : lmax;
    if (sizeof(KT) == 4) {
      p = __builtin_amdgcn_ubfe((uint32_t)low, poff, pwid);
      t = __builtin_amdgcn_ubfe((uint32_t)low, rsh, twid) << lsh;
    } else {
      const uint64_t l64 = (uint64_t)low;
      p = (uint32_t)(l64 >> poff) & ((1u << pwid) - 1u);
      t = ((uint32_t)(l64 >> rsh) & ((1u << twid) - 1u)) << lsh;
    }
  }
  // stage 1 with one shift for t: finish() multiplies only the low 24 bits
  // of t, and those bits of low << lsh (sub-fields of s <= 24 bits, rsh = 0)
  // or of low >> rsh (s > 24, lsh = 0) equal the normalised sub-field, so
  // the bucket is the same as split()'s.  LEFT must be s <= 24.
  template <bool LEFT>
  __device__ __forceinline__ void split1(KT key, uint32_t& p, uint32_t& t) const {
    KT low = (key > kb ? key : kb) - kb;
    low = low < lmax ? low : lmax;
    if (sizeof(KT) == 4) {
      p = __builtin_amdgcn_ubfe((uint32_t)low, poff, pwid);
      t = LEFT ? (uint32_t)low << lsh : (uint32_t)low >> rsh;
    } else {
      const uint64_t l64 = (uint64_t)low;
      p = (uint32_t)(l64 >> poff) & ((1u << pwid) - 1u);
      t = LEFT ? (uint32_t)l64 << lsh : (uint32_t)(l64 >> rsh);
    }
  }
  // stage 2: bucket from the LUT entry e = lut[p]
  __device__ __forceinline__ uint32_t finish(uint32_t e, uint32_t t) const {
    const uint32_t nbsh = (e >> 16) << 8;  // nb * 2^8 < 2^24
    const uint32_t j = (uint32_t)(((uint64_t)(t & 0xFFFFFFu) * (nbsh & 0xFFFFFFu)) >> 32);
    return (e & 0xFFFFu) + j;
  }
  __device__ __forceinline__ uint32_t operator()(KT key, const uint32_t* lut) const {
    uint32_t p, t;
    split(key, p, t);
    const uint32_t b = finish(lut[p], t);
    return hvon ? heavy(b, key) : b;
  }
};

// Bit c of an LDS bitmap.  v_bfe_u32 (and the intrinsic's definition) take
// the bit offset modulo 32, so c needs no mask.
__device__ __forceinline__ uint32_t bitmap_test(const uint32_t* bm, uint32_t c) {
  return __builtin_amdgcn_ubfe(bm[c >> 5], c, 1u);
}

// Buckets of N keys with the N look-ups issued back to back, so a wavefront
// waits on LDS once per batch rather than once per key.
template <typename KT, int C, int N>
__device__ __forceinline__ void bucket_batch(const BucketFn<KT>& fn, const uint32_t* lut,
                                             const KT (&raw)[N], uint32_t (&b)[N]) {
  uint32_t p[N], t[N], e[N];
#pragma unroll
  for (int i = 0; i < N; ++i) fn.split(to_key<C>(raw[i]), p[i], t[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) e[i] = lut[p[i]];
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = fn.finish(e[i], t[i]);
  if (fn.hvon) {  // uniform
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = fn.heavy(b[i], to_key<C>(raw[i]));
  }
}

// bucket_batch for the streaming passes: split1 with the shift direction
// and the heavy-value term fixed at compile time (each pass runs its loop in
// the instance chosen by the segment's state, so the loop body has no
// branch).
template <typename KT, int C, int N, bool LEFT, bool HV = false>
__device__ __forceinline__ void bucket_batch1(const BucketFn<KT>& fn, const uint32_t* lut,
                                              const KT (&raw)[N], uint32_t (&b)[N]) {
  uint32_t p[N], t[N], e[N];
#pragma unroll
  for (int i = 0; i < N; ++i) fn.template split1<LEFT>(to_key<C>(raw[i]), p[i], t[i]);
#pragma unroll
  for (int i = 0; i < N; ++i) e[i] = lut[p[i]];
#pragma unroll
  for (int i = 0; i < N; ++i) b[i] = fn.finish(e[i], t[i]);
  if (HV) {
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = fn.heavy(b[i], to_key<C>(raw[i]));
  }
}

// Streaming 16-byte load of input data: non-temporal, so the read stream
// does not displace lines the kernel reuses (LUT, histogram rows, outputs).
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 stream_load(const uint4* p) {
  const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Unpack R 16-byte vectors into their keys.
template <typename KT, int R>
__device__ __forceinline__ void unpack(const uint4 (&v)[R], KT (&k)[R * 16 / sizeof(KT)]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (sizeof(KT) == 4) {
      k[r * 4 + 0] = (KT)v[r].x; k[r * 4 + 1] = (KT)v[r].y;
      k[r * 4 + 2] = (KT)v[r].z; k[r * 4 + 3] = (KT)v[r].w;
    } else {
      k[r * 2 + 0] = (KT)(((uint64_t)v[r].y << 32) | v[r].x);
      k[r * 2 + 1] = (KT)(((uint64_t)v[r].w << 32) | v[r].z);
    }
  }
}

// Key range [lo, hi] of a bucket, found by binary search over keys with the
// same device function the passes use, so it can never disagree with them.
template <typename KT>
struct RangeOf {
  BucketFn<KT> fn;
  const uint32_t* lut;
  uint32_t B, P, s;
  uint32_t hb;  // F's bucket of the heavy value (fn.hvon)
  // Closed-form candidate for the smallest key of F-bucket b: with a heavy
  // value, buckets up to hb - 1 are L's, hb is {hv}, hb + 1 starts at
  // hv + 1 and the rest are L's shifted by two.
  __device__ bool candidate(uint32_t b, KT& out) const {
    if (fn.hvon) {
      if (b == hb) { out = fn.hv; return true; }
      if (b == hb + 1) {
        if (fn.hv == KeyBits<KT>::MAX) return false;
        out = (KT)(fn.hv + 1);
        return true;
      }
      if (b > hb + 1) b -= 2;
    }
    return candidate_l(b, out);
  }
  // Closed-form candidate for the smallest key of L-bucket b: find the
  // prefix p holding b, then the smallest t with floor(t * nb / 2^24) >= j
  // and the smallest sub-field value giving that t.
  __device__ bool candidate_l(uint32_t b, KT& out) const {
    uint32_t lo = 0, hi = P - 1;  // smallest prefix whose bucket range ends past b
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      uint32_t e = lut[mid];
      uint32_t nb = e >> 16;
      if ((e & 0xFFFFu) + (nb > 1 ? nb : 1u) > b) hi = mid; else lo = mid + 1;
    }
    const uint32_t e = lut[lo];
    const uint32_t nb = e >> 16, j = b - (e & 0xFFFFu);
    uint64_t sub = 0;
    if (j > 0 && nb > 1) {
      const uint64_t T = (((uint64_t)j << 24) + nb - 1) / nb;  // smallest t
      sub = fn.rsh ? (T << fn.rsh) : ((T + (1ull << fn.lsh) - 1) >> fn.lsh);
    }
    const unsigned __int128 low = ((unsigned __int128)lo << s) + sub;
    if (low > (unsigned __int128)fn.lmax) return false;
    out = (KT)((KT)low + fn.kb);
    return true;
  }
  // smallest key whose bucket is >= b, or false if no key reaches b
  __device__ bool first_key(uint32_t b, KT& out) const {
    if (b == 0) { out = 0; return true; }
    KT c;
    if (candidate(b, c) && fn(c, lut) >= b && (c == 0 || fn((KT)(c - 1), lut) < b)) {
      out = c;
      return true;
    }
    if (fn(KeyBits<KT>::MAX, lut) < b) return false;
    KT lo = 0, hi = KeyBits<KT>::MAX;
    while (lo < hi) {
      KT mid = lo + ((hi - lo) >> 1);
      if (fn(mid, lut) >= b) hi = mid; else lo = mid + 1;
    }
    out = lo;
    return true;
  }
  __device__ void operator()(uint32_t b, KT& lo, KT& hi) const {
    first_key(b, lo);
    KT nxt;
    hi = (b + 1 < B && first_key(b + 1, nxt)) ? (KT)(nxt - 1) : KeyBits<KT>::MAX;
  }
};

// Visit every element of p[0, len) with f(raw), 16-byte vector loads for the
// aligned body, four loads in flight per thread.
template <typename KT, typename F>
__device__ __forceinline__ void for_each_elem(const KT* __restrict__ p, uint64_t len, F&& f) {
  constexpr uint32_t VEC = 16 / sizeof(KT);
  const uint32_t tid = threadIdx.x;
  uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
  uint64_t head = mis ? (VEC - mis) : 0;
  if (head > len) head = len;
  if (tid < head) f(p[tid]);
  const uint64_t nvec = (len - head) / VEC;
  const uint4* __restrict__ pv = reinterpret_cast<const uint4*>(p + head);
  uint64_t i = tid;
  auto apply = [&](const uint4& v) {
    if (sizeof(KT) == 4) {
      f((KT)v.x); f((KT)v.y); f((KT)v.z); f((KT)v.w);
    } else {
      f((KT)(((uint64_t)v.y << 32) | v.x));
      f((KT)(((uint64_t)v.w << 32) | v.z));
    }
  };
  for (; i + 3 * kThreads < nvec; i += 4 * kThreads) {
    uint4 a = pv[i], b = pv[i + kThreads], c = pv[i + 2 * kThreads], d = pv[i + 3 * kThreads];
    apply(a); apply(b); apply(c); apply(d);
  }
  for (; i < nvec; i += kThreads) apply(pv[i]);
  const uint64_t done = head + nvec * VEC;
  if (done + tid < len) f(p[done + tid]);
}

// ----------------------------------------------------------------- kernels

// Query list from a device k list: rank = n - k + 1 (k-th largest -> 1-based
// ascending rank, bucketMultiselect.cu:845-846).  Bad k raises a status bit
// and is clamped so every later index stays in bounds.
__global__ void k_queries(const uint32_t* __restrict__ k, uint32_t nk, uint64_t n,
                          Query* __restrict__ q, uint32_t* __restrict__ status) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  uint32_t kk = k[i];
  if (kk < 1 || (uint64_t)kk > n) {
    atomicOr(status, (uint32_t)STATUS_BAD_K);
    kk = 1;
  }
  q[i].rank = (uint32_t)(n - kk + 1);
  q[i].out = i;
}

// Top-K as ranks: k = 1..K (the reference's bucket top-k,
// multiselectTimingFunctions.cu:204-231), so out[j] is the (j+1)-th largest.
__global__ void k_topk_queries(uint32_t K, uint64_t n, Query* __restrict__ q) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  q[i].rank = (uint32_t)(n - i);
  q[i].out = i;
}

// Seeded random sample of each segment: one work-group per block of 1024
// samples of a segment (the planner's dense table; every sample count is a
// multiple of 1024), kSampleThreads threads of kSamplePer independent
// gathers each.  Each work-group also writes a record of four keys for its
// 1024 samples: their min and max (so the LUT kernel reads 64 records
// instead of scanning the sample for its range), and a heavy-value
// candidate with its count: every wavefront proposes its first sample, the
// block counts each proposal's copies among its 1024 samples, and keeps the
// most frequent (the smaller key on a tie).
constexpr uint32_t kSampleThreads = kThreads;  // one sample a thread
constexpr uint32_t kSamplePer = kThreads / kSampleThreads;
constexpr uint32_t kSampleWaves = kSampleThreads / 64;
template <typename KT, int C>
__global__ __launch_bounds__(kSampleThreads) void k_sample(const KT* __restrict__ src,
                                                           const SegDesc* __restrict__ segs,
                                                           const SampleBlock* __restrict__ sblocks,
                                                           KT* __restrict__ samples,
                                                           KT* __restrict__ minmax, uint64_t seed,
                                                           SegState* __restrict__ states,
                                                           uint32_t* __restrict__ lut_pool,
                                                           uint32_t* __restrict__ hgram) {
  __shared__ KT red_lo[kSampleWaves], red_hi[kSampleWaves], cand[kSampleWaves];
  __shared__ uint32_t hl[1u << kMaxPrefixBits];  // the block's keys by their top pbits bits
  __shared__ uint32_t ccnt[kSampleWaves][kSampleWaves];  // [wave][candidate]
  const SampleBlock sb = sblocks[blockIdx.x];
  const SegDesc& sd = segs[sb.seg];
  if (sd.mode == MODE_RANGE || sb.blk * kThreads >= sd.nsample) return;  // uniform
  if (sb.blk == 0 && lut_parts(sd.mode, sd.nsample, sizeof(KT)) > 1) {
    // the split LUT build adds its partial sample histograms into the LUT
    // region and counts its parts in the state: both start at zero
    for (uint32_t i = threadIdx.x; i < (1u << sd.pbits); i += kSampleThreads) lut_pool[sd.lut_off + i] = 0;
    if (threadIdx.x == 0) states[sb.seg].arrive = 0;
  }
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  KT k[kSamplePer];
#pragma unroll
  for (uint32_t j = 0; j < kSamplePer; ++j) {  // sample i of the segment
    const uint32_t i = sb.blk * kThreads + j * kSampleThreads + threadIdx.x;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)sb.seg * 0x100000001ull + i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    k[j] = src[sd.data_off + __umul64hi(z, sd.len)];
  }
  // With a histogram (one sampled segment) the sample is also counted by
  // the top pbits bits of the key, into hgram: where the sampled range spans
  // the top bit (full-width ranges, such as floats of both signs) the LUT's
  // prefixes are exactly these bins, and k_lut reads 2^pbits counts instead
  // of histogramming the whole sample on one CU.
  const bool hg = hgram != nullptr && sd.mode == MODE_SAMPLE;  // uniform
  const uint32_t hsh = KeyBits<KT>::W - sd.pbits;
  if (hg)  // zeroed while the gathers are in flight
    for (uint32_t i = threadIdx.x; i < (1u << sd.pbits); i += kSampleThreads) hl[i] = 0;
  KT lo = KeyBits<KT>::MAX, hi = 0;
#pragma unroll
  for (uint32_t j = 0; j < kSamplePer; ++j) {
    const KT x = to_key<C>(k[j]);
    k[j] = x;
    samples[sd.sample_off + sb.blk * kThreads + j * kSampleThreads + threadIdx.x] = x;
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
  }
  if (lane == 0) cand[w] = k[0];
  for (int d = 32; d >= 1; d >>= 1) {
    const KT a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (lane == 0) {
    red_lo[w] = lo;
    red_hi[w] = hi;
  }
  __syncthreads();  // cand, red_lo / red_hi and the zeroed hl
  if (hg) {
#pragma unroll
    for (uint32_t j = 0; j < kSamplePer; ++j) atomicAdd(&hl[(uint32_t)(k[j] >> hsh)], 1u);
  }
  // copies of every candidate in this wavefront's samples (ballot counts)
#pragma unroll
  for (uint32_t c = 0; c < kSampleWaves; ++c) {
    const KT cv = cand[c];
    uint32_t m = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSamplePer; ++j) m += (uint32_t)__popcll(__ballot(k[j] == cv));
    if (lane == 0) ccnt[w][c] = m;
  }
  __syncthreads();
  if (w == 0) {
    // lane c < kSampleWaves: candidate c's copies in the block, and the
    // block's min and max; the most frequent candidate (the smaller key on
    // a tie) by a wavefront reduction
    uint32_t t = 0;
    KT cv = KeyBits<KT>::MAX, l = KeyBits<KT>::MAX, h = 0;
    if (lane < kSampleWaves) {
      for (uint32_t v = 0; v < kSampleWaves; ++v) t += ccnt[v][lane];
      cv = cand[lane];
      l = red_lo[lane];
      h = red_hi[lane];
    }
    for (int d = 32; d >= 1; d >>= 1) {
      const KT la = __shfl_xor(l, d, 64), ha = __shfl_xor(h, d, 64);
      l = la < l ? la : l;
      h = ha > h ? ha : h;
      const uint32_t ot = (uint32_t)__shfl_xor((int)t, d, 64);
      const KT oc = __shfl_xor(cv, d, 64);
      if (ot > t || (ot == t && oc < cv)) {
        t = ot;
        cv = oc;
      }
    }
    if (lane == 0) {
      KT* mm = minmax + 4 * (sd.sample_off / kThreads + sb.blk);
      mm[0] = l;
      mm[1] = h;
      mm[2] = cv;
      mm[3] = (KT)t;
    }
  }
  if (hg) {  // hl is complete: the barrier above follows every count
    for (uint32_t i = threadIdx.x; i < (1u << sd.pbits); i += kSampleThreads)
      if (hl[i]) atomicAdd(&hgram[i], hl[i]);
  }
}

// Finisher shapes that k_findk plans for.  k_finish_small takes children of
// at most kWaveRanks ranks with a work-group of kFThreads.
constexpr uint32_t kWaveRanks = 4;
constexpr uint64_t kWaveMaxLen = 1ull << 15;  // longest child k_finish_small takes
// Children with few ranks are finished by k_finish_small (one 256-thread
// work-group each); the others by k_finish (a 1024-thread work-group with
// the child in LDS).
__host__ __device__ __forceinline__ bool wave_child(uint64_t len, uint32_t q_cnt) {
  return q_cnt <= kWaveRanks && len <= kWaveMaxLen;
}
__device__ __forceinline__ bool wave_child(const Child& c) { return wave_child(c.len, c.q_cnt); }
// Children with more than kWaveRanks ranks and at most kSortCap keys (top-K
// buckets, mostly) are sorted by k_finish_small too; k_finish takes what
// fits neither.
constexpr uint32_t kSortCap = 1024;
__host__ __device__ __forceinline__ bool sort_child(uint64_t len, uint32_t q_cnt) {
  return !wave_child(len, q_cnt) && len <= kSortCap;
}
__device__ __forceinline__ bool sort_child(const Child& c) { return sort_child(c.len, c.q_cnt); }
// A child with ranks that neither fits: k_finish's.
__host__ __device__ __forceinline__ bool big_child(uint64_t len, uint32_t q_cnt) {
  return q_cnt != 0 && !wave_child(len, q_cnt) && !sort_child(len, q_cnt);
}
constexpr uint32_t kFThreads = 256;
// 16-byte words per thread held by k_finish_small: 36 KB of 4-byte keys,
// 32 KB of 8-byte keys in registers (the most that fit in 128 VGPRs without
// spilling for every key type), and kFLdsWords more per thread in LDS (28 KB
// per work-group: with the rest of its LDS, four work-groups still share a
// CU), so children up to 64 KB of 4-byte keys and 60 KB of 8-byte keys are
// read from memory once.  At 2^28 float32 keys with 2^15 buckets about one
// child in eight is longer than the registers alone hold (mean 8.4 K keys,
// largest 13 K); at 2^27 doubles about half (mean 4.5 K).
template <typename KT>
constexpr int kFRegWords = sizeof(KT) == 4 ? 9 : 8;
constexpr int kFLdsWords = 7;
// ... and by k_finish_small_lite (4-byte keys): 4 words (16 KB of keys) in at
// most 80 VGPRs and none in LDS, so that six work-groups share a CU instead
// of four (plans whose children are small: batches of many segments,
// mostly).  The 8-byte instance would spill at that size and is not built.
constexpr int kFLiteWords = 4;
template <typename KT>
constexpr uint64_t finish_small_reg_keys(bool lite) {
  return (uint64_t)(lite ? kFLiteWords : kFRegWords<KT> + kFLdsWords) * kFThreads *
         (16 / sizeof(KT));
}

// A sampled key is a heavy value when it is at least this many of one
// block's kThreads samples (3.1 %): it then gets a bucket of its own.
constexpr uint32_t kHeavyMinCount = 32;

// Build one segment's bucket function (one block per segment).
//  MODE_SAMPLE  prefixes of 2^s keys over the sampled range, sub-buckets in
//               proportion to each prefix's share of the sample; a run of
//               empty prefixes shares one bucket.
//  MODE_RADIX   MSD radix digits: the digit is the next radix_bits bits below
//               the longest common prefix of the sampled keys (the digit
//               passes the reference skips when every key shares the digit,
//               RadixSelect_api.h:590-592, are skipped up front), each
//               digit value exactly its buckets; keys outside the sampled
//               window fall into a guard bucket below or above every digit.
//  MODE_RANGE   equal-width buckets over the segment's known key range.
// A heavy value of the sample (SegState) takes two more buckets.
template <typename KT>
__global__ __launch_bounds__(kThreads) void k_lut(const SegDesc* __restrict__ segs,
                                                  const KT* __restrict__ samples,
                                                  const KT* __restrict__ minmax,
                                                  SegState* __restrict__ states,
                                                  uint32_t* __restrict__ lut_pool,
                                                  uint32_t* __restrict__ l2count,
                                                  uint32_t* __restrict__ bigw,
                                                  uint32_t* __restrict__ hgram) {
  // the second level's counters (heavy children, chunks) start at zero in
  // every run; k_findk, which counts them, runs after this kernel
  if (l2count && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2) l2count[threadIdx.x] = 0;
  // k_findk counts the children k_finish takes here
  if (bigw && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *bigw = 0;
  __shared__ uint32_t cnt[1u << kMaxPrefixBits];
  // the sample histogram in four copies, prefix p's in words 4p .. 4p + 3,
  // a lane counting into copy lane & 3: samples concentrate in a few
  // prefixes (the dense middle of a normal), and lanes of one instruction
  // that hit the same prefix then spread over four addresses instead of
  // serialising on one
  __shared__ __align__(16) uint32_t cnt4[4u << kMaxPrefixBits];
  __shared__ uint32_t wsum[17];
  __shared__ KT red_lo[16], red_hi[16];
  __shared__ KT sh_hv;
  __shared__ uint32_t sh_hvon, sh_last;
  constexpr uint32_t W = KeyBits<KT>::W;
  const SegDesc sd = segs[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  // k_sample's prefix counts, loaded now (16 bytes a thread) so that they
  // arrive while the sample's range is reduced; used for full-width ranges
  const bool hload = hgram && sd.mode == MODE_SAMPLE && tid < (1u << sd.pbits) / 4;
  const uint4 h4 = hload ? reinterpret_cast<const uint4*>(hgram)[tid] : make_uint4(0, 0, 0, 0);
  // part blockIdx.y of the segment's LUT build: its parts, at most the
  // grid's y extent (1 keeps every segment's build in one work-group)
  const uint32_t nparts = min(lut_parts(sd.mode, sd.nsample, sizeof(KT)), gridDim.y), part = blockIdx.y;
  if (part >= nparts) return;  // uniform
  uint32_t* lut = lut_pool + sd.lut_off;
  KT kmin, kmax;
  uint32_t pb = sd.pbits;
  uint32_t hvon = 0;
  KT hv = 0;

  if (sd.mode == MODE_SAMPLE || sd.mode == MODE_RADIX) {
    KT lo = KeyBits<KT>::MAX, hi = 0, hc = 0;
    uint32_t hn = 0;
    const uint32_t npairs = sd.nsample / kThreads;  // <= 64: wavefront 0 holds them all
    if (tid < npairs) {
      const KT* mm = minmax + 4 * (sd.sample_off / kThreads + tid);
      lo = mm[0];
      hi = mm[1];
      hc = mm[2];
      hn = (uint32_t)mm[3];
    }
    for (int d = 32; d >= 1; d >>= 1) {
      KT a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
      const KT oc = __shfl_xor(hc, d, 64);
      const uint32_t on = (uint32_t)__shfl_xor((int)hn, d, 64);
      if (on > hn || (on == hn && oc < hc)) {
        hn = on;
        hc = oc;
      }
    }
    if (lane_id() == 0) { red_lo[tid >> 6] = lo; red_hi[tid >> 6] = hi; }
    if (tid == 0) {
      sh_hvon = hn >= kHeavyMinCount ? 1u : 0u;
      sh_hv = hc;
    }
    __syncthreads();
    kmin = red_lo[0]; kmax = red_hi[0];
    for (int w = 1; w < 16; ++w) {
      kmin = red_lo[w] < kmin ? red_lo[w] : kmin;
      kmax = red_hi[w] > kmax ? red_hi[w] : kmax;
    }
    hvon = sh_hvon;
    hv = sh_hv;
  } else {
    kmin = (KT)sd.lo;
    kmax = (KT)sd.hi;
  }
  const KT R = kmax - kmin;
  const int bl = bitlen<KT>(R);
  KT kb = kmin;
  uint32_t s, P;
  // radix: digit window, guard prefixes (below / above the window), digit
  // prefixes of 2^s keys holding nbd exact sub-digits each
  uint32_t glo = 0, ndp = 0, nbd = 1;
  if (sd.mode == MODE_RADIX) {
    const uint32_t D = (uint32_t)bitlen<KT>(kmin ^ kmax);  // low bits the sampled keys differ in
    const uint32_t db = D < sd.radix_bits ? D : sd.radix_bits;
    const uint32_t dpb = db < kMaxPrefixBits - 1 ? db : kMaxPrefixBits - 1;
    s = D - dpb;
    nbd = 1u << (db - dpb);
    ndp = 1u << dpb;
    const KT base = D >= W ? KT(0) : (KT)((kmin >> D) << D);  // common prefix of the window
    const KT top = D >= W ? KeyBits<KT>::MAX : (KT)(base + ((KT(1) << D) - 1));
    glo = base > 0 ? 1u : 0u;
    kb = (KT)(base - ((KT)glo << s));
    P = glo + ndp + (top < KeyBits<KT>::MAX ? 1u : 0u);
    pb = (uint32_t)bitlen<uint32_t>(P - 1);
  } else if (sd.mode == MODE_RANGE) {
    pb = 0;
    s = (uint32_t)bl;
    P = 1;
  } else if (hgram && bl == (int)W) {
    // full-width range with k_sample's histogram: prefixes aligned to
    // multiples of 2^s, so the prefix of a key is (key >> s) - (kb >> s)
    // and its count is hgram[key >> s]
    s = W - pb;
    kb = (KT)((kmin >> s) << s);
    P = (uint32_t)((kmax >> s) - (kb >> s)) + 1;  // <= 2^pb
  } else {
    s = bl > (int)pb ? (uint32_t)(bl - (int)pb) : 0u;
    P = (uint32_t)(((uint64_t)R >> s) + 1);  // <= 2^pb
    // guard prefixes below kmin and above kmax's prefix when the LUT has
    // room (narrow key ranges): keys outside the sampled range then get
    // buckets of their own instead of widening the extreme sampled keys'
    // buckets to the ends of the key space (so those stay single-key when
    // s = 0).  Each guard is added only where keys can lie beyond the range.
    if (P + 2 <= (1u << pb)) {
      const uint32_t glo_s = kmin >= (KT(1) << s) ? 1u : 0u;
      kb = (KT)(kmin - ((KT)glo_s << s));
      const bool ghi_s = (KT)(KeyBits<KT>::MAX - kb) >> s >= (KT)(P + glo_s);
      P += glo_s + (ghi_s ? 1u : 0u);
    }
  }
  const uint32_t Bt = sd.bmax - (hvon ? 2u : 0u);  // bucket budget of L
  const uint32_t cap = s >= 16 ? 65535u : min(65535u, 1u << s);
  uint32_t Btot;

  if (sd.mode == MODE_RADIX) {
    for (uint32_t p = tid; p < P; p += kThreads) {
      if (p < glo) lut[p] = 0u | (1u << 16);
      else if (p < glo + ndp) lut[p] = (glo + (p - glo) * nbd) | (nbd << 16);
      else lut[p] = (glo + ndp * nbd) | (1u << 16);
    }
    Btot = glo + ndp * nbd + (P - glo - ndp);
  } else if (sd.mode == MODE_RANGE) {
    const uint32_t nb = min(Bt, cap);
    if (tid == 0) lut[0] = nb << 16;
    Btot = nb;
  } else {
    const bool useh = hgram && bl == (int)W;  // uniform
    if (useh && part != 0) return;  // uniform: one work-group reads the counts
    for (uint32_t p = tid; p < P; p += kThreads) reinterpret_cast<uint4*>(cnt4)[p] = make_uint4(0, 0, 0, 0);
    __syncthreads();
    const uint32_t copy = tid & 3u;
    // the four copies summed into cnt
    auto combine = [&]() {
      for (uint32_t p = tid; p < P; p += kThreads) {
        const uint4 c4 = reinterpret_cast<const uint4*>(cnt4)[p];
        cnt[p] = c4.x + c4.y + c4.z + c4.w;
      }
    };
    if (useh) {
      if (hload) reinterpret_cast<uint4*>(cnt4)[tid] = h4;  // the counts by key >> s
      __syncthreads();
      const uint32_t h0 = (uint32_t)(kb >> s);
      for (uint32_t p = tid; p < P; p += kThreads) cnt[p] = cnt4[h0 + p];
      __syncthreads();
    } else if (nparts == 1) {
      // the whole sample in one round of loads (64 per thread for 4-byte keys)
      strided_for<64 * 4 / sizeof(KT)>(samples + sd.sample_off, sd.nsample, [&](uint32_t, KT k) {
        atomicAdd(&cnt4[((uint32_t)((KT)(k - kb) >> s) << 2) | copy], 1u);
      });
      __syncthreads();
      combine();
      __syncthreads();
    } else {
      // this part's share of the sample; the partial counts are added into
      // the segment's LUT region (zeroed by k_sample), and the last part to
      // arrive goes on with their sums
      const uint32_t per = sd.nsample / nparts;  // the last part takes the remainder
      const uint32_t mine = part + 1 < nparts ? per : sd.nsample - part * per;
      strided_for<kLutPartBytes / sizeof(KT) / kThreads>(samples + sd.sample_off + part * per, mine,
                                              [&](uint32_t, KT k) {
        atomicAdd(&cnt4[((uint32_t)((KT)(k - kb) >> s) << 2) | copy], 1u);
      });
      __syncthreads();
      combine();
      __syncthreads();
      for (uint32_t p = tid; p < P; p += kThreads)
        if (cnt[p]) atomicAdd(&lut[p], cnt[p]);
      __threadfence();  // release: the sums before the arrival count
      __syncthreads();
      if (tid == 0) sh_last = atomicAdd(&states[blockIdx.x].arrive, 1u) == nparts - 1 ? 1u : 0u;
      __syncthreads();
      if (!sh_last) return;  // uniform
      __threadfence();  // acquire: every part's sums
      for (uint32_t p = tid; p < P; p += kThreads) cnt[p] = __hip_atomic_load(&lut[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
    }
    // prefixes holding no sample: a run of them shares one bucket
    uint32_t nonempty = 0, runs = 0, csum = 0;
    for (uint32_t p = tid; p < P; p += kThreads) {
      csum += cnt[p];
      if (cnt[p]) nonempty++;
      else if (p == 0 || cnt[p - 1]) runs++;
    }
    uint32_t t12, t3;  // nonempty prefixes and empty runs (< 2^16 each) in one scan
    block_excl_scan(nonempty | (runs << 16), wsum, t12);
    block_excl_scan(csum, wsum, t3);
    const uint32_t t1 = t12 & 0xFFFFu, t2 = t12 >> 16;
    const uint32_t A = Bt > t1 + t2 ? Bt - t1 - t2 : 0u;
    // the counts' own total (the sample size; the sub-bucket shares below
    // stay within the budget whatever the counts sum to)
    const uint32_t ns = t3 > 0 ? t3 : 1u;
    const uint32_t per = (P + kThreads - 1) / kThreads;
    const uint32_t p0 = tid * per, p1 = min(P, p0 + per);
    auto width = [&](uint32_t p, uint32_t& nb) -> uint32_t {
      const uint32_t c = cnt[p];
      if (c) { nb = min(cap, 1u + (uint32_t)(((uint64_t)c * A) / ns)); return nb; }
      nb = 0;
      return (p == 0 || cnt[p - 1]) ? 1u : 0u;
    };
    uint32_t local = 0, nbv;
    for (uint32_t p = p0; p < p1; ++p) local += width(p, nbv);
    uint32_t tot;
    uint32_t run = block_excl_scan(local, wsum, tot);
    for (uint32_t p = p0; p < p1; ++p) {
      const uint32_t w = width(p, nbv);
      const uint32_t base = (w == 0) ? run - 1 : run;  // a run's later prefixes
      lut[p] = base | (nbv << 16);
      run += w;
    }
    Btot = tot;
  }
  __syncthreads();  // the LUT is complete (thread 0 reads hv's entry)
  if (tid == 0) {
    SegState st = {};
    st.kb = (uint64_t)kb;
    st.kmax = (uint64_t)kmax;
    // lmax = min(P * 2^s - 1, KMAX - kb): every key maps into a prefix < P
    const unsigned __int128 span = ((unsigned __int128)P << s) - 1;
    const unsigned __int128 room = (unsigned __int128)(KeyBits<KT>::MAX - kb);
    st.lmax = (uint64_t)(span < room ? span : room);
    st.s = s;
    st.P = P;
    st.B = Btot;
    st.nact = 0;
    st.pwid = P > 1 ? pb : 0;
    st.poff = P > 1 ? s : 0;
    st.twid = s < 24 ? s : 24;
    st.rsh = s > 24 ? s - 24 : 0;
    st.lsh = s < 24 ? 24 - s : 0;
    st.hvon = 0;
    st.hv = 0;
    st.hb = 0;
    if (hvon) {  // hv's bucket is L(hv) + 1; two more buckets in all
      BucketFn<KT> fn;
      fn.init(st);
      st.hvon = 1;
      st.hv = (uint64_t)hv;
      st.hb = fn(hv, lut) + 1;
      st.B = Btot + 2;
    }
    states[blockIdx.x] = st;
  }
  // k_sample counts into a zeroed histogram: clear it for the next select
  // (every read of it is behind the barriers above)
  if (hgram && sd.mode == MODE_SAMPLE)
    for (uint32_t i = tid; i < (1u << sd.pbits); i += kThreads) hgram[i] = 0;
}

// Pass 1: per-chunk LDS histogram of bucket ids.
constexpr uint32_t kLaneSlots = 64;  // k_count's per-lane heavy-bin slots
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_count(const KT* __restrict__ src,
                                                    const SegDesc* __restrict__ segs,
                                                    const SegState* __restrict__ states,
                                                    const Chunk* __restrict__ chunks,
                                                    const uint32_t* __restrict__ lut_pool,
                                                    uint32_t* __restrict__ rows) {
  extern __shared__ uint32_t lds[];
  const Chunk ch = chunks[blockIdx.x];
  const SegDesc& sd = segs[ch.seg];
  const SegState st = states[ch.seg];
  const uint32_t bmax = sd.bmax;
  // LDS: the histogram (bmax bins), one slot per lane for the heavy value's
  // bin (kLaneSlots), the LUT
  uint32_t* hist = lds;
  const uint32_t hslot = bmax + lane_id();  // this lane's heavy-bin slot
  uint32_t* lut = lds + bmax + kLaneSlots;
  constexpr uint32_t VEC = 16 / sizeof(KT);
  constexpr int R = 4, N = R * VEC;
  const KT* p = src + sd.data_off + ch.start;
  const uint64_t len = ch.len;
  const uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
  uint64_t head = mis ? (VEC - mis) : 0;
  if (head > len) head = len;
  const uint64_t nvec = (len - head) / VEC;
  const uint64_t done = head + nvec * VEC;
  const uint32_t tid = threadIdx.x;
  const uint4* __restrict__ pv = reinterpret_cast<const uint4*>(p + head);
  const uint64_t per_round = (uint64_t)R * kThreads;
  uint64_t r0 = 0;
  // the first round's loads go out before the histogram is cleared and the
  // LUT loaded, so they arrive during that prologue
  uint4 v[R];
  if (per_round <= nvec) {  // uniform
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = stream_load(&pv[(uint64_t)t * kThreads + tid]);
  }
  for (uint32_t i = threadIdx.x; i < st.B; i += kThreads) hist[i] = 0;
  if (threadIdx.x < kLaneSlots) hist[bmax + threadIdx.x] = 0;
  for (uint32_t i = threadIdx.x; i < st.P; i += kThreads) lut[i] = lut_pool[sd.lut_off + i];
  __syncthreads();
  BucketFn<KT> bk;
  bk.init(st);
  if (tid < head) atomicAdd(&hist[bk(to_key<C>(p[tid]), lut)], 1u);
  if (done + tid < len) atomicAdd(&hist[bk(to_key<C>(p[done + tid]), lut)], 1u);
  // full rounds, software pipelined: the next round's loads are in flight
  // while a round is bucketed and counted
  // With a heavy value its bucket takes up to most keys: each lane counts
  // those in a slot of its own (a select of the address, no lanes of one
  // atomic on one address, no change of the execution mask), and the slots
  // are added into the bucket at the end.
  const uint32_t hbk = st.hb;
  auto full_rounds = [&](auto left, auto hvt) {
    if (r0 + per_round > nvec) return;  // uniform (round 0's loads are in v)
    for (; r0 + per_round <= nvec; r0 += per_round) {
      KT k[N];
      unpack<KT, R>(v, k);
      if (r0 + 2 * per_round <= nvec) {  // uniform
#pragma unroll
        for (int t = 0; t < R; ++t)
          v[t] = stream_load(&pv[r0 + per_round + (uint64_t)t * kThreads + tid]);
      }
      uint32_t b[N];
      bucket_batch1<KT, C, N, decltype(left)::value, decltype(hvt)::value>(bk, lut, k, b);
      if (decltype(hvt)::value) {
#pragma unroll
        for (int i = 0; i < N; ++i) atomicAdd(&hist[b[i] == hbk ? hslot : b[i]], 1u);
      } else {
#pragma unroll
        for (int i = 0; i < N; ++i) atomicAdd(&hist[b[i]], 1u);
      }
    }
  };
  // uniform: the sub-field shift direction and the heavy-value term
  if (!st.hvon) {
    if (st.s <= 24) full_rounds(std::true_type{}, std::false_type{});
    else full_rounds(std::false_type{}, std::false_type{});
  } else {
    if (st.s <= 24) full_rounds(std::true_type{}, std::true_type{});
    else full_rounds(std::false_type{}, std::true_type{});
  }
  for (uint64_t i = r0 + tid; i < nvec; i += kThreads) {  // last partial round
    uint4 v[1] = {stream_load(&pv[i])};
    KT k[VEC];
    unpack<KT, 1>(v, k);
    uint32_t b[VEC];
    bucket_batch<KT, C, VEC>(bk, lut, k, b);
#pragma unroll
    for (uint32_t j = 0; j < VEC; ++j) atomicAdd(&hist[b[j]], 1u);
  }
  if (st.hvon) {  // uniform
    __syncthreads();
    if (threadIdx.x < kLaneSlots) {
      const uint32_t v = hist[bmax + threadIdx.x];
      const uint32_t t = wave_last(wave_incl_scan(v));
      if (threadIdx.x == 0) hist[hbk] += t;
    }
  }
  __syncthreads();
  uint32_t* row = rows + sd.row_base + (uint64_t)ch.idx * bmax;
  for (uint32_t i = threadIdx.x; i < st.B; i += kThreads) row[i] = hist[i];
}

// Sum the chunk rows of an item's buckets; rewrite each row entry as the
// chunk's exclusive offset inside its bucket; write the bucket prefixes
// inside each group of kReduceBins and the group sums.  The 16 wavefronts
// of a work-group are (chunk parts) x (2^gi_log2 groups of 64 bins): a segment with
// many chunks splits them over 16 parts, one with few chunks covers up to 16
// groups per work-group instead.  Each lane keeps its (at most kReduceRegs)
// row entries in registers between the sum and the rewrite.
constexpr uint32_t kReduceWaves = kThreads / kReduceBins;  // 16
constexpr uint32_t kReduceRegs = kMaxChunks / kReduceWaves;  // chunks per part

__global__ __launch_bounds__(kThreads) void k_reduce(const SegDesc* __restrict__ segs,
                                                     const SegState* __restrict__ states,
                                                     const ReduceItem* __restrict__ items,
                                                     uint32_t* __restrict__ rows,
                                                     uint32_t* __restrict__ totals,
                                                     uint32_t* __restrict__ gsums) {
  __shared__ uint32_t part[kReduceWaves][kReduceBins];
  const ReduceItem it = items[blockIdx.x];
  const SegDesc& sd = segs[it.seg];
  const uint32_t B = states[it.seg].B;
  const uint32_t gl = it.gi_log2, pl = 4 - gl;  // log2 of groups and of chunk parts
  const uint32_t parts = 1u << pl;
  const uint32_t lane = threadIdx.x % kReduceBins, w = threadIdx.x / kReduceBins;
  const uint32_t g = w & ((1u << gl) - 1u), q = w >> gl;  // this wavefront's group and part
  const uint32_t group = it.group + g;
  const uint32_t bin = group * kReduceBins + lane;
  const bool valid = bin < B;
  const uint32_t nch = sd.nchunks;  // <= kMaxChunks (planner)
  const uint32_t c0 = (nch * q) >> pl;
  const uint32_t c1 = (nch * (q + 1)) >> pl;
  uint32_t* base = rows + sd.row_base + bin;
  const uint64_t stride = sd.bmax;
  uint32_t v[kReduceRegs];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t u = 0; u < kReduceRegs; ++u) {
    v[u] = (valid && c0 + u < c1) ? base[(uint64_t)(c0 + u) * stride] : 0u;
    sum += v[u];
  }
  part[w][lane] = sum;
  __syncthreads();
  uint32_t off = 0, total = 0;
#pragma unroll
  for (uint32_t j = 0; j < kReduceWaves; ++j) {
    if (j < parts) {
      const uint32_t x = part[(j << gl) + g][lane];
      off += j < q ? x : 0u;
      total += x;
    }
  }
  uint32_t run = off;
#pragma unroll
  for (uint32_t u = 0; u < kReduceRegs; ++u) {
    if (valid && c0 + u < c1) base[(uint64_t)(c0 + u) * stride] = run;
    run += v[u];
  }
  // the part-0 wavefront of each group holds its 64 bins: their exclusive
  // prefix inside the group, and the group's sum (k_findk scans the group
  // sums only); groups past the segment's bucket budget are not written
  if (q == 0 && group * kReduceBins < sd.bmax) {
    const uint32_t t = valid ? total : 0u;
    const uint32_t inc = wave_incl_scan(t);
    if (valid) totals[sd.tot_off + bin] = inc - t;
    if (lane == kReduceBins - 1) gsums[sd.tot_off / kReduceBins + group] = inc;
  }
}

// The i-th query of a segment: explicit, from a raw k list (k-th largest ->
// rank len - k + 1, output slot i; a bad k raises a status bit and is clamped
// to k = 1), or rank len - i for a top-K.
__device__ __forceinline__ Query seg_query(const SegDesc& sd, const Query* __restrict__ queries,
                                           const uint32_t* __restrict__ kraw, uint32_t i,
                                           uint32_t* __restrict__ status,
                                           const uint32_t* kpre = nullptr) {
  Query q;
  if (sd.qsrc == QSRC_QUERIES) return queries[sd.q_off + i];
  if (sd.qsrc == QSRC_K) {
    uint32_t k = kpre ? *kpre : kraw[sd.q_off + i];
    if (k < 1 || (uint64_t)k > sd.len) {
      atomicOr(status, (uint32_t)STATUS_BAD_K);
      k = 1;
    }
    q.rank = (uint32_t)(sd.len - k + 1);
  } else {
    q.rank = (uint32_t)(sd.len - i);
  }
  q.out = sd.out_off + i;
  return q;
}

// Output slot of the i-th query of a segment (no k list read).
__device__ __forceinline__ uint32_t seg_out(const SegDesc& sd, const Query* __restrict__ queries,
                                            uint32_t i) {
  return sd.qsrc == QSRC_QUERIES ? queries[sd.q_off + i].out : sd.out_off + i;
}

// Per segment: ranks -> buckets, children, the child map for the compaction
// pass, and the ranks grouped by child.  TV is the number of 16-byte words
// of bucket prefixes a thread holds: the full instance takes up to
// kMaxBuckets buckets, the small one up to 4096 in at most 64 VGPRs, so that
// two work-groups share a CU (batches of many segments).
template <typename KT, int C, uint32_t TV>
__device__ __forceinline__ void findk_segment(const SegDesc* __restrict__ segs,
                                              SegState* __restrict__ states,
                                              const uint32_t* __restrict__ lut_pool,
                                              const uint32_t* __restrict__ totals,
                                              const uint32_t* __restrict__ gsums,
                                              const Query* __restrict__ queries,
                                              const uint32_t* __restrict__ kraw,
                                              uint32_t* __restrict__ status,
                                              uint32_t* __restrict__ qb,
                                              uint32_t* __restrict__ qr,
                                              Child* __restrict__ children,
                                              Query* __restrict__ cq,
                                              uint32_t* __restrict__ bits_pool,
                                              uint32_t* __restrict__ wbase_pool,
                                              KT* __restrict__ out,
                                              uint32_t* __restrict__ l2count,
                                              Heavy* __restrict__ heavy,
                                              Chunk2* __restrict__ chunks2, uint32_t hmax,
                                              uint32_t* __restrict__ bigw,
                                              uint2* __restrict__ cinfo) {
  extern __shared__ __align__(16) uint32_t lds[];
  const SegDesc sd = segs[blockIdx.x];
  // the first k of this thread, read now so its latency hides under the
  // table loads and scans
  const uint32_t kfirst =
      (sd.qsrc == QSRC_K && threadIdx.x < sd.q_cnt) ? kraw[sd.q_off + threadIdx.x] : 0u;
  const uint32_t wcap = (sd.bmax + 31) / 32;
  uint32_t* hist = lds;                          // bmax
  uint32_t* lut = hist + sd.bmax;                // 2^pbits
  uint32_t* bits = lut + (1u << sd.pbits);       // wcap
  uint32_t* wb = bits + wcap;                    // wcap: first child of a word
  uint32_t* wc = wb + wcap;                      // wcap: first reduced slot of a word
  uint32_t* wsum = wc + wcap;                    // 17 (+ pad)
  const uint32_t tid = threadIdx.x;
  // bucket prefixes inside groups of kReduceBins, the group sums and the LUT:
  // they need only the descriptor, so their loads (16 bytes each, all of a
  // thread's at once) go out before the state.  The table sizes are
  // multiples of 4 words at 16-byte aligned offsets.  The bucket prefixes
  // stay in registers until the group bases are scanned and are stored once,
  // base added.  Entries past the segment's bucket count are never used.
  const uint32_t nt4 = sd.bmax / 4, ng4 = sd.bmax / kReduceBins / 4, nl4 = (1u << sd.pbits) / 4;
  const uint4* tot4 = reinterpret_cast<const uint4*>(totals + sd.tot_off);
  uint4 tv[TV];
#pragma unroll
  for (uint32_t u = 0; u < TV; ++u)
    if (tid + u * kThreads < nt4) tv[u] = tot4[tid + u * kThreads];
  uint4 gv = make_uint4(0, 0, 0, 0), lv = make_uint4(0, 0, 0, 0);
  if (tid < ng4) gv = reinterpret_cast<const uint4*>(gsums + sd.tot_off / kReduceBins)[tid];
  if (tid < nl4) lv = reinterpret_cast<const uint4*>(lut_pool + sd.lut_off)[tid];
  __shared__ SegState st;
  if (threadIdx.x == 0) st = states[blockIdx.x];
  if (tid < ng4) reinterpret_cast<uint4*>(wc)[tid] = gv;
  if (tid < nl4) reinterpret_cast<uint4*>(lut)[tid] = lv;
  __syncthreads();
  const uint32_t B = st.B;
  const uint32_t nwords = (B + 31) / 32;
  const uint32_t ngroups = (B + kReduceBins - 1) / kReduceBins;  // <= wcap
  for (uint32_t i = tid; i < nwords; i += kThreads) bits[i] = 0;
  const uint32_t len = block_scan_array(wc, ngroups, wsum);  // group bases (synchronises)
#pragma unroll
  for (uint32_t u = 0; u < TV; ++u) {
    const uint32_t i4 = tid + u * kThreads;
    if (i4 < nt4) {
      const uint32_t g = wc[i4 * 4 / kReduceBins];
      reinterpret_cast<uint4*>(hist)[i4] =
          make_uint4(tv[u].x + g, tv[u].y + g, tv[u].z + g, tv[u].w + g);
    }
  }
  __syncthreads();  // hist = exclusive prefix of the bucket counts
  RangeOf<KT> range;
  range.fn.init(st);
  range.lut = lut;
  range.B = B;
  range.P = st.P;
  range.s = st.s;
  range.hb = st.hb;
  auto count_of = [&](uint32_t b) { return (b + 1 < B ? hist[b + 1] : len) - hist[b]; };

  // A: bucket of every rank; single-key buckets are answered here
  const uint32_t NONE = 0xFFFFFFFFu;
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    const Query q = seg_query(sd, queries, kraw, i, status, i == tid ? &kfirst : nullptr);
    const uint32_t r1 = q.rank - 1;
    uint32_t lo = 0, hi = B - 1;  // largest b with hist[b] <= r1
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (hist[mid] <= r1) lo = mid; else hi = mid - 1;
    }
    // a bucket inside a prefix of 2^s keys split nb < 2^16 ways spans at
    // least floor(2^s / nb) >= 2 keys when s >= 17 (edge buckets only
    // widen), so its exact range is needed only for smaller prefixes, and
    // for the heavy value's bucket (one key by construction)
    KT klo = 0, khi = 1;
    if (st.s < 17 || (st.hvon && lo == st.hb)) range(lo, klo, khi);
    if (klo == khi) {
      out[q.out] = from_key<C>(klo);
      qb[sd.q_off + i] = NONE;
    } else {
      lds_or(bits, lo >> 5, 1u << (lo & 31));
      qb[sd.q_off + i] = lo;
      qr[sd.q_off + i] = q.rank - hist[lo];
    }
  }
  __syncthreads();

  // B: per bitmap word its first child and first reduced slot; the child map
  uint32_t pc = 0, csum = 0, m = 0;
  if (tid < nwords) {
    m = bits[tid];
    pc = __popc(m);
    for (uint32_t mm = m; mm; mm &= mm - 1) csum += count_of(tid * 32 + __ffs(mm) - 1);
  }
  uint32_t U, mtot;
  const uint32_t ubase = block_excl_scan(pc, wsum, U);
  const uint32_t cbase = block_excl_scan(csum, wsum, mtot);
  if (tid < nwords) {
    wb[tid] = ubase;
    wc[tid] = cbase;
  }
  if (tid < nwords) {
    bits_pool[sd.bits_off + tid] = bits[tid];
    wbase_pool[sd.bits_off + tid] = ubase;
  }
  __syncthreads();
  const bool l2 = (sd.flags & SEG_LEVEL2) && l2count;
  // one thread per child: its bucket and reduced slice (its key range once
  // its ranks are counted, below)
  for (uint32_t u = tid; u < U; u += kThreads) {
    uint32_t a = 0, z = nwords - 1;  // largest word with wb <= u
    while (a < z) {
      uint32_t mid = (a + z + 1) >> 1;
      if (wb[mid] <= u) a = mid; else z = mid - 1;
    }
    uint32_t mw = bits[a], start = wc[a];
    for (uint32_t r = u - wb[a]; r; --r) {  // skip earlier children of the word
      start += count_of(a * 32 + __ffs(mw) - 1);
      mw &= mw - 1;
    }
    const uint32_t b = a * 32 + __ffs(mw) - 1;
    Child c;
    c.start = sd.red_base + start;
    c.len = count_of(b);
    c.lo = 0;
    c.hi = 0;
    c.need_range = 1;
    c.q_off = 0;
    c.q_cnt = 0;
    c.src = 1;
    c.bucket = b;
    c.seg = blockIdx.x;
    children[sd.q_off + u] = c;
    cinfo[sd.q_off + u] = make_uint2(start, b);
  }
  __syncthreads();

  // C: group ranks by child (hist is free now: counters, then offsets)
  __shared__ uint32_t lsplit;  // child slots taken past U by split children
  __shared__ uint32_t lbig;    // children for k_finish
  if (tid == 0) {
    lsplit = 0;
    lbig = 0;
  }
  for (uint32_t u = tid; u < U; u += kThreads) hist[u] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    uint32_t b = qb[sd.q_off + i];
    if (b == NONE) continue;
    uint32_t w = b >> 5, mw = bits[w];
    uint32_t u = wb[w] + __popc(mw & ((1u << (b & 31)) - 1u));
    lds_inc(hist, u);
    qb[sd.q_off + i] = u;
  }
  __syncthreads();
  const uint32_t nq = block_scan_array(hist, U, wsum);  // hist[u] = first slot
  for (uint32_t u = tid; u < U; u += kThreads) {
    uint32_t first = hist[u];
    uint32_t next = (u + 1 < U) ? hist[u + 1] : nq;
    Child* cp = &children[sd.q_off + u];  // this thread wrote the record above
    cp->q_off = sd.q_off + first;
    cp->q_cnt = next - first;
    const uint64_t clen = cp->len;
    // the key range where a later step needs it before reading the keys (a
    // child the second level may take, or one k_finish walks from memory);
    // a child of few ranks held in the 256-thread finisher's registers gets
    // the exact min and max of its keys there, one pass over registers,
    // which is cheaper than inverting the bucket function here
    if ((l2 && clen > kHeavyKeys) || clen > finish_small_reg_keys<KT>(false) ||
        next - first > kWaveRanks) {
      KT klo, khi;
      range(cp->bucket, klo, khi);
      cp->lo = klo;
      cp->hi = khi;
      cp->need_range = 0;
    }
    // a child of 2..kWaveRanks ranks for the 256-thread finisher (and no
    // second level, which takes the free slots) is split into one-rank
    // records, the extra ones in free child slots: its ranks are then
    // finished side by side by as many work-groups, not one after another
    if (!l2 && next - first > 1 && next - first <= kWaveRanks && clen <= kWaveMaxLen) {
      const uint32_t extra = next - first - 1;
      const uint32_t slot = U + atomicAdd(&lsplit, extra);
      // the record copied as 8-byte words held in registers (a Child
      // object here lived in scratch memory), q_off and q_cnt in word 4
      static_assert(sizeof(Child) == 56 && offsetof(Child, q_off) == 32 &&
                        offsetof(Child, q_cnt) == 36, "Child layout");
      const uint2* cw = reinterpret_cast<const uint2*>(cp);
      uint2 rec[7];
#pragma unroll
      for (int w = 0; w < 7; ++w) rec[w] = cw[w];
      rec[4].y = 1;
#pragma unroll
      for (uint32_t j = 0; j < kWaveRanks - 1; ++j) {
        if (j < extra) {
          rec[4].x = sd.q_off + first + 1 + j;
          uint2* dw = reinterpret_cast<uint2*>(&children[sd.q_off + slot + j]);
#pragma unroll
          for (int w = 0; w < 7; ++w) dw[w] = rec[w];
        }
      }
      cp->q_cnt = 1;
    }
    if (l2 && clen > kHeavyKeys) {
      // a heavy child goes to the second level: a heavy-child record and its
      // chunks; its slot finishes nothing until k_count2 writes a
      // second-level child there (the capacity is the planner's exact
      // bound, so a refused registration leaves an ordinary child)
      const uint32_t h = atomicAdd(&l2count[0], 1u);
      if (h < hmax) {
        const uint32_t nch = (uint32_t)((clen + kChunk2 - 1) / kChunk2);
        const uint32_t cb = atomicAdd(&l2count[1], nch);
        Heavy H;
        H.start = cp->start;
        H.len = clen;
        H.lo = cp->lo;
        H.hi = cp->hi;
        H.seg = blockIdx.x;
        H.slot = sd.q_off + u;
        H.q_off = sd.q_off + first;
        H.q_cnt = next - first;
        H.chunk_base = cb;
        H.nchunks = nch;
        const int rb = bitlen<KT>((KT)(cp->hi - cp->lo));
        H.shift = rb > (int)kBits2 ? (uint32_t)(rb - (int)kBits2) : 0u;
        H.done = 0;
        H.nact = 0;
        H.slot_base = 0;
        H.pad[0] = H.pad[1] = 0;
        heavy[h] = H;
        for (uint32_t i = 0; i < nch; ++i) {
          Chunk2 c2;
          c2.h = h;
          c2.idx = i;
          c2.start = (uint64_t)i * kChunk2;
          c2.len = min(kChunk2, clen - c2.start);
          chunks2[cb + i] = c2;
        }
        cp->q_cnt = 0;
      }
    }
    if (big_child(cp->len, cp->q_cnt)) atomicAdd(&lbig, 1u);
  }
  if (l2 && tid == 0) states[blockIdx.x].nfree = U;  // first free child slot
  __syncthreads();
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    uint32_t u = qb[sd.q_off + i];
    if (u == NONE) continue;
    uint32_t pos = lds_inc(hist, u);
    Query q;
    q.rank = qr[sd.q_off + i];
    q.out = seg_out(sd, queries, i);
    cq[sd.q_off + pos] = q;
  }
  // unused child slots finish nothing
  __syncthreads();  // lsplit is final
  for (uint32_t u = U + lsplit + tid; u < sd.q_cnt; u += kThreads) {
    children[sd.q_off + u].q_cnt = 0;
    children[sd.q_off + u].len = 0;
  }
  if (tid == 0) {
    states[blockIdx.x].nact = U;
    states[blockIdx.x].nred = mtot;
    if (bigw && lbig) atomicAdd(bigw, lbig);
  }
}

#define GGKS_FINDK_PARAMS                                                                      \
  const SegDesc *__restrict__ segs, SegState *__restrict__ states,                             \
      const uint32_t *__restrict__ lut_pool, const uint32_t *__restrict__ totals,              \
      const uint32_t *__restrict__ gsums, const Query *__restrict__ queries,                   \
      const uint32_t *__restrict__ kraw, uint32_t *__restrict__ status, uint32_t *__restrict__ qb, \
      uint32_t *__restrict__ qr, Child *__restrict__ children, Query *__restrict__ cq,         \
      uint32_t *__restrict__ bits_pool, uint32_t *__restrict__ wbase_pool, KT *__restrict__ out, \
      uint32_t *__restrict__ l2count, Heavy *__restrict__ heavy, Chunk2 *__restrict__ chunks2,  \
      uint32_t hmax, uint32_t *__restrict__ bigw, uint2 *__restrict__ cinfo
#define GGKS_FINDK_ARGS                                                                        \
  segs, states, lut_pool, totals, gsums, queries, kraw, status, qb, qr, children, cq, bits_pool, \
      wbase_pool, out, l2count, heavy, chunks2, hmax, bigw, cinfo
constexpr uint32_t kFindkSmallBuckets = 4 * kThreads;  // one 16-byte word per thread
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_findk(GGKS_FINDK_PARAMS) {
  findk_segment<KT, C, kMaxBuckets / 4 / kThreads>(GGKS_FINDK_ARGS);
}
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_findk_small(
    GGKS_FINDK_PARAMS) {
  findk_segment<KT, C, 1>(GGKS_FINDK_ARGS);
}
#undef GGKS_FINDK_PARAMS
#undef GGKS_FINDK_ARGS

// Pass 2 (k_compact): stream the chunk once and move the keys of active
// buckets into their children's ranges (each chunk owns a private slice of
// every child: its running offsets start at the chunk's row entries).  Every
// key is bucketed exactly as in pass 1 (LUT in static LDS) and tested in a
// byte-per-bucket copy of the active-bucket bitmap (one LDS byte load, no bit
// arithmetic); the active keys of a round are appended straight from
// registers to a work-group buffer in LDS (one LDS atomic per wavefront).
// The main loop issues no global stores, so its waits cover only its own
// loads.  After each round the work-group synchronises; once the buffer
// holds more than a few thousand keys (or could not take another round) it
// is flushed: every buffered key is placed (child, LDS cursor, store) in one
// short burst between rounds.
// Table sizes of k_compact.  The full instance serves any plan.  The small
// one serves plans of at most 4096 buckets and 2^10 LUT prefixes per segment
// in under 62 KB of LDS, so two work-groups share a CU and one chunk's table
// loads and last flush hide behind another chunk's streaming (batches of
// many short segments, whose chunks are 2^14..2^17 keys).
//
// The dynamic LDS holds the children's cursors and the key buffer.  The
// first `cursors` children of a segment keep their cursor in LDS; children
// beyond place each key with a global atomic on the chunk's row.  With
// 4-byte keys the full instance's buffer still takes a round above the
// flush threshold with 4096 cursors, so it has room for them (selects of a
// few thousand ranks keep every cursor in LDS); 8-byte keys need the room
// for the buffer.
template <bool SMALL>
struct CompactCfg {
  static constexpr uint32_t lut_words = SMALL ? (1u << 10) : (1u << kMaxPrefixBits);
  static constexpr uint32_t buckets = SMALL ? 4096u : kMaxBuckets;
  static constexpr int round = SMALL ? 2 : 4;  // 16-byte loads per lane per round
  static constexpr uint32_t dyn_bytes = SMALL ? 52u * 1024 : 100u * 1024;  // cursors + key buffer
  template <typename KT>
  static constexpr uint32_t cursors = (!SMALL && sizeof(KT) == 4) ? 4096u : 1024u;
  template <typename KT>
  static constexpr uint32_t buf_bytes = dyn_bytes - cursors<KT> * 4;  // buffered active keys
};
inline bool compact_small(uint32_t pbits_max, uint32_t bmax_max) {
  return (1u << pbits_max) <= CompactCfg<true>::lut_words && bmax_max <= CompactCfg<true>::buckets;
}

template <bool SMALL>
__host__ __device__ constexpr size_t compact_lds_bytes() {
  // dynamic part: cursors and the key buffer (the LUT and bitmaps are static)
  return CompactCfg<SMALL>::dyn_bytes;
}

template <typename KT, int C, bool SMALL, bool WIN>
__device__ __forceinline__ void compact_chunk(const KT* __restrict__ src,
                                              const SegDesc* __restrict__ segs,
                                              const SegState* __restrict__ states,
                                              const Chunk* __restrict__ chunks,
                                              const uint32_t* __restrict__ lut_pool,
                                              uint32_t* __restrict__ rows,
                                              const uint32_t* __restrict__ bits_pool,
                                              const uint32_t* __restrict__ wbase_pool,
                                              const uint2* __restrict__ cinfo,
                                              KT* __restrict__ red) {
  using Cfg = CompactCfg<SMALL>;
  constexpr int kRound = Cfg::round;
  extern __shared__ __align__(16) uint32_t lds[];
  // static LDS at fixed addresses: no base add in the per-key address math
  __shared__ uint32_t lut[Cfg::lut_words];
  __shared__ uint32_t abits[Cfg::buckets / 32];  // active buckets
  __shared__ uint32_t wbase[Cfg::buckets / 32];  // first child of each bitmap word
  __shared__ uint32_t abyte[Cfg::buckets / 4];   // active buckets, a byte (0 or 1) each
  __shared__ uint32_t nbuf;                      // keys in the buffer
  __shared__ uint32_t dsum[kThreads / 64];       // dense rounds: wavefront totals
  const Chunk ch = chunks[blockIdx.x];
  const SegDesc& sd = segs[ch.seg];
  const SegState st = states[ch.seg];
  const uint32_t U = st.nact;
  if (U == 0) return;
  constexpr uint32_t VEC = 16 / sizeof(KT);
  constexpr int N = kRound * VEC;  // keys per lane per round
  constexpr uint32_t CURS = Cfg::template cursors<KT>;
  constexpr uint32_t CAP = Cfg::template buf_bytes<KT> / sizeof(KT);
  constexpr uint32_t ROUND_MAX = kThreads * N;  // most keys one round can append
  // a round starts with at most CAP - ROUND_MAX keys buffered, so it fits
  static_assert(CAP > ROUND_MAX, "buffer must hold more than a round");
  // The buffer is flushed once it holds more than FLUSH_AT keys.  Short
  // bursts keep the load stream going between them: at 2^28 f32 keys,
  // flushing above 4096 keys instead of 8192 took the pass from 233 to
  // 225 us, and above 2048 or 3072 (f64: 2048, 3072 instead of 4096) was
  // slower again.
  constexpr uint32_t FLUSH_AT = CAP - ROUND_MAX < 4096u ? CAP - ROUND_MAX : 4096u;
  const uint32_t nwords = (st.B + 31) / 32;
  uint32_t* cur = lds;
  KT* kb = reinterpret_cast<KT*>(cur + CURS);
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const KT* p = src + sd.data_off + ch.start;
  const uint64_t len = ch.len;
  uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
  uint64_t head = mis ? (VEC - mis) : 0;
  if (head > len) head = len;
  const uint64_t nvec = (len - head) / VEC;
  const uint64_t done = head + nvec * VEC;
  const uint4* pv = reinterpret_cast<const uint4*>(p + head);
  const uint64_t per_round = (uint64_t)kRound * kThreads;
  const uint64_t rounds = (nvec + per_round - 1) / per_round;  // uniform
  const uint64_t full_rounds = nvec / per_round;  // rounds with every load in range
  // software pipeline: the loads of round r + 1 are in flight while round r
  // is filtered and appended
  uint4 v[kRound];
  uint32_t okmask = 0;
  auto load_round = [&](uint64_t r) {
    const uint4* q = pv + r * per_round + tid;
    if (r < full_rounds) {  // uniform
      okmask = (N >= 32) ? 0xFFFFFFFFu : ((1u << N) - 1u);
#pragma unroll
      for (int t = 0; t < kRound; ++t) v[t] = stream_load(q + (uint64_t)t * kThreads);
      return;
    }
    okmask = 0;
#pragma unroll
    for (int t = 0; t < kRound; ++t) {
      const uint64_t idx = r * per_round + (uint64_t)t * kThreads + tid;
      const bool ok = idx < nvec;
      okmask |= ok ? (((1u << VEC) - 1u) << (t * VEC)) : 0u;
      v[t] = ok ? stream_load(q + (uint64_t)t * kThreads) : make_uint4(0, 0, 0, 0);
    }
  };
  const uint2* kinfo = cinfo + sd.q_off;  // per child: slice start in the segment's part, bucket
  uint32_t* row = rows + sd.row_base + (uint64_t)ch.idx * sd.bmax;
  // Cursors are positions inside the segment's part of the reduced buffer
  // (child start - red_base < segment length < 2^32), so they stay 32-bit
  // however far into the buffer the segment sits (batches and served groups
  // of more than 2^32 keys).
  KT* const rseg = red + sd.red_base;
  // Cursor windows.  A segment of more children than LDS cursors either
  // places the keys of the children beyond through global atomics on the
  // chunk's row, or streams the chunk once per window of CURS children,
  // each pass placing only its window's keys.  A key placed through a
  // global atomic costs about what streaming 200 bytes does (at 2^28 f32
  // keys, 16384 k: ~94 M such keys took ~2.9 ms), so the windows are taken
  // when those keys times 200 bytes exceed the extra passes' bytes.
  constexpr uint64_t kWindowKeyBytes = 200;
  uint32_t nwin = 1;
  if (WIN && U > CURS) {  // uniform
    const uint64_t beyond = st.nred - kinfo[CURS].x;  // keys of children CURS..U-1
    const uint32_t w = (U + CURS - 1) / CURS;
    if (beyond * kWindowKeyBytes > (uint64_t)(w - 1) * sd.len * sizeof(KT)) nwin = w;
  }
  uint32_t wlo = 0;                       // the window's first child
  uint32_t ucur = U < CURS ? U : CURS;    // its children with an LDS cursor
  const uint8_t* const act_of = reinterpret_cast<const uint8_t*>(abyte);
  // the window's tables: the byte table of its active buckets, its cursors
  auto window_tables = [&]() {
    for (uint32_t i = tid; i < nwords; i += kThreads) {
      uint32_t w = abits[i];
      if (nwin > 1) {  // uniform: only the bits of children in [wlo, wlo + ucur)
        uint32_t c = wbase[i], m = 0;
        for (uint32_t mm = w; mm; mm &= mm - 1, ++c)
          if (c >= wlo && c < wlo + ucur) m |= mm & (0u - mm);
        w = m;
      }
      // bits 4j..4j+3 -> bytes of abyte[8i + j]: x * 0x204081 puts bit m of a
      // nibble x at bit 8m, and no two shifted copies overlap
#pragma unroll
      for (uint32_t j = 0; j < 8; ++j) abyte[i * 8 + j] = (((w >> (4 * j)) & 0xFu) * 0x204081u) & 0x01010101u;
    }
    for (uint32_t u = tid; u < ucur; u += kThreads)
      cur[u] = kinfo[wlo + u].x + row[kinfo[wlo + u].y];
    if (tid == 0) nbuf = 0;
  };
  // round 0's loads go out before the tables are loaded and the cursors
  // set, so they arrive during that prologue
  if (rounds) load_round(0);
  strided_for(lut_pool + sd.lut_off, st.P, [&](uint32_t i, uint32_t x) { lut[i] = x; });
  for (uint32_t i = tid; i < nwords; i += kThreads) {
    abits[i] = bits_pool[sd.bits_off + i];
    wbase[i] = wbase_pool[sd.bits_off + i];
  }
  window_tables();  // (each thread reads back only the words it wrote)
  __syncthreads();
  BucketFn<KT> bk;
  bk.init(st);
  // moves a key of an active bucket to its child's slice
  auto place = [&](KT raw) {
    const uint32_t b = bk(to_key<C>(raw), lut);
    if (act_of[b]) {
      const uint32_t bits = abits[b >> 5];
      const uint32_t u = wbase[b >> 5] + __popc(bits & ((1u << (b & 31u)) - 1u));
      const uint32_t pos = u - wlo < CURS
                               ? atomicAdd(&cur[u - wlo], 1u)
                               : kinfo[u].x + atomicAdd(&row[kinfo[u].y], 1u);
      rseg[pos] = raw;
    }
  };
  // every thread, with the same n: place the buffered keys, empty the buffer.
  // Buffered keys are active by construction; F of them per thread go
  // through each step together so their LDS round trips overlap.
  // Sorted placement (full instance, more than kDenseChildren children, at
  // most SORT_MAX keys): each key takes
  // its slot from its child's cursor as below, and the keys are then
  // written to LDS in child order and stored from there, so consecutive
  // lanes store consecutive slots of a child and a wavefront's stores cover
  // a few lines instead of one line per key.  The cursors at the flush's
  // start (read by the threads owning them) and the children's counts in
  // this flush give each child's first sorted slot; while the sorted keys
  // are stored, a child's cursor word holds (its start - its first slot).
  // The sorted keys and their children sit in the buffer behind the
  // flushed keys.
  constexpr uint32_t SORT_MAX =
      (CAP * sizeof(KT) / (2 * sizeof(KT) + 2)) / kThreads * kThreads;
  auto sorted_place = [&](uint32_t n) {
    if constexpr (!SMALL) {
      constexpr int FS = SORT_MAX / kThreads;
      constexpr uint32_t CPT = CURS / kThreads;  // cursors per owning thread
      constexpr uint32_t NONE32 = 0xFFFFFFFFu;
      static_assert(FS >= 1 && CURS % kThreads == 0 && CURS <= 65536, "sorted flush shape");
      KT* const srt = kb + SORT_MAX;
      uint16_t* const su = reinterpret_cast<uint16_t*>(kb + 2 * SORT_MAX);
      static_assert(2 * SORT_MAX * sizeof(KT) + 2 * SORT_MAX <= CAP * sizeof(KT), "sorted flush fits");
      const uint32_t UC = ucur, wid = tid >> 6;
      uint32_t cb[CPT];
#pragma unroll
      for (uint32_t i = 0; i < CPT; ++i) {
        const uint32_t u = CPT * tid + i;
        cb[i] = u < UC ? cur[u] : 0u;
      }
      __syncthreads();  // the cursors are read before they advance
      KT kk[FS];
      uint32_t bb[FS], pp[FS];
#pragma unroll
      for (int f = 0; f < FS; ++f) kk[f] = kb[min(tid + f * kThreads, n - 1)];
      bucket_batch<KT, C, FS>(bk, lut, kk, bb);
#pragma unroll
      for (int f = 0; f < FS; ++f) {
        const uint32_t w = abits[bb[f] >> 5];
        const uint32_t u = wbase[bb[f] >> 5] + __popc(w & ((1u << (bb[f] & 31u)) - 1u));
        bb[f] = NONE32;
        if (tid + f * kThreads < n) {
          if (u - wlo < CURS) {
            pp[f] = atomicAdd(&cur[u - wlo], 1u);
            bb[f] = u - wlo;
          } else {
            rseg[kinfo[u].x + atomicAdd(&row[kinfo[u].y], 1u)] = kk[f];
          }
        }
      }
      __syncthreads();  // the cursors have advanced by this flush's keys
      uint32_t c[CPT], sum = 0;
#pragma unroll
      for (uint32_t i = 0; i < CPT; ++i) {
        const uint32_t u = CPT * tid + i;
        c[i] = u < UC ? cur[u] - cb[i] : 0u;
        sum += c[i];
      }
      const uint32_t inc = wave_incl_scan(sum);
      if (lane_id() == 63) dsum[wid] = inc;
      __syncthreads();  // wavefront totals written
      const uint32_t ws = lane_id() < kThreads / 64 ? dsum[lane_id()] : 0u;
      const uint32_t wi = wave_incl_scan(ws);
      uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)(wi - ws), (int)wid) + inc - sum;
      const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)wi, kThreads / 64 - 1);
#pragma unroll
      for (uint32_t i = 0; i < CPT; ++i) {
        const uint32_t u = CPT * tid + i;
        if (u < UC) cur[u] = cb[i] - t;
        t += c[i];
      }
      __syncthreads();  // (start - first slot) of every child written
#pragma unroll
      for (int f = 0; f < FS; ++f) {
        if (bb[f] != NONE32) {
          const uint32_t j = pp[f] - cur[bb[f]];
          srt[j] = kk[f];
          su[j] = (uint16_t)bb[f];
        }
      }
      __syncthreads();  // the flush is sorted
      uint32_t j = tid;
      for (; j + 3 * kThreads < tot; j += 4 * kThreads) {
        uint32_t uu[4], dd[4];
        KT ks[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          uu[f] = su[j + f * kThreads];
          ks[f] = srt[j + f * kThreads];
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) dd[f] = cur[uu[f]];
#pragma unroll
        for (int f = 0; f < 4; ++f) rseg[dd[f] + j + f * kThreads] = ks[f];
      }
      for (; j < tot; j += kThreads) rseg[cur[su[j]] + j] = srt[j];
      __syncthreads();  // every store has read its child's word
#pragma unroll
      for (uint32_t i = 0; i < CPT; ++i) {
        const uint32_t u = CPT * tid + i;
        if (u < UC) cur[u] = cb[i] + c[i];
      }
    }
  };
  auto flush = [&](uint32_t n) {
    // (sorting pays with many children; at 1024 k of a 2^28 N(0,1) vector
    // it costs 6 us, at 4096 k it saves 70 us)
    if (!SMALL && U > kDenseChildren && n <= SORT_MAX) {  // uniform
      if (n) sorted_place(n);
      __syncthreads();
      if (tid == 0) nbuf = 0;
      __syncthreads();
      return;
    }
    constexpr int F = SMALL ? 4 : 8;
    uint32_t i = tid;
    for (; i + (F - 1) * kThreads < n; i += F * kThreads) {
      KT kk[F];
      uint32_t bb[F], w[F], u[F];
#pragma unroll
      for (int f = 0; f < F; ++f) kk[f] = kb[i + f * kThreads];
      bucket_batch<KT, C, F>(bk, lut, kk, bb);
#pragma unroll
      for (int f = 0; f < F; ++f) w[f] = abits[bb[f] >> 5];
#pragma unroll
      for (int f = 0; f < F; ++f) u[f] = wbase[bb[f] >> 5] + __popc(w[f] & ((1u << (bb[f] & 31u)) - 1u));
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const uint32_t pos = u[f] - wlo < CURS
                                 ? atomicAdd(&cur[u[f] - wlo], 1u)
                                 : kinfo[u[f]].x + atomicAdd(&row[kinfo[u[f]].y], 1u);
        rseg[pos] = kk[f];
      }
    }
    for (; i < n; i += kThreads) place(kb[i]);
    __syncthreads();
    if (tid == 0) nbuf = 0;
    __syncthreads();
  };
  // head and tail: fewer than VEC keys each (a thread may own one of each)
  auto edges = [&]() {
    if (tid < head) place(p[tid]);
    if (done + tid < len) place(p[done + tid]);
  };
  auto main_loop = [&](auto left, auto hvt) {
    for (uint64_t r = 0; r < rounds; ++r) {
      KT k[N];
      unpack<KT, kRound>(v, k);
      uint32_t b[N], act[N];
      bucket_batch1<KT, C, N, decltype(left)::value, decltype(hvt)::value>(bk, lut, k, b);
#pragma unroll
      for (int i = 0; i < N; ++i) act[i] = act_of[b[i]];
      if (r >= full_rounds) {  // uniform: the chunk's last, partial round
#pragma unroll
        for (int i = 0; i < N; ++i) act[i] &= (okmask >> i) & 1u;
      }
      uint32_t cl = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) cl += act[i];
      const uint32_t inc = wave_incl_scan(cl);
      const uint32_t tot = wave_last(inc);  // uniform
      if (tot) {
        uint32_t wb = 0;
        if (lane == 0) wb = atomicAdd(&nbuf, tot);
        KT* dst = kb + ((uint32_t)__builtin_amdgcn_readfirstlane((int)wb) + inc - cl);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          if (act[i]) *dst++ = k[i];
        }
      }
      if (r + 1 < rounds) load_round(r + 1);
      __syncthreads();
      const uint32_t nb = nbuf;
      __syncthreads();  // every thread has read nbuf before the next round appends
      if (nb > FLUSH_AT) flush(nb);  // uniform
    }
  };
  // Dense segments (most keys active): a key-per-line scatter would cost one
  // L2 write request per key.  Instead each round's active keys are counted
  // per child (the LDS atomic's return value is the key's rank in the
  // child), the children's counts are scanned, the keys are written to LDS
  // in child order, and the sorted round is stored with consecutive
  // threads on consecutive slots of a child, so a wavefront's stores cover
  // a few lines.  Children beyond the LDS cursors place their keys directly.
  auto dense_loop = [&](auto left, auto hvt) {
    if constexpr (!SMALL) {
      constexpr uint32_t NONE32 = 0xFFFFFFFFu;
      constexpr uint32_t DC = kDenseChildren;
      uint32_t* ts = reinterpret_cast<uint32_t*>(kb);    // child's first slot in the sorted round
      uint32_t* dl = ts + DC;                             // child's slice position - first slot
      uint32_t* cnt = dl + DC;                            // child's keys in the round
      KT* srt = reinterpret_cast<KT*>(cnt + DC);          // the round's keys, by child
      uint16_t* su = reinterpret_cast<uint16_t*>(abyte);  // child of each sorted key
      static_assert(3 * DC * 4 + ROUND_MAX * sizeof(KT) <= Cfg::template buf_bytes<KT>,
                    "dense round must fit the key buffer");
      static_assert(ROUND_MAX * 2 <= sizeof(abyte), "dense round's child ids must fit the byte table");
      static_assert(ROUND_MAX <= 65536 && DC <= CURS && DC <= kThreads, "16-bit ranks, children");
      const uint32_t UC = U < DC ? U : DC;
      const uint32_t wid = tid >> 6;
      if (tid < UC) cnt[tid] = 0;
      __syncthreads();
      for (uint64_t r = 0; r < rounds; ++r) {
        KT k[N];
        unpack<KT, kRound>(v, k);
        if (r + 1 < rounds) load_round(r + 1);
        uint32_t b[N];
        bucket_batch1<KT, C, N, decltype(left)::value, decltype(hvt)::value>(bk, lut, k, b);
        const bool part = r >= full_rounds;  // uniform
        // b[i] becomes (child << 16) | rank in the child, or NONE32.  (The
        // counters were cleared before the previous round's last barrier,
        // and every thread has finished that round's stores before any
        // passes the next barrier, so three barriers a round suffice.)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const uint32_t w = abits[b[i] >> 5], sh = b[i] & 31u;
          const bool a = ((w >> sh) & 1u) && (!part || ((okmask >> i) & 1u));
          const uint32_t u = wbase[b[i] >> 5] + __popc(w & ((1u << sh) - 1u));
          uint32_t x = NONE32;
          if (a) {
            if (u < DC) {
              x = (u << 16) | atomicAdd(&cnt[u], 1u);
            } else {
              rseg[kinfo[u].x + atomicAdd(&row[kinfo[u].y], 1u)] = k[i];
            }
          }
          b[i] = x;
        }
        __syncthreads();  // counts final
        const uint32_t c = tid < UC ? cnt[tid] : 0u;
        const uint32_t inc = wave_incl_scan(c);
        if (lane == 63) dsum[wid] = inc;
        __syncthreads();  // wavefront totals written
        const uint32_t ws = lane < kThreads / 64 ? dsum[lane] : 0u;
        const uint32_t wi = wave_incl_scan(ws);
        const uint32_t wpre = (uint32_t)__builtin_amdgcn_readlane((int)(wi - ws), (int)wid);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)wi, kThreads / 64 - 1);
        if (tid < UC) {
          const uint32_t s = wpre + inc - c;
          ts[tid] = s;
          dl[tid] = cur[tid] - s;
          cur[tid] += c;
          cnt[tid] = 0;
        }
        __syncthreads();  // first slots written
#pragma unroll
        for (int i = 0; i < N; ++i) {
          if (b[i] != NONE32) {
            const uint32_t u = b[i] >> 16;
            const uint32_t j = ts[u] + (b[i] & 0xFFFFu);
            srt[j] = k[i];
            su[j] = (uint16_t)u;
          }
        }
        __syncthreads();  // the round is sorted
        // 8-byte keys go out with non-temporal stores: a wavefront's run
        // covers whole lines, and the second level, which reads the reduced
        // set next, then finds no write-back pending (2^27 f64 radix: count2
        // and compact2 -67 us, compaction +9 us).  4-byte keys' runs end in
        // partial lines, which non-temporal stores write one by one (2^28
        // f32 radix: compaction +300 us), so they are stored plainly.
        auto put = [](KT* a, KT x) {
          if constexpr (sizeof(KT) == 8) __builtin_nontemporal_store(x, a);
          else *a = x;
        };
        // four slots per thread per step, so their LDS round trips overlap
        uint32_t j = tid;
        for (; j + 3 * kThreads < tot; j += 4 * kThreads) {
          uint32_t uu[4], dd[4];
          KT kk[4];
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            uu[f] = su[j + f * kThreads];
            kk[f] = srt[j + f * kThreads];
          }
#pragma unroll
          for (int f = 0; f < 4; ++f) dd[f] = dl[uu[f]];
#pragma unroll
          for (int f = 0; f < 4; ++f) put(&rseg[dd[f] + j + f * kThreads], kk[f]);
        }
        for (; j < tot; j += kThreads) put(&rseg[dl[su[j]] + j], srt[j]);
      }
    }
  };
  // dense: at least a quarter of the keys active, at most 1024 children
  // (so one cursor window)
  const bool dense = U <= kDenseChildren && ((uint64_t)st.nred << kDenseLog2) >= sd.len;
  if (!SMALL && dense) {  // uniform
    edges();
    if (!st.hvon) {
      if (st.s <= 24) dense_loop(std::true_type{}, std::false_type{});
      else dense_loop(std::false_type{}, std::false_type{});
    } else {
      if (st.s <= 24) dense_loop(std::true_type{}, std::true_type{});
      else dense_loop(std::false_type{}, std::true_type{});
    }
    return;
  }
  // one pass over the chunk: edges, rounds, the last flush
  auto pass = [&]() {
    edges();
    // uniform: the sub-field shift direction and the heavy-value term
    if (!st.hvon) {
      if (st.s <= 24) main_loop(std::true_type{}, std::false_type{});
      else main_loop(std::false_type{}, std::false_type{});
    } else {
      if (st.s <= 24) main_loop(std::true_type{}, std::true_type{});
      else main_loop(std::false_type{}, std::true_type{});
    }
    const uint32_t nb = nbuf;  // stable: nothing appends any more
    __syncthreads();
    flush(nb);  // ends with a barrier: every key of the pass is placed
  };
  if constexpr (!WIN) {
    pass();
  } else {
    for (uint32_t win = 0;;) {  // uniform
      // the chunk's address taken afresh in each window (an empty asm the
      // compiler cannot see through), so per-lane load addresses are not
      // kept live from one window to the next
      {
        uint64_t a = (uint64_t)(uintptr_t)pv;
        asm volatile("" : "+s"(a));
        pv = reinterpret_cast<const uint4*>((uintptr_t)a);
      }
      pass();
      if (++win == nwin) break;
      wlo = win * CURS;
      ucur = U - wlo < CURS ? U - wlo : CURS;
      if (rounds) load_round(0);
      window_tables();
      __syncthreads();
    }
  }
}

#define GGKS_COMPACT_PARAMS                                                                   \
  const KT *__restrict__ src, const SegDesc *__restrict__ segs,                               \
      const SegState *__restrict__ states, const Chunk *__restrict__ chunks,                  \
      const uint32_t *__restrict__ lut_pool, uint32_t *__restrict__ rows,                     \
      const uint32_t *__restrict__ bits_pool, const uint32_t *__restrict__ wbase_pool,        \
      const uint2 *__restrict__ cinfo, KT *__restrict__ red
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_compact(GGKS_COMPACT_PARAMS) {
  compact_chunk<KT, C, false, false>(src, segs, states, chunks, lut_pool, rows, bits_pool,
                                     wbase_pool, cinfo, red);
}
// with cursor windows: for plans with a segment of more ranks than LDS
// cursors (a separate instance, so the window loop costs the others nothing)
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_compact_win(GGKS_COMPACT_PARAMS) {
  compact_chunk<KT, C, false, true>(src, segs, states, chunks, lut_pool, rows, bits_pool,
                                    wbase_pool, cinfo, red);
}
// at most 64 VGPRs: two 1024-thread work-groups per CU
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_compact_small(
    GGKS_COMPACT_PARAMS) {
  compact_chunk<KT, C, true, false>(src, segs, states, chunks, lut_pool, rows, bits_pool,
                                    wbase_pool, cinfo, red);
}
#undef GGKS_COMPACT_PARAMS

// ------------------------------------------------------------ second level
// Work-groups of both kernels walk the chunk table (count in l2count[1],
// written by k_findk) with a grid stride, so one launch of a fixed grid
// serves any number of heavy children.

// Resolve heavy child h from its chunk rows (run by the work-group that
// counted its last chunk): bucket totals and their prefix, each rank's
// bucket (a single-key bucket answers the rank), the second-level children
// (slots: the heavy child's own, then free slots of its segment), their
// ranks regrouped into cq2, each chunk's start inside every active bucket,
// and the active bitmap with its child bases for k_compact2.
template <typename KT, int C>
__device__ void resolve_heavy(uint32_t h, Heavy* __restrict__ heavy, uint32_t* __restrict__ rows2,
                              const Query* __restrict__ cq, Query* __restrict__ cq2,
                              uint32_t* __restrict__ qb, uint32_t* __restrict__ qr,
                              Child* __restrict__ children, SegState* __restrict__ states,
                              const SegDesc* __restrict__ segs, uint32_t* __restrict__ bits2,
                              uint32_t* __restrict__ wb2, KT* __restrict__ out,
                              uint32_t* hist, uint32_t* cnt, uint32_t* bits, uint32_t* wb,
                              uint32_t* wsum, uint32_t* misc) {
  constexpr uint32_t NW = kBuckets2 / 32;
  const uint32_t tid = threadIdx.x;
  const Heavy H = heavy[h];
  const KT lo = (KT)H.lo, hi = (KT)H.hi;
  const uint32_t sh = H.shift;
  for (uint32_t b = tid; b < kBuckets2; b += kThreads) {
    uint32_t t = 0;
    for (uint32_t c = 0; c < H.nchunks; ++c) t += rows2[(uint64_t)(H.chunk_base + c) * kBuckets2 + b];
    hist[b] = t;
  }
  for (uint32_t i = tid; i < NW; i += kThreads) bits[i] = 0;
  __syncthreads();
  const uint32_t total = block_scan_array(hist, kBuckets2, wsum);  // hist = prefix (synchronises)
  const uint32_t NONE = 0xFFFFFFFFu;
  auto count_of = [&](uint32_t b) { return (b + 1 < kBuckets2 ? hist[b + 1] : total) - hist[b]; };
  auto range_of = [&](uint32_t b, KT& a, KT& z) {
    a = (KT)(lo + ((KT)b << sh));
    const KT w = (KT)((KT(1) << sh) - 1);
    z = (KT)(hi - a) > w ? (KT)(a + w) : hi;
  };
  for (uint32_t i = tid; i < H.q_cnt; i += kThreads) {
    const Query q = cq[H.q_off + i];
    const uint32_t r1 = q.rank - 1;
    uint32_t a = 0, z = kBuckets2 - 1;  // largest b with hist[b] <= r1
    while (a < z) {
      const uint32_t mid = (a + z + 1) >> 1;
      if (hist[mid] <= r1) a = mid; else z = mid - 1;
    }
    KT klo, khi;
    range_of(a, klo, khi);
    if (klo == khi) {
      out[q.out] = from_key<C>(klo);
      qb[H.q_off + i] = NONE;
    } else {
      atomicOr(&bits[a >> 5], 1u << (a & 31));
      qb[H.q_off + i] = a;
      qr[H.q_off + i] = q.rank - hist[a];
    }
  }
  __syncthreads();
  uint32_t U2;
  const uint32_t ub = block_excl_scan(tid < NW ? (uint32_t)__popc(bits[tid]) : 0u, wsum, U2);
  if (tid < NW) {
    wb[tid] = ub;
    bits2[(uint64_t)h * NW + tid] = bits[tid];
    wb2[(uint64_t)h * NW + tid] = ub;
  }
  if (tid == 0) {
    misc[0] = U2 > 1 ? atomicAdd(&states[H.seg].nfree, U2 - 1) : 0u;
    heavy[h].nact = U2;
  }
  __syncthreads();
  if (U2 == 0) return;
  const uint32_t sbase = misc[0];
  if (tid == 0) heavy[h].slot_base = sbase;
  const uint32_t sq = segs[H.seg].q_off;
  auto slot_of = [&](uint32_t u) { return u == 0 ? H.slot : sq + sbase + u - 1; };
  // children: one thread per child finds its bucket from the bitmap
  for (uint32_t u = tid; u < U2; u += kThreads) {
    uint32_t a = 0, z = NW - 1;  // largest word with wb <= u
    while (a < z) {
      const uint32_t mid = (a + z + 1) >> 1;
      if (wb[mid] <= u) a = mid; else z = mid - 1;
    }
    uint32_t mw = bits[a];
    for (uint32_t r = u - wb[a]; r; --r) mw &= mw - 1;
    const uint32_t b = a * 32 + __ffs(mw) - 1;
    cnt[u] = 0;
    Child c;
    c.start = H.start + hist[b];  // the second reduced buffer mirrors the first
    c.len = count_of(b);
    KT klo, khi;
    range_of(b, klo, khi);
    c.lo = klo;
    c.hi = khi;
    c.q_off = 0;
    c.q_cnt = 0;
    c.src = 2;
    c.need_range = 0;
    c.bucket = b;
    c.seg = H.seg;
    children[slot_of(u)] = c;
    // each chunk's start inside this bucket's slice (relative to H.start)
    uint32_t run = hist[b];
    for (uint32_t ch = 0; ch < H.nchunks; ++ch) {
      uint32_t* e = &rows2[(uint64_t)(H.chunk_base + ch) * kBuckets2 + b];
      const uint32_t v = *e;
      *e = run;
      run += v;
    }
  }
  __syncthreads();
  // ranks grouped by child into cq2 (same index range as the heavy child's)
  for (uint32_t i = tid; i < H.q_cnt; i += kThreads) {
    const uint32_t b = qb[H.q_off + i];
    if (b == NONE) continue;
    const uint32_t u = wb[b >> 5] + __popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
    atomicAdd(&cnt[u], 1u);
    qb[H.q_off + i] = u;
  }
  __syncthreads();
  const uint32_t nq2 = block_scan_array(cnt, U2, wsum);  // cnt[u] = first position
  for (uint32_t u = tid; u < U2; u += kThreads) {
    const uint32_t first = cnt[u], next = u + 1 < U2 ? cnt[u + 1] : nq2;
    Child* cp = &children[slot_of(u)];  // written by this thread above
    cp->q_off = H.q_off + first;
    cp->q_cnt = next - first;
  }
  __syncthreads();
  for (uint32_t i = tid; i < H.q_cnt; i += kThreads) {
    const uint32_t u = qb[H.q_off + i];
    if (u == NONE) continue;
    const uint32_t pos = atomicAdd(&cnt[u], 1u);
    Query q;
    q.rank = qr[H.q_off + i];
    q.out = cq[H.q_off + i].out;
    cq2[H.q_off + pos] = q;
  }
}

// Second-level count: per chunk of a heavy child, a histogram of its keys
// (in the first reduced buffer) over the child's digit (key - lo) >> shift,
// written as the chunk's row.  The children are resolved by k_resolve2
// after the kernel boundary (a release fence and an arrival counter per
// chunk, the last chunk resolving its child, made the count pass of a
// gigabyte of heavy children 7x slower than its compaction pass).
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_count2(
    const KT* __restrict__ red, const Heavy* __restrict__ heavy,
    const Chunk2* __restrict__ chunks2, const uint32_t* __restrict__ l2count, uint32_t cmax,
    uint32_t* __restrict__ rows2) {
  __shared__ uint32_t hist[kBuckets2];
  const uint32_t tid = threadIdx.x;
  const uint32_t nchunks = min(l2count[1], cmax);
  for (uint32_t w = blockIdx.x; w < nchunks; w += gridDim.x) {  // uniform
    const Chunk2 ck = chunks2[w];
    const KT lo = (KT)heavy[ck.h].lo;
    const uint32_t sh = heavy[ck.h].shift;
    const uint64_t hstart = heavy[ck.h].start;
    for (uint32_t b = tid; b < kBuckets2; b += kThreads) hist[b] = 0;
    __syncthreads();
    for_each_elem(red + hstart + ck.start, ck.len, [&](KT raw) {
      atomicAdd(&hist[(uint32_t)((KT)(to_key<C>(raw) - lo) >> sh)], 1u);
    });
    __syncthreads();
    uint32_t* row = rows2 + (uint64_t)w * kBuckets2;
    for (uint32_t b = tid; b < kBuckets2; b += kThreads) row[b] = hist[b];
    __syncthreads();
  }
}

// Second-level resolution: one work-group per heavy child (grid-stride)
// sums its chunks' rows, answers single-key digits, lays out its
// second-level children and rewrites the rows as chunk offsets.
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_resolve2(
    Heavy* __restrict__ heavy, const uint32_t* __restrict__ l2count, uint32_t hmax,
    uint32_t* __restrict__ rows2, const Query* __restrict__ cq, Query* __restrict__ cq2,
    uint32_t* __restrict__ qb, uint32_t* __restrict__ qr, Child* __restrict__ children,
    SegState* __restrict__ states, const SegDesc* __restrict__ segs, uint32_t* __restrict__ bits2,
    uint32_t* __restrict__ wb2, KT* __restrict__ out) {
  __shared__ uint32_t hist[kBuckets2];
  __shared__ uint32_t cnt[kBuckets2];
  __shared__ uint32_t bits[kBuckets2 / 32], wb[kBuckets2 / 32];
  __shared__ uint32_t wsum[17], misc[2];
  const uint32_t nh = min(l2count[0], hmax);
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {  // uniform
    resolve_heavy<KT, C>(h, heavy, rows2, cq, cq2, qb, qr, children, states, segs, bits2, wb2, out,
                         hist, cnt, bits, wb, wsum, misc);
    __syncthreads();
  }
}

// Second-level compaction: the keys of each heavy child's active buckets,
// from the first reduced buffer into the second at the same offsets, each
// chunk at its own start inside every bucket (LDS cursors).
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_compact2(
    const KT* __restrict__ red, const Heavy* __restrict__ heavy,
    const Chunk2* __restrict__ chunks2, const uint32_t* __restrict__ l2count, uint32_t cmax,
    const uint32_t* __restrict__ rows2, const uint32_t* __restrict__ bits2,
    const uint32_t* __restrict__ wb2, KT* __restrict__ red2) {
  constexpr uint32_t NW = kBuckets2 / 32;
  __shared__ uint32_t cur[kBuckets2];
  __shared__ uint32_t bits[NW], wb[NW];
  const uint32_t tid = threadIdx.x;
  const uint32_t nchunks = min(l2count[1], cmax);
  for (uint32_t w = blockIdx.x; w < nchunks; w += gridDim.x) {  // uniform
    const Chunk2 ck = chunks2[w];
    if (heavy[ck.h].nact == 0) continue;  // uniform
    const KT lo = (KT)heavy[ck.h].lo;
    const uint32_t sh = heavy[ck.h].shift;
    const uint64_t hstart = heavy[ck.h].start;
    if (tid < NW) {
      const uint32_t m = bits2[(uint64_t)ck.h * NW + tid];
      uint32_t u = wb2[(uint64_t)ck.h * NW + tid];
      bits[tid] = m;
      wb[tid] = u;
      const uint32_t* row = rows2 + (uint64_t)w * kBuckets2 + tid * 32;
      for (uint32_t mm = m; mm; mm &= mm - 1) cur[u++] = row[__ffs(mm) - 1];
    }
    __syncthreads();
    KT* dst = red2 + hstart;
    for_each_elem(red + hstart + ck.start, ck.len, [&](KT raw) {
      const uint32_t b = (uint32_t)((KT)(to_key<C>(raw) - lo) >> sh);
      const uint32_t m = bits[b >> 5], bit = b & 31;
      if ((m >> bit) & 1u) {
        const uint32_t u = wb[b >> 5] + __popc(m & ((1u << bit) - 1u));
        dst[atomicAdd(&cur[u], 1u)] = raw;
      }
    });
    __syncthreads();
  }
}

// Plain streaming copy of nv 16-byte words (grid-stride).
__global__ __launch_bounds__(kThreads) void k_copy16(const uint4* __restrict__ src,
                                                     uint4* __restrict__ dst, uint64_t nv) {
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < nv; i += stride)
    dst[i] = stream_load(&src[i]);
}

// k_finish sorts a child held in LDS when it has many ranks or is short
// enough that a sort beats the digit passes.
constexpr uint32_t kSortLen = 4096;
constexpr uint32_t kFDigit = 11;                      // first digit: 2048 bins
constexpr uint32_t kFWords = (1u << kFDigit) + (1u << kFDigit) / 32;  // padded: bin b at b + b/32
constexpr uint32_t kFCollectWords = 1024;             // keys gathered after the first digit

// Visit keys p[0, len) with a work-group of kFThreads: 16-byte loads for the
// aligned body (eight in flight per thread), scalar head and tail.
template <typename KT, typename F>
__device__ __forceinline__ void group_for_each(const KT* __restrict__ p, uint64_t len, F&& f) {
  constexpr uint32_t VEC = 16 / sizeof(KT);
  constexpr int D = 8;
  const uint32_t tid = threadIdx.x;
  uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
  uint64_t head = mis ? (VEC - mis) : 0;
  if (head > len) head = len;
  if (tid < head) f(p[tid]);
  const uint64_t nvec = (len - head) / VEC;
  const uint4* __restrict__ pv = reinterpret_cast<const uint4*>(p + head);
  auto apply = [&](const uint4& v) {
    if (sizeof(KT) == 4) {
      f((KT)v.x); f((KT)v.y); f((KT)v.z); f((KT)v.w);
    } else {
      f((KT)(((uint64_t)v.y << 32) | v.x));
      f((KT)(((uint64_t)v.w << 32) | v.z));
    }
  };
  uint64_t i = tid;
  for (; i + (D - 1) * kFThreads < nvec; i += D * kFThreads) {
    uint4 v[D];
#pragma unroll
    for (int u = 0; u < D; ++u) v[u] = pv[i + u * kFThreads];
#pragma unroll
    for (int u = 0; u < D; ++u) apply(v[u]);
  }
  for (; i < nvec; i += kFThreads) apply(pv[i]);
  const uint64_t done = head + nvec * VEC;
  if (done + tid < len) f(p[done + tid]);
}

// A child's keys held by a work-group of kFThreads: R 16-byte words per
// thread in registers and X more in LDS (each thread's own slots of `ext`,
// X * kFThreads words: no other thread reads them), plus at most one
// unaligned head key and one tail key, so a child of up to (R + X) *
// kFThreads * 16 bytes is read from memory once for all passes.  Visiting
// order differs from group_for_each; no pass depends on it.
template <typename KT, int R, int X = 0>
struct GroupKeys {
  static constexpr uint32_t VEC = 16 / sizeof(KT);
  uint4 v[R];
  uint4* ext;
  KT hk, tk;
  uint32_t nv;  // valid words of this thread: v[0, min(nv, R)), then ext
  bool hh, ht;
  // whether a child at p of len keys fits (uniform)
  __device__ static bool fits(const KT* p, uint64_t len) {
    const uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
    uint64_t head = mis ? (VEC - mis) : 0;
    if (head > len) head = len;
    return (len - head) / VEC <= (uint64_t)(R + X) * kFThreads;
  }
  __device__ void load(const KT* __restrict__ p, uint64_t len, uint4* lds_ext = nullptr) {
    const uint32_t tid = threadIdx.x;
    const uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
    uint64_t head = mis ? (VEC - mis) : 0;
    if (head > len) head = len;
    const uint32_t nvec = (uint32_t)((len - head) / VEC);  // <= (R + X) * kFThreads
    const uint4* __restrict__ pv = reinterpret_cast<const uint4*>(p + head);
    nv = nvec > tid ? (nvec - tid + kFThreads - 1) / kFThreads : 0u;
    ext = lds_ext;
    if (X > 0 && nv > (uint32_t)R) {
      // the LDS words first, stored before the register words are loaded, so
      // the two sets are never live in registers together (a thread with no
      // LDS words skips this; only children longer than the registers hold
      // pay the second round trip)
#pragma unroll
      for (int x = 0; x < X; ++x)
        if ((uint32_t)(R + x) < nv) ext[x * kFThreads + tid] = pv[tid + (R + x) * kFThreads];
      __asm__ __volatile__("" ::: "memory");
    }
#pragma unroll
    for (int u = 0; u < R; ++u) v[u] = (uint32_t)u < nv ? pv[tid + u * kFThreads] : make_uint4(0, 0, 0, 0);
    hh = tid < head;
    hk = hh ? p[tid] : KT(0);
    const uint64_t done = head + (uint64_t)nvec * VEC;
    ht = done + tid < len;
    tk = ht ? p[done + tid] : KT(0);
  }
  template <typename F>
  __device__ __forceinline__ void for_each(F&& f) const {
    if (hh) f(hk);
    auto apply = [&](const uint4& w) {
      if (sizeof(KT) == 4) {
        f((KT)w.x); f((KT)w.y); f((KT)w.z); f((KT)w.w);
      } else {
        f((KT)(((uint64_t)w.y << 32) | w.x));
        f((KT)(((uint64_t)w.w << 32) | w.z));
      }
    };
#pragma unroll
    for (int u = 0; u < R; ++u)
      if ((uint32_t)u < nv) apply(v[u]);
    if (X > 0 && nv > (uint32_t)R) {
      // read the LDS words afresh in every pass: kept live across passes
      // (the compiler may merge the identical reads) they would spill
      __asm__ __volatile__("" ::: "memory");
#pragma unroll
      for (int x = 0; x < X; ++x)
        if ((uint32_t)(R + x) < nv) apply(ext[x * kFThreads + threadIdx.x]);
    }
    if (ht) f(tk);
  }
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rank r (0-based) among 256 LDS bins h[0, 2^d) counted by a wavefront: the
// bin holding it and the number of values in lower bins.
__device__ __forceinline__ void wave_pick256(const uint32_t* h, uint32_t lane, uint32_t r,
                                             uint32_t& bin, uint32_t& below) {
  const uint32_t v0 = h[lane * 4], v1 = h[lane * 4 + 1], v2 = h[lane * 4 + 2], v3 = h[lane * 4 + 3];
  const uint32_t sum = v0 + v1 + v2 + v3;
  const uint32_t e0 = wave_incl_scan(sum) - sum;
  const uint32_t e1 = e0 + v0, e2 = e1 + v1, e3 = e2 + v2;
  int hit = -1;
  uint32_t start = 0;
  if (r >= e0 && r < e1) { hit = 0; start = e0; }
  if (r >= e1 && r < e2) { hit = 1; start = e1; }
  if (r >= e2 && r < e3) { hit = 2; start = e2; }
  if (r >= e3 && r < e3 + v3) { hit = 3; start = e3; }
  const int src = __ffsll((unsigned long long)__ballot(hit >= 0)) - 1;
  bin = (uint32_t)__shfl((int)(lane * 4 + hit), src, 64);
  below = (uint32_t)__shfl((int)start, src, 64);
}

// Finisher for children with at most kWaveRanks ranks, and for children
// with many ranks that fit kSortCap keys: one 256-thread work-group per child
// (several per CU).  The first kind takes two passes over the child's keys
// (held in registers when they fit).  Pass 1 builds an 11-bit top-digit
// histogram of (key - lo) in a padded LDS table (lane l's 32 bins are bank
// disjoint); wavefront 0 then finds each rank's top digit.  Pass 2 gathers,
// per rank, the keys holding that digit (a handful, as the child is one
// bucket) into LDS, where wavefront 0 resolves the rank: by counting for up
// to 64 keys, by 8-bit radix passes over LDS up to the gather capacity, and
// beyond it (heavy duplicates of one digit) by 8-bit radix passes over the
// child in memory.  The second kind is sorted in LDS and every rank read off.
// R is the number of 16-byte words of a child a thread holds in registers.
template <typename KT, int C, int R, int X>
__device__ __forceinline__ void finish_small_body(const KT* __restrict__ input,
                                                  const KT* __restrict__ red,
                                                  const KT* __restrict__ red2,
                                                  const Child* __restrict__ children,
                                                  uint32_t nchildren,
                                                  const Query* __restrict__ cq0,
                                                  const Query* __restrict__ cq2,
                                                  KT* __restrict__ out) {
  constexpr uint32_t W = KeyBits<KT>::W;
  constexpr uint32_t CAP = kFCollectWords * 4 / sizeof(KT);  // gathered keys
  static_assert(kSortCap * sizeof(KT) <= kFWords * 4, "a sorted child fits the histogram table");
  __shared__ __align__(16) uint32_t h[kFWords];  // histogram, gathered keys, or a child to sort
  __shared__ KT red_lo[kFThreads / 64], red_hi[kFThreads / 64];
  __shared__ uint32_t res[4 * kWaveRanks + 2];  // per rank: digit, residual, count; gathered
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const bool w0 = tid < 64;
  KT* buf = reinterpret_cast<KT*>(h);  // gathered keys (after the histogram is read)
  uint32_t* bins = h + kFCollectWords;  // 256 bins for the LDS radix passes
  __shared__ uint32_t list[kFThreads + 1];  // this window's children for this finisher; count
  __shared__ uint4 gk_ext[X > 0 ? X * kFThreads : 1];  // the child's keys past the registers
  // this work-group's children are blockIdx.x, + gridDim.x, ...; they are
  // screened kFThreads at a time, one per thread, and only those for this
  // finisher are walked
  const uint32_t ncand =
      nchildren > blockIdx.x ? (nchildren - blockIdx.x + gridDim.x - 1) / gridDim.x : 0u;
  for (uint32_t wstart = 0; wstart < ncand; wstart += kFThreads) {  // uniform
    if (tid == 0) list[kFThreads] = 0;
    __syncthreads();
    if (wstart + tid < ncand) {
      const uint32_t ci = blockIdx.x + (wstart + tid) * gridDim.x;
      const Child cc = children[ci];
      if (cc.q_cnt != 0 && (wave_child(cc) || sort_child(cc)))
        list[atomicAdd(&list[kFThreads], 1u)] = ci;
    }
    __syncthreads();
    const uint32_t nl = list[kFThreads];
    for (uint32_t li = 0; li < nl; ++li) {
      const Child c = children[list[li]];
      // source: the input, the first or the second reduced buffer; a
      // second-level child's ranks are in cq2
      const KT* srcp = (c.src == 0 ? input : (c.src == 1 ? red : red2)) + c.start;
      const Query* __restrict__ cq = c.src == 2 ? cq2 : cq0;
      if (sort_child(c)) {  // uniform
        // many ranks: the converted keys sorted ascending in LDS (bitonic,
        // one compare-exchange pair per thread step); rank r is the r-th key
        KT* keys = reinterpret_cast<KT*>(h);
        const uint32_t len = (uint32_t)c.len;
        uint32_t np = 1;
        while (np < len) np <<= 1;
        for (uint32_t i = tid; i < np; i += kFThreads)
          keys[i] = i < len ? to_key<C>(srcp[i]) : KeyBits<KT>::MAX;
        __syncthreads();
        lds_bitonic_sort<kFThreads, kSortCap / 2 / kFThreads>(keys, np);
        for (uint32_t i = tid; i < c.q_cnt; i += kFThreads) {
          const Query q = cq[c.q_off + i];
          out[q.out] = from_key<C>(keys[q.rank - 1]);
        }
        __syncthreads();
        continue;
      }
      // the passes over the child's keys, given a visitor of its keys
      auto body = [&](auto&& each) {
        KT lo = (KT)c.lo, hi = (KT)c.hi;
        if (c.need_range) {
          KT a = KeyBits<KT>::MAX, b = 0;
          each([&](KT raw) {
            const KT k = to_key<C>(raw);
            a = k < a ? k : a;
            b = k > b ? k : b;
          });
          for (int d = 32; d >= 1; d >>= 1) {
            const KT x = __shfl_xor(a, d, 64), y = __shfl_xor(b, d, 64);
            a = x < a ? x : a;
            b = y > b ? y : b;
          }
          if (lane == 0) {
            red_lo[tid >> 6] = a;
            red_hi[tid >> 6] = b;
          }
          __syncthreads();
          lo = red_lo[0];
          hi = red_hi[0];
          for (uint32_t w = 1; w < kFThreads / 64; ++w) {
            lo = red_lo[w] < lo ? red_lo[w] : lo;
            hi = red_hi[w] > hi ? red_hi[w] : hi;
          }
        }
        const int nbits = bitlen<KT>((KT)(hi - lo));
        if (nbits == 0) {  // every key equal
          for (uint32_t qi = tid; qi < c.q_cnt; qi += kFThreads)
            out[cq[c.q_off + qi].out] = from_key<C>(lo);
          return;
        }
        // pass 1: histogram of the top digit
        const int d1 = nbits < (int)kFDigit ? nbits : (int)kFDigit;
        const int sh1 = nbits - d1;
        for (uint32_t j = tid; j < kFWords; j += kFThreads) h[j] = 0;
        __syncthreads();
        each([&](KT raw) {
          const uint32_t bn = (uint32_t)((KT)(to_key<C>(raw) - lo) >> sh1);
          atomicAdd(&h[bn + (bn >> 5)], 1u);
        });
        __syncthreads();
        if (w0) {  // each rank's top digit, residual rank and digit count
          uint32_t lsum = 0;  // lane's bins lane*32 .. lane*32 + 31
    #pragma unroll 8
          for (uint32_t j = 0; j < 32; ++j) lsum += h[lane * 33 + j];
          const uint32_t lex = wave_incl_scan(lsum) - lsum;
          for (uint32_t qi = 0; qi < c.q_cnt; ++qi) {
            const uint32_t r = cq[c.q_off + qi].rank - 1;
            const int L = __ffsll((unsigned long long)__ballot(r >= lex && r < lex + lsum)) - 1;
            const uint32_t eL = (uint32_t)__shfl((int)lex, L, 64);
            const uint32_t v = lane < 32 ? h[(uint32_t)L * 33 + lane] : 0u;
            const uint32_t ex = wave_incl_scan(v) - v + eL;
            const int j = __ffsll((unsigned long long)__ballot(lane < 32 && r >= ex && r < ex + v)) - 1;
            const uint32_t exj = (uint32_t)__shfl((int)ex, j, 64);  // every lane takes part
            const uint32_t vj = (uint32_t)__shfl((int)v, j, 64);
            if (lane == 0) {
              res[3 * qi] = (uint32_t)L * 32 + (uint32_t)j;
              res[3 * qi + 1] = r - exj;
              res[3 * qi + 2] = vj;
            }
          }
        }
        __syncthreads();
        const KT lowmask = sh1 == 0 ? KT(0) : (KT)(KeyBits<KT>::MAX >> (W - sh1));
        for (uint32_t qi = 0; qi < c.q_cnt; ++qi) {
          const Query q = cq[c.q_off + qi];
          const KT top = (KT)res[3 * qi];
          uint32_t r = res[3 * qi + 1];
          const uint32_t n2 = res[3 * qi + 2];
          KT low = 0;  // the rank's key minus lo, below the top digit
          if (sh1 > 0 && n2 <= CAP) {
            // pass 2: gather the keys with this top digit
            if (tid == 0) res[4 * kWaveRanks] = 0;
            __syncthreads();
            each([&](KT raw) {
              const KT rel = (KT)(to_key<C>(raw) - lo);
              if ((KT)(rel >> sh1) == top) buf[atomicAdd(&res[4 * kWaveRanks], 1u)] = rel & lowmask;
            });
            __syncthreads();
            if (w0) {
              if (n2 <= 64) {
                // rank by counting: lane i holds key i; ties broken by position
                const KT mine = lane < n2 ? buf[lane] : KeyBits<KT>::MAX;
                uint32_t below = 0;
                for (uint32_t j = 0; j < n2; ++j) {
                  const KT o = buf[j];
                  below += (o < mine || (o == mine && j < lane)) ? 1u : 0u;
                }
                const int src = __ffsll((unsigned long long)__ballot(lane < n2 && below == r)) - 1;
                low = (KT)__shfl(mine, src, 64);
              } else {
                int rem = sh1;
                KT pre = 0;
                while (rem > 0) {
                  const int d = rem < 8 ? rem : 8;
                  const int shift = rem - d;
                  const int hshift = shift + d;
    #pragma unroll
                  for (int j = 0; j < 4; ++j) bins[lane * 4 + j] = 0;
                  wave_sync();
                  for (uint32_t i = lane; i < n2; i += 64) {
                    const KT v = buf[i];
                    const KT t = hshift >= (int)W ? KT(0) : (KT)(v >> hshift);
                    if (t == pre) atomicAdd(&bins[(uint32_t)(v >> shift) & ((1u << d) - 1u)], 1u);
                  }
                  wave_sync();
                  uint32_t bn, below;
                  wave_pick256(bins, lane, r, bn, below);
                  r -= below;
                  pre = (KT)((pre << d) | (KT)bn);
                  rem = shift;
                  wave_sync();
                }
                low = pre;
              }
            }
            __syncthreads();
          } else if (sh1 > 0) {
            // many keys share the top digit: 8-bit passes over the child in memory
            int rem = sh1;
            KT pre = top;
            while (rem > 0) {
              const int d = rem < 8 ? rem : 8;
              const int shift = rem - d;
              const int hshift = shift + d;
              if (tid < 256) bins[tid] = 0;
              __syncthreads();
              each([&](KT raw) {
                const KT rel = (KT)(to_key<C>(raw) - lo);
                if ((KT)(rel >> hshift) == pre)
                  atomicAdd(&bins[(uint32_t)(rel >> shift) & ((1u << d) - 1u)], 1u);
              });
              __syncthreads();
              if (w0) {
                uint32_t bn, below;
                wave_pick256(bins, lane, r, bn, below);
                if (lane == 0) {
                  res[4 * kWaveRanks + 1] = bn;
                  res[4 * kWaveRanks] = below;
                }
              }
              __syncthreads();
              r -= res[4 * kWaveRanks];
              pre = (KT)((pre << d) | (KT)res[4 * kWaveRanks + 1]);
              rem = shift;
              __syncthreads();
            }
            low = (KT)(pre & lowmask);
          }
          if (tid == 0) out[q.out] = from_key<C>((KT)(lo + ((KT)(top << sh1) | low)));
        }
      };
      // from registers when the child fits (read from memory once), else from
      // memory in every pass
      if (GroupKeys<KT, R, X>::fits(srcp, c.len)) {  // uniform
        GroupKeys<KT, R, X> gk;
        gk.load(srcp, c.len, gk_ext);
        body([&](auto&& f) { gk.for_each(f); });
      } else {
        body([&](auto&& f) { group_for_each(srcp, c.len, f); });
      }
      __syncthreads();
    }
    __syncthreads();  // every thread is done with the list before it is refilled
  }
}

#define GGKS_FINISH_PARAMS                                                                    \
  const KT *__restrict__ input, const KT *__restrict__ red, const KT *__restrict__ red2,      \
      const Child *__restrict__ children, uint32_t nchildren, const Query *__restrict__ cq,   \
      const Query *__restrict__ cq2, KT *__restrict__ out
template <typename KT, int C>
__global__ __launch_bounds__(kFThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_finish_small(
    GGKS_FINISH_PARAMS) {
  finish_small_body<KT, C, kFRegWords<KT>, kFLdsWords>(input, red, red2, children, nchildren, cq, cq2,
                                                        out);
}
template <typename KT, int C>
__global__ __launch_bounds__(kFThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_finish_small_lite(
    GGKS_FINISH_PARAMS) {
  finish_small_body<KT, C, kFLiteWords, 0>(input, red, red2, children, nchildren, cq, cq2, out);
}
#undef GGKS_FINISH_PARAMS

// Finisher: exact multi-rank radix select inside each child.
// LDS layout (byte offsets, all 16-aligned): keys | hist | queries | query
// prefixes | group prefixes | min/max scratch | group starts | misc.
constexpr uint32_t kFinishOffHist = kFinishLdsBytes;
constexpr uint32_t kFinishOffQv = kFinishOffHist + kGroupMax * 256 * 4;
constexpr uint32_t kFinishOffQpre = kFinishOffQv + kQCap * 8;
constexpr uint32_t kFinishOffGpre = kFinishOffQpre + kQCap * 8;
constexpr uint32_t kFinishOffRed = kFinishOffGpre + (kGroupMax + 2) * 8;
constexpr uint32_t kFinishOffGfirst = kFinishOffRed + 32 * 8;
constexpr uint32_t kFinishOffMisc = kFinishOffGfirst + ((kQCap + 1) * 4 + 15) / 16 * 16;
constexpr uint32_t kFinishOffList = kFinishOffMisc + 32 * 4;
constexpr uint32_t kFinishLdsTotal = kFinishOffList + (kThreads + 4) * 4;
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_finish(const KT* __restrict__ input,
                                                     const KT* __restrict__ red,
                                                     const KT* __restrict__ red2,
                                                     const Child* __restrict__ children,
                                                     uint32_t nchildren,
                                                     const Query* __restrict__ cq0,
                                                     const Query* __restrict__ cq2,
                                                     KT* __restrict__ out,
                                                     const uint32_t* __restrict__ bigw) {
  // nothing to do when k_findk counted no child for this finisher
  if (bigw && *bigw == 0) return;  // uniform
  constexpr uint32_t W = KeyBits<KT>::W;
  constexpr uint32_t CAP = kFinishLdsBytes / sizeof(KT);
  extern __shared__ __align__(16) uint32_t lds[];
  char* base = reinterpret_cast<char*>(lds);
  KT* keys = reinterpret_cast<KT*>(base);
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + kFinishOffHist);
  uint64_t* qv = reinterpret_cast<uint64_t*>(base + kFinishOffQv);  // (rank << 32) | out
  KT* qpre = reinterpret_cast<KT*>(base + kFinishOffQpre);           // relative prefix
  KT* gpre = reinterpret_cast<KT*>(base + kFinishOffGpre);           // group prefixes
  KT* red_lo = reinterpret_cast<KT*>(base + kFinishOffRed);          // 16
  KT* red_hi = red_lo + 16;                                          // 16
  uint32_t* gfirst = reinterpret_cast<uint32_t*>(base + kFinishOffGfirst);  // kQCap + 1
  uint32_t* misc = reinterpret_cast<uint32_t*>(base + kFinishOffMisc);      // 32 words
  uint32_t* list = reinterpret_cast<uint32_t*>(base + kFinishOffList);      // kThreads + count
  uint32_t& nlist = list[kThreads];
  const uint32_t tid = threadIdx.x;

  // this work-group's children are blockIdx.x, + gridDim.x, ...; they are
  // screened kThreads at a time, one per thread, and only those for this
  // finisher are walked (a table of mostly small children costs one round
  // of loads per window, not one dependent load per child)
  const uint32_t ncand =
      nchildren > blockIdx.x ? (nchildren - blockIdx.x + gridDim.x - 1) / gridDim.x : 0u;
  for (uint32_t w0 = 0; w0 < ncand; w0 += kThreads) {  // uniform
    if (tid == 0) nlist = 0;
    __syncthreads();
    if (w0 + tid < ncand) {
      const uint32_t ci = blockIdx.x + (w0 + tid) * gridDim.x;
      const Child c = children[ci];
      if (c.q_cnt != 0 && !wave_child(c) && !sort_child(c)) list[atomicAdd(&nlist, 1u)] = ci;
    }
    __syncthreads();
    const uint32_t nl = nlist;
    for (uint32_t li = 0; li < nl; ++li) {
      const Child c = children[list[li]];
      const KT* srcp = (c.src == 0 ? input : (c.src == 1 ? red : red2)) + c.start;
      const Query* __restrict__ cq = c.src == 2 ? cq2 : cq0;
      const bool inl = c.len <= CAP;
      KT lo = (KT)c.lo, hi = (KT)c.hi;
      if (inl) {  // the child into LDS, eight loads per thread in flight
        constexpr uint32_t D = 8;
        for (uint32_t i0 = 0; i0 < c.len; i0 += D * kThreads) {
          KT v[D];
#pragma unroll
          for (uint32_t u = 0; u < D; ++u) {
            const uint32_t i = i0 + u * kThreads + tid;
            if (i < c.len) v[u] = srcp[i];
          }
#pragma unroll
          for (uint32_t u = 0; u < D; ++u) {
            const uint32_t i = i0 + u * kThreads + tid;
            if (i < c.len) keys[i] = to_key<C>(v[u]);
          }
        }
      }
      if (c.need_range) {
        KT a = KeyBits<KT>::MAX, b = 0;
        auto mm = [&](KT k) {
          a = k < a ? k : a;
          b = k > b ? k : b;
        };
        if (inl) {
          for (uint32_t i = tid; i < c.len; i += kThreads) mm(keys[i]);
        } else {  // 16-byte loads, four in flight per thread
          for_each_elem(srcp, c.len, [&](KT raw) { mm(to_key<C>(raw)); });
        }
        for (int d = 32; d >= 1; d >>= 1) {
          KT x = __shfl_xor(a, d, 64), y = __shfl_xor(b, d, 64);
          a = x < a ? x : a;
          b = y > b ? y : b;
        }
        if (lane_id() == 0) { red_lo[tid >> 6] = a; red_hi[tid >> 6] = b; }
        __syncthreads();
        lo = red_lo[0]; hi = red_hi[0];
        for (int w = 1; w < 16; ++w) {
          lo = red_lo[w] < lo ? red_lo[w] : lo;
          hi = red_hi[w] > hi ? red_hi[w] : hi;
        }
      }
      __syncthreads();
      if (inl) {
        for (uint32_t i = tid; i < c.len; i += kThreads) keys[i] -= lo;
      }
      const int nbits = bitlen<KT>(hi - lo);

      if (inl && (c.q_cnt > kSortRanks || c.len <= kSortLen)) {
        // many ranks or a short child: sort its keys in LDS, then every rank is a lookup
        const uint32_t np = ((uint32_t)c.len + kThreads - 1) / kThreads * kThreads;
        for (uint32_t i = c.len + tid; i < np; i += kThreads) keys[i] = KeyBits<KT>::MAX;
        __syncthreads();
        lds_radix_sort<KT, CAP / kThreads>(keys, np, nbits, hist, misc);
        for (uint32_t i = tid; i < c.q_cnt; i += kThreads) {
          Query q = cq[c.q_off + i];
          out[q.out] = from_key<C>((KT)(lo + keys[q.rank - 1]));
        }
        __syncthreads();
        continue;
      }

      for (uint32_t q0 = 0; q0 < c.q_cnt; q0 += kQCap) {
        const uint32_t nq = min(kQCap, c.q_cnt - q0);
        // queries sorted by rank: bitonic sort of the padded power of two
        uint32_t np = 1;
        while (np < nq) np <<= 1;
        for (uint32_t i = tid; i < np; i += kThreads) {
          if (i < nq) {
            Query q = cq[c.q_off + q0 + i];
            qv[i] = ((uint64_t)q.rank << 32) | q.out;
          } else {
            qv[i] = ~0ull;
          }
        }
        __syncthreads();
        lds_bitonic_sort<kThreads, (kQCap / 2 + kThreads - 1) / kThreads>(qv, np);
        for (uint32_t i = tid; i < nq; i += kThreads) qpre[i] = 0;
        __syncthreads();

        // digit levels from the top of the relative range
        int rem = nbits;
        while (rem > 0) {
          const int d = rem < 8 ? rem : 8;
          const int shift = rem - d;
          const int hshift = shift + d;  // bits above the digit
          // groups = runs of equal prefix (queries are rank-sorted, so prefixes
          // are non-decreasing)
          if (tid == 0) {
            uint32_t g = 0;
            for (uint32_t i = 0; i < nq; ++i)
              if (i == 0 || qpre[i] != qpre[i - 1]) gfirst[g++] = i;
            gfirst[g] = nq;
            misc[0] = g;
          }
          __syncthreads();
          const uint32_t ng = misc[0];
          for (uint32_t g0 = 0; g0 < ng; g0 += kGroupMax) {
            const uint32_t gn = min(kGroupMax, ng - g0);
            for (uint32_t i = tid; i < gn * 256; i += kThreads) hist[i] = 0;
            for (uint32_t g = tid; g < gn; g += kThreads) gpre[g] = qpre[gfirst[g0 + g]];
            __syncthreads();
            auto visit = [&](KT rel) {
              KT pre = hshift >= (int)W ? KT(0) : (rel >> hshift);
              uint32_t g;
              if (gn == 1) {
                if (pre != gpre[0]) return;
                g = 0;
              } else {
                uint32_t a = 0, b = gn - 1;
                while (a < b) {
                  uint32_t mid = (a + b) >> 1;
                  if (gpre[mid] < pre) a = mid + 1; else b = mid;
                }
                if (gpre[a] != pre) return;
                g = a;
              }
              atomicAdd(&hist[g * 256 + (uint32_t)((rel >> shift) & (KT)((1u << d) - 1))], 1u);
            };
            if (inl) {
              for (uint32_t i = tid; i < c.len; i += kThreads) visit(keys[i]);
            } else {  // 16-byte loads, four in flight per thread
              for_each_elem(srcp, c.len, [&](KT raw) { visit((KT)(to_key<C>(raw) - lo)); });
            }
            __syncthreads();
            // exclusive scan of each group's 256 bins: one wave per group
            const uint32_t wv = tid >> 6, lane = lane_id();
            for (uint32_t g = wv; g < gn; g += kThreads / 64) {
              uint32_t* h = hist + g * 256;
              uint32_t v0 = h[lane * 4], v1 = h[lane * 4 + 1], v2 = h[lane * 4 + 2], v3 = h[lane * 4 + 3];
              uint32_t s = v0 + v1 + v2 + v3;
              uint32_t e = wave_incl_scan(s) - s;
              h[lane * 4] = e;
              h[lane * 4 + 1] = e + v0;
              h[lane * 4 + 2] = e + v0 + v1;
              h[lane * 4 + 3] = e + v0 + v1 + v2;
            }
            __syncthreads();
            // every query of these groups picks its digit
            const uint32_t qa = gfirst[g0], qz = gfirst[g0 + gn];
            for (uint32_t i = qa + tid; i < qz; i += kThreads) {
              uint32_t a = 0, b = gn - 1;  // group of query i
              while (a < b) {
                uint32_t mid = (a + b + 1) >> 1;
                if (gfirst[g0 + mid] <= i) a = mid; else b = mid - 1;
              }
              const uint32_t* h = hist + a * 256;
              uint32_t r1 = (uint32_t)(qv[i] >> 32) - 1;
              uint32_t lo_d = 0, hi_d = (1u << d) - 1;  // largest digit with h <= r1
              while (lo_d < hi_d) {
                uint32_t mid = (lo_d + hi_d + 1) >> 1;
                if (h[mid] <= r1) lo_d = mid; else hi_d = mid - 1;
              }
              uint32_t nr = r1 - h[lo_d] + 1;
              qv[i] = ((uint64_t)nr << 32) | (uint32_t)qv[i];
              qpre[i] = (qpre[i] << d) | (KT)lo_d;
            }
            __syncthreads();
          }
          rem = shift;
        }
        for (uint32_t i = tid; i < nq; i += kThreads)
          out[(uint32_t)qv[i]] = from_key<C>((KT)(lo + qpre[i]));
        __syncthreads();
      }
    }
    __syncthreads();  // every thread is done with the list before it is refilled
  }
}

}  // namespace ggks
