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
//   k_finish  per child: multi-rank MSD radix select on 8-bit digits of
//             (key - lo), keys resident in LDS when they fit
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

#include "ggks_keys.hpp"
#include "ggks_types.hpp"

namespace ggks {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
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

// In-place exclusive scan of a[0, L) in LDS by one 1024-thread block.
__device__ uint32_t block_scan_array(uint32_t* a, uint32_t L, uint32_t* wsum) {
  const uint32_t per = (L + kThreads - 1) / kThreads;
  const uint32_t b0 = threadIdx.x * per;
  const uint32_t b1 = min(L, b0 + per);
  uint32_t local = 0;
  for (uint32_t i = b0; i < b1; ++i) local += a[i];
  uint32_t total;
  uint32_t run = block_excl_scan(local, wsum, total);
  for (uint32_t i = b0; i < b1; ++i) {
    uint32_t v = a[i];
    a[i] = run;
    run += v;
  }
  __syncthreads();
  return total;
}

// f(i, p[i]) for i = threadIdx.x, +kThreads, ... < n with 16 loads in flight
// per thread (the loops that read another kernel's output are latency bound).
template <typename T, typename F>
__device__ __forceinline__ void strided_for(const T* __restrict__ p, uint32_t n, F&& f) {
  constexpr uint32_t D = 16;
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

// The bucket function of a segment.  bucket 0 holds keys below kb, bucket B-1
// keys above the last prefix; interior bucket = 1 + base[p] + j with
// p = (key - kb) >> s and j = floor(sub * nb[p] / 2^s), sub the low s bits.
// j is computed as mulhi(sub << (W - s), nb) which is exact.
template <typename KT>
struct Bucketer {
  KT kb, smask;
  uint32_t s1, s2, P, Bm1, lsh;
  __device__ __forceinline__ void init(const SegState& st) {
    constexpr uint32_t W = KeyBits<KT>::W;
    kb = (KT)st.kb;
    s1 = st.s >> 1;
    s2 = st.s - s1;
    P = st.P;
    Bm1 = st.B - 1;
    smask = st.s >= W ? KeyBits<KT>::MAX : ((KT(1) << st.s) - KT(1));
    lsh = st.s == 0 ? 0u : (W - st.s);
  }
  // Branch-free: the two tail cases are selected after the look-up, so a
  // wavefront issues all its LDS reads without exec-mask changes.
  __device__ __forceinline__ uint32_t operator()(KT key, const uint32_t* lut) const {
    const bool below = key < kb;
    const KT low = key - kb;
    const KT p = (low >> s1) >> s2;
    const bool above = p >= (KT)P;
    const uint32_t pi = (below | above) ? 0u : (uint32_t)p;
    const uint32_t e = lut[pi];
    const uint32_t j = mulhi_nb<KT>((low & smask) << lsh, e >> 16);
    const uint32_t b = 1u + (e & 0xFFFFu) + j;
    return below ? 0u : (above ? Bm1 : b);
  }
};

// ceil(j * 2^s / nb) for j < nb < 2^16, s <= 64, without overflow.
__device__ __forceinline__ uint64_t ceil_j_pow2_div(uint32_t j, uint32_t s, uint32_t nb) {
  if (j == 0 || nb <= 1) return 0;
  if (s <= 47) {
    uint64_t num = ((uint64_t)j << s) + nb - 1;
    return num / nb;
  }
  // j*2^s = (j*2^16) * 2^(s-16); split j*2^16 = q*nb + r.
  uint64_t A = (uint64_t)j << 16;
  uint64_t q = A / nb, r = A % nb;
  uint32_t t = s - 16;  // <= 48
  return (q << t) + (((r << t) + nb - 1) / nb);
}

// Key range [lo, hi] of bucket b of a segment (b must hold elements).
// Interior buckets: lo is the smallest key mapped to b.  Because the bucket
// function is monotone, hi is the smallest key of the next bucket minus one.
template <typename KT>
struct RangeOf {
  const SegState* st;
  const uint32_t* lut;
  __device__ unsigned __int128 first_key(uint32_t bi) const {
    // smallest prefix p with base[p] + max(nb[p], 1) > bi
    uint32_t lo = 0, hi = st->P - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      uint32_t e = lut[mid];
      uint32_t nb = e >> 16;
      uint32_t end = (e & 0xFFFFu) + (nb > 1 ? nb : 1u);
      if (end > bi) hi = mid; else lo = mid + 1;
    }
    uint32_t e = lut[lo];
    uint32_t nb = e >> 16;
    uint32_t j = bi - (e & 0xFFFFu);
    return (unsigned __int128)st->kb + ((unsigned __int128)lo << st->s) +
           ceil_j_pow2_div(j, st->s, nb);
  }
  __device__ void operator()(uint32_t b, KT& lo, KT& hi) const {
    constexpr unsigned __int128 KMAX = (unsigned __int128)KeyBits<KT>::MAX;
    unsigned __int128 end = (unsigned __int128)st->kb + ((unsigned __int128)st->P << st->s);
    if (b == 0) {
      lo = 0;
      hi = (KT)(st->kb - 1);
      return;
    }
    if (b == st->B - 1) {
      lo = (KT)(end > KMAX ? KMAX : end);
      hi = KeyBits<KT>::MAX;
      return;
    }
    uint32_t bi = b - 1;
    unsigned __int128 l = first_key(bi);
    unsigned __int128 h = (bi + 1 < st->Bi) ? first_key(bi + 1) - 1 : end - 1;
    if (h > KMAX) h = KMAX;
    lo = (KT)l;
    hi = (KT)h;
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

// Seeded random sample of each segment: grid (segments, sample blocks).
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_sample(const KT* __restrict__ src,
                                                     const SegDesc* __restrict__ segs,
                                                     KT* __restrict__ samples, uint64_t seed) {
  const SegDesc& sd = segs[blockIdx.x];
  if (sd.mode != MODE_SAMPLE) return;
  uint32_t i = blockIdx.y * kThreads + threadIdx.x;
  if (i >= sd.nsample) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)blockIdx.x * 0x100000001ull + i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  uint64_t pos = __umul64hi(z, sd.len);
  samples[sd.sample_off + i] = to_key<C>(src[sd.data_off + pos]);
}

// Build one segment's bucket function (one block per segment).
template <typename KT>
__global__ __launch_bounds__(kThreads) void k_lut(const SegDesc* __restrict__ segs,
                                                  const KT* __restrict__ samples,
                                                  SegState* __restrict__ states,
                                                  uint32_t* __restrict__ lut_pool) {
  __shared__ uint32_t cnt[1u << kMaxPrefixBits];
  __shared__ uint32_t wsum[17];
  __shared__ KT red_lo[16], red_hi[16];
  constexpr uint32_t W = KeyBits<KT>::W;
  const SegDesc sd = segs[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  uint32_t* lut = lut_pool + sd.lut_off;
  SegState st = {};
  KT kmin, kmax;
  uint32_t pb = sd.pbits;

  if (sd.mode == MODE_SAMPLE) {
    KT lo = KeyBits<KT>::MAX, hi = 0;
    strided_for(samples + sd.sample_off, sd.nsample, [&](uint32_t, KT k) {
      lo = k < lo ? k : lo;
      hi = k > hi ? k : hi;
    });
    for (int d = 32; d >= 1; d >>= 1) {
      KT a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
    }
    if (lane_id() == 0) { red_lo[tid >> 6] = lo; red_hi[tid >> 6] = hi; }
    __syncthreads();
    kmin = red_lo[0]; kmax = red_hi[0];
    for (int w = 1; w < 16; ++w) {
      kmin = red_lo[w] < kmin ? red_lo[w] : kmin;
      kmax = red_hi[w] > kmax ? red_hi[w] : kmax;
    }
  } else if (sd.mode == MODE_RADIX) {
    kmin = 0;
    kmax = KeyBits<KT>::MAX;
  } else {
    kmin = (KT)sd.lo;
    kmax = (KT)sd.hi;
  }
  const KT R = kmax - kmin;
  const int bl = bitlen<KT>(R);
  uint32_t s, P;
  if (sd.mode == MODE_RADIX) {
    uint32_t rb = sd.radix_bits;
    pb = rb < pb ? rb : pb;
    s = W - pb;
    P = 1u << pb;
  } else if (sd.mode == MODE_RANGE) {
    pb = 0;
    s = (uint32_t)bl;
    P = 1;
  } else {
    s = bl > (int)pb ? (uint32_t)(bl - (int)pb) : 0u;
    P = (uint32_t)(((uint64_t)R >> s) + 1);  // s >= 1 whenever R >= 2^pb
  }
  const uint32_t Bt = sd.bmax - 2;  // interior budget
  const uint32_t cap = s >= 16 ? 65535u : min(65535u, 1u << s);

  if (sd.mode == MODE_RADIX) {
    // every prefix gets 2^(radix_bits - pb) buckets: the next bits exactly
    uint32_t nb = 1u << (sd.radix_bits - pb);
    for (uint32_t p = tid; p < P; p += kThreads) lut[p] = (p * nb) | (nb << 16);
    st.Bi = P * nb;
  } else if (sd.mode == MODE_RANGE) {
    uint32_t nb = min(Bt, cap);
    if (tid == 0) lut[0] = nb << 16;
    st.Bi = nb;
  } else {
    for (uint32_t p = tid; p < P; p += kThreads) cnt[p] = 0;
    __syncthreads();
    strided_for(samples + sd.sample_off, sd.nsample, [&](uint32_t, KT k) {
      atomicAdd(&cnt[(uint32_t)((k - kmin) >> s)], 1u);
    });
    __syncthreads();
    // prefixes holding no sample: a run of them shares one bucket
    uint32_t nonempty = 0, runs = 0;
    for (uint32_t p = tid; p < P; p += kThreads) {
      if (cnt[p]) nonempty++;
      else if (p == 0 || cnt[p - 1]) runs++;
    }
    uint32_t t1, t2;
    block_excl_scan(nonempty, wsum, t1);
    block_excl_scan(runs, wsum, t2);
    const uint32_t A = Bt > t1 + t2 ? Bt - t1 - t2 : 0u;
    const uint32_t ns = sd.nsample;
    // width of each prefix in buckets, then an exclusive scan gives the base
    const uint32_t per = (P + kThreads - 1) / kThreads;
    const uint32_t p0 = tid * per, p1 = min(P, p0 + per);
    uint32_t local = 0;
    for (uint32_t p = p0; p < p1; ++p) {
      uint32_t c = cnt[p];
      uint32_t w;
      if (c) w = min(cap, 1u + (uint32_t)(((uint64_t)c * A) / ns));
      else w = (p == 0 || cnt[p - 1]) ? 1u : 0u;
      local += w;
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(local, wsum, tot);
    for (uint32_t p = p0; p < p1; ++p) {
      uint32_t c = cnt[p];
      uint32_t nb, w, base;
      if (c) { nb = min(cap, 1u + (uint32_t)(((uint64_t)c * A) / ns)); w = nb; base = run; }
      else if (p == 0 || cnt[p - 1]) { nb = 0; w = 1; base = run; }
      else { nb = 0; w = 0; base = run - 1; }
      lut[p] = base | (nb << 16);
      run += w;
    }
    st.Bi = tot;
  }
  if (tid == 0) {
    st.kb = (uint64_t)kmin;
    st.kmax = (uint64_t)kmax;
    st.s = s;
    st.P = P;
    st.B = st.Bi + 2;
    st.nact = 0;
    states[blockIdx.x] = st;
  }
}

// Pass 1: per-chunk LDS histogram of bucket ids.
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
  uint32_t* hist = lds;
  uint32_t* lut = lds + bmax;
  for (uint32_t i = threadIdx.x; i < st.B; i += kThreads) hist[i] = 0;
  for (uint32_t i = threadIdx.x; i < st.P; i += kThreads) lut[i] = lut_pool[sd.lut_off + i];
  __syncthreads();
  Bucketer<KT> bk;
  bk.init(st);
  for_each_elem<KT>(src + sd.data_off + ch.start, ch.len, [&](KT raw) {
    atomicAdd(&hist[bk(to_key<C>(raw), lut)], 1u);
  });
  __syncthreads();
  uint32_t* row = rows + sd.row_base + (uint64_t)ch.idx * bmax;
  for (uint32_t i = threadIdx.x; i < st.B; i += kThreads) row[i] = hist[i];
}

// Sum the chunk rows of kReduceBins buckets; rewrite each row entry as the
// chunk's exclusive offset inside its bucket.  1024 threads = 16 groups of
// chunks x 64 bins; each thread keeps its (at most kReduceRegs) row entries in
// registers between the sum and the rewrite.
constexpr uint32_t kReduceParts = kThreads / kReduceBins;  // 16
constexpr uint32_t kReduceRegs = kMaxChunks / kReduceParts;  // chunks per part

__global__ __launch_bounds__(kThreads) void k_reduce(const SegDesc* __restrict__ segs,
                                                     const SegState* __restrict__ states,
                                                     const ReduceItem* __restrict__ items,
                                                     uint32_t* __restrict__ rows,
                                                     uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[kReduceParts][kReduceBins];
  const ReduceItem it = items[blockIdx.x];
  const SegDesc& sd = segs[it.seg];
  const uint32_t B = states[it.seg].B;
  const uint32_t lane = threadIdx.x % kReduceBins, q = threadIdx.x / kReduceBins;
  const uint32_t bin = it.group * kReduceBins + lane;
  const bool valid = bin < B;
  const uint32_t nch = sd.nchunks;  // <= kMaxChunks (planner)
  const uint32_t c0 = (uint32_t)((uint64_t)nch * q / kReduceParts);
  const uint32_t c1 = (uint32_t)((uint64_t)nch * (q + 1) / kReduceParts);
  uint32_t* base = rows + sd.row_base + bin;
  const uint64_t stride = sd.bmax;
  uint32_t v[kReduceRegs];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t u = 0; u < kReduceRegs; ++u) {
    v[u] = (valid && c0 + u < c1) ? base[(uint64_t)(c0 + u) * stride] : 0u;
    sum += v[u];
  }
  part[q][lane] = sum;
  __syncthreads();
  uint32_t off = 0, total = 0;
#pragma unroll
  for (uint32_t j = 0; j < kReduceParts; ++j) {
    uint32_t w = part[j][lane];
    off += j < q ? w : 0u;
    total += w;
  }
  uint32_t run = off;
#pragma unroll
  for (uint32_t u = 0; u < kReduceRegs; ++u) {
    if (valid && c0 + u < c1) base[(uint64_t)(c0 + u) * stride] = run;
    run += v[u];
  }
  if (valid && q == 0) totals[sd.tot_off + bin] = total;
}

// Per segment: ranks -> buckets, children, and the ranks grouped by child.
// The i-th query of a segment: explicit, from a raw k list (k-th largest ->
// rank len - k + 1, output slot i; a bad k raises a status bit and is clamped
// to k = 1), or rank len - i for a top-K.
__device__ __forceinline__ Query seg_query(const SegDesc& sd, const Query* __restrict__ queries,
                                           const uint32_t* __restrict__ kraw, uint32_t i,
                                           uint32_t* __restrict__ status) {
  Query q;
  if (sd.qsrc == QSRC_QUERIES) return queries[sd.q_off + i];
  if (sd.qsrc == QSRC_K) {
    uint32_t k = kraw[sd.q_off + i];
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

template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_findk(const SegDesc* __restrict__ segs,
                                                    SegState* __restrict__ states,
                                                    const uint32_t* __restrict__ lut_pool,
                                                    const uint32_t* __restrict__ totals,
                                                    const Query* __restrict__ queries,
                                                    const uint32_t* __restrict__ kraw,
                                                    uint32_t* __restrict__ status,
                                                    uint32_t* __restrict__ qb,
                                                    uint32_t* __restrict__ qr,
                                                    Child* __restrict__ children,
                                                    Query* __restrict__ cq,
                                                    uint32_t* __restrict__ bits_pool,
                                                    uint32_t* __restrict__ wbase_pool,
                                                    KT* __restrict__ out) {
  extern __shared__ uint32_t lds[];
  const SegDesc sd = segs[blockIdx.x];
  __shared__ SegState st;
  if (threadIdx.x == 0) st = states[blockIdx.x];
  __syncthreads();
  const uint32_t B = st.B;
  const uint32_t nwords = (B + 31) / 32;
  uint32_t* hist = lds;                          // bmax
  uint32_t* lut = hist + sd.bmax;                // 2^pbits
  uint32_t* bits = lut + (1u << sd.pbits);       // bmax / 32
  uint32_t* wb = bits + sd.bmax / 32;            // bmax / 32
  uint32_t* wsum = wb + sd.bmax / 32;            // 17 (+ pad)
  const uint32_t tid = threadIdx.x;
  strided_for(totals + sd.tot_off, B, [&](uint32_t i, uint32_t v) { hist[i] = v; });
  strided_for(lut_pool + sd.lut_off, st.P, [&](uint32_t i, uint32_t v) { lut[i] = v; });
  for (uint32_t i = tid; i < nwords; i += kThreads) bits[i] = 0;
  __syncthreads();
  const uint32_t len = block_scan_array(hist, B, wsum);  // hist = exclusive
  RangeOf<KT> range{&st, lut};
  auto count_of = [&](uint32_t b) { return (b + 1 < B ? hist[b + 1] : len) - hist[b]; };

  // A: bucket of every rank; single-key buckets are answered here
  const uint32_t NONE = 0xFFFFFFFFu;
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    const Query q = seg_query(sd, queries, kraw, i, status);
    const uint32_t r1 = q.rank - 1;
    uint32_t lo = 0, hi = B - 1;  // largest b with hist[b] <= r1
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (hist[mid] <= r1) lo = mid; else hi = mid - 1;
    }
    KT klo, khi;
    range(lo, klo, khi);
    if (klo == khi) {
      out[q.out] = from_key<C>(klo);
      qb[sd.q_off + i] = NONE;
    } else {
      atomicOr(&bits[lo >> 5], 1u << (lo & 31));
      qb[sd.q_off + i] = lo;
      qr[sd.q_off + i] = q.rank - hist[lo];
    }
  }
  __syncthreads();

  // B: active-bucket index per bitmap word, child layout in bucket order
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
    bits_pool[sd.bits_off + tid] = m;
    wbase_pool[sd.bits_off + tid] = ubase;
    uint32_t u = ubase;
    uint64_t start = sd.red_base + cbase;
    for (uint32_t mm = m; mm; mm &= mm - 1) {
      uint32_t b = tid * 32 + __ffs(mm) - 1;
      Child c;
      c.start = start;
      c.len = count_of(b);
      KT klo, khi;
      range(b, klo, khi);
      c.lo = klo;
      c.hi = khi;
      c.q_off = 0;
      c.q_cnt = 0;
      c.src = 1;
      c.need_range = 0;
      c.bucket = b;
      c.seg = blockIdx.x;
      children[sd.q_off + u] = c;
      start += c.len;
      ++u;
    }
  }
  __syncthreads();

  // C: group ranks by child (hist is free now: counters, then offsets)
  for (uint32_t u = tid; u < U; u += kThreads) hist[u] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    uint32_t b = qb[sd.q_off + i];
    if (b == NONE) continue;
    uint32_t w = b >> 5, mw = bits[w];
    uint32_t u = wb[w] + __popc(mw & ((1u << (b & 31)) - 1u));
    atomicAdd(&hist[u], 1u);
    qb[sd.q_off + i] = u;
  }
  __syncthreads();
  const uint32_t nq = block_scan_array(hist, U, wsum);  // hist[u] = first slot
  for (uint32_t u = tid; u < U; u += kThreads) {
    uint32_t first = hist[u];
    uint32_t next = (u + 1 < U) ? hist[u + 1] : nq;
    children[sd.q_off + u].q_off = sd.q_off + first;
    children[sd.q_off + u].q_cnt = next - first;
  }
  __syncthreads();
  for (uint32_t i = tid; i < sd.q_cnt; i += kThreads) {
    uint32_t u = qb[sd.q_off + i];
    if (u == NONE) continue;
    uint32_t pos = atomicAdd(&hist[u], 1u);
    Query q;
    q.rank = qr[sd.q_off + i];
    q.out = seg_query(sd, queries, kraw, i, status).out;
    cq[sd.q_off + pos] = q;
  }
  // unused child slots finish nothing
  for (uint32_t u = U + tid; u < sd.q_cnt; u += kThreads) {
    children[sd.q_off + u].q_cnt = 0;
    children[sd.q_off + u].len = 0;
  }
  if (tid == 0) states[blockIdx.x].nact = U;
}

// Pass 2: copy the keys of active buckets into their children's ranges.
//
// Every wavefront stages its active keys in LDS (slot = ballot prefix count,
// no atomics) and flushes them in batches: one LDS cursor atomic per staged
// key, 64 lanes at a time, then the global store.  The body runs in uniform
// rounds so every lane of a wavefront takes part in every ballot.
constexpr uint32_t kStageCap = 512;  // staged keys per wavefront (>= 2 groups of 256)
constexpr uint32_t kWaves = kThreads / 64;

template <typename KT>
__host__ __device__ constexpr size_t compact_lds_bytes(uint32_t pbits, uint32_t bmax) {
  return (size_t)(1u << pbits) * 4 + (size_t)((bmax + 31) / 32) * 8 + (size_t)kCursorCap * 4 +
         (size_t)kWaves * kStageCap * (sizeof(KT) + 2);
}

template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_compact(const KT* __restrict__ src,
                                                      const SegDesc* __restrict__ segs,
                                                      const SegState* __restrict__ states,
                                                      const Chunk* __restrict__ chunks,
                                                      const uint32_t* __restrict__ lut_pool,
                                                      uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ bits_pool,
                                                      const uint32_t* __restrict__ wbase_pool,
                                                      const Child* __restrict__ children,
                                                      KT* __restrict__ red) {
  extern __shared__ __align__(16) uint32_t lds[];
  const Chunk ch = chunks[blockIdx.x];
  const SegDesc& sd = segs[ch.seg];
  const SegState st = states[ch.seg];
  const uint32_t U = st.nact;
  if (U == 0) return;
  const uint32_t nwords = (st.B + 31) / 32;
  char* base = reinterpret_cast<char*>(lds);
  uint32_t* lut = reinterpret_cast<uint32_t*>(base);
  uint2* bw = reinterpret_cast<uint2*>(base + (size_t)(1u << sd.pbits) * 4);  // (bits, wbase)
  uint32_t* cur = reinterpret_cast<uint32_t*>(bw + (sd.bmax + 31) / 32);
  KT* sval = reinterpret_cast<KT*>(cur + kCursorCap);
  uint16_t* su = reinterpret_cast<uint16_t*>(sval + kWaves * kStageCap);
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = lane_id();
  for (uint32_t i = tid; i < st.P; i += kThreads) lut[i] = lut_pool[sd.lut_off + i];
  for (uint32_t i = tid; i < nwords; i += kThreads)
    bw[i] = make_uint2(bits_pool[sd.bits_off + i], wbase_pool[sd.bits_off + i]);
  uint32_t* row = rows + sd.row_base + (uint64_t)ch.idx * sd.bmax;
  const bool lds_cur = U <= kCursorCap;
  if (lds_cur) {
    for (uint32_t u = tid; u < U; u += kThreads) {
      const Child& c = children[sd.q_off + u];
      cur[u] = (uint32_t)c.start + row[c.bucket];
    }
  }
  __syncthreads();
  Bucketer<KT> bk;
  bk.init(st);
  const Child* kids = children + sd.q_off;
  // child index of key `raw` or ~0u when its bucket is not active
  auto child_of = [&](KT raw, uint32_t& b) -> uint32_t {
    b = bk(to_key<C>(raw), lut);
    const uint2 e = bw[b >> 5];
    const uint32_t sh = b & 31u;
    return ((e.x >> sh) & 1u) ? e.y + __popc(e.x & ((1u << sh) - 1u)) : ~0u;
  };
  auto place = [&](uint32_t u, uint32_t b, KT raw) {
    uint32_t pos = lds_cur ? atomicAdd(&cur[u], 1u) : (uint32_t)kids[u].start + atomicAdd(&row[b], 1u);
    red[pos] = raw;
  };
  const KT* p = src + sd.data_off + ch.start;
  const uint64_t len = ch.len;
  constexpr uint32_t VEC = 16 / sizeof(KT);
  uint64_t mis = ((uintptr_t)p & 15u) / sizeof(KT);
  uint64_t head = mis ? (VEC - mis) : 0;
  if (head > len) head = len;
  const uint64_t nvec = (len - head) / VEC;
  const uint64_t done = head + nvec * VEC;
  // head and tail: fewer than VEC keys each, placed directly (a thread may
  // own one of each)
  if (tid < head) {
    const KT raw = p[tid];
    uint32_t b, u = child_of(raw, b);
    if (u != ~0u) place(u, b, raw);
  }
  if (done + tid < len) {
    const KT raw = p[done + tid];
    uint32_t b, u = child_of(raw, b);
    if (u != ~0u) place(u, b, raw);
  }
  const uint4* __restrict__ pv = reinterpret_cast<const uint4*>(p + head);
  KT* wv = sval + wave * kStageCap;
  uint16_t* wu = su + wave * kStageCap;
  uint32_t cnt = 0;  // staged keys of this wavefront (uniform)
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t i = lane; i < cnt; i += 64) {
      uint32_t u = wu[i];
      KT raw = wv[i];
      uint32_t b = 0;
      if (!lds_cur) b = kids[u].bucket;
      place(u, b, raw);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    cnt = 0;
  };
  auto stage = [&](KT raw, bool valid) {
    uint32_t b, u = child_of(raw, b);
    const bool act = valid && u != ~0u;
    const uint64_t m = __ballot(act);
    if (act) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      wv[cnt + below] = raw;
      wu[cnt + below] = (uint16_t)u;
    }
    cnt += (uint32_t)__popcll(m);
  };
  const uint64_t per_round = 4ull * kThreads;
  const uint64_t rounds = (nvec + per_round - 1) / per_round;
  for (uint64_t r = 0; r < rounds; ++r) {
    uint4 v[4];
    bool ok[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      uint64_t idx = r * per_round + (uint64_t)t * kThreads + tid;
      ok[t] = idx < nvec;
      v[t] = ok[t] ? pv[idx] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      // a 16-byte group stages at most VEC * 64 keys per wavefront
      if (cnt > kStageCap - VEC * 64) flush();
      if (sizeof(KT) == 4) {
        stage((KT)v[t].x, ok[t]); stage((KT)v[t].y, ok[t]);
        stage((KT)v[t].z, ok[t]); stage((KT)v[t].w, ok[t]);
      } else {
        stage((KT)(((uint64_t)v[t].y << 32) | v[t].x), ok[t]);
        stage((KT)(((uint64_t)v[t].w << 32) | v[t].z), ok[t]);
      }
    }
  }
  if (cnt) flush();
}

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
constexpr uint32_t kFinishLdsTotal = kFinishOffMisc + 32 * 4;
template <typename KT, int C>
__global__ __launch_bounds__(kThreads) void k_finish(const KT* __restrict__ input,
                                                     const KT* __restrict__ red,
                                                     const Child* __restrict__ children,
                                                     uint32_t nchildren,
                                                     const Query* __restrict__ cq,
                                                     KT* __restrict__ out) {
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
  const uint32_t tid = threadIdx.x;

  for (uint32_t ci = blockIdx.x; ci < nchildren; ci += gridDim.x) {
    const Child c = children[ci];
    if (c.q_cnt == 0) continue;
    const KT* srcp = (c.src ? red : input) + c.start;
    const bool inl = c.len <= CAP;
    KT lo = (KT)c.lo, hi = (KT)c.hi;
    if (inl) {
      for (uint32_t i = tid; i < c.len; i += kThreads) keys[i] = to_key<C>(srcp[i]);
    }
    if (c.need_range) {
      KT a = KeyBits<KT>::MAX, b = 0;
      for (uint32_t i = tid; i < c.len; i += kThreads) {
        KT k = inl ? keys[i] : to_key<C>(srcp[i]);
        a = k < a ? k : a;
        b = k > b ? k : b;
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

    if (inl && c.q_cnt > kSortRanks) {
      // many ranks: sort the child's keys in LDS, then every rank is a lookup
      uint32_t np = 1;
      while (np < c.len) np <<= 1;
      for (uint32_t i = c.len + tid; i < np; i += kThreads) keys[i] = KeyBits<KT>::MAX;
      __syncthreads();
      for (uint32_t kk = 2; kk <= np; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
          for (uint32_t i = tid; i < np; i += kThreads) {
            uint32_t ix = i ^ j;
            if (ix > i) {
              KT a = keys[i], b = keys[ix];
              bool up = (i & kk) == 0;
              if ((a > b) == up) { keys[i] = b; keys[ix] = a; }
            }
          }
          __syncthreads();
        }
      }
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
      for (uint32_t kk = 2; kk <= np; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
          for (uint32_t i = tid; i < np; i += kThreads) {
            uint32_t ix = i ^ j;
            if (ix > i) {
              uint64_t a = qv[i], b = qv[ix];
              bool up = (i & kk) == 0;
              if ((a > b) == up) { qv[i] = b; qv[ix] = a; }
            }
          }
          __syncthreads();
        }
      }
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
          } else {
            for (uint64_t i = tid; i < c.len; i += kThreads) visit(to_key<C>(srcp[i]) - lo);
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
}

}  // namespace ggks
