/*
 * oracle.c -- CPU restatement of GGKS's selection path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  Nothing in the product library calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) lives in
 * tests/test_oracle.py; every function here is pinned there.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* key conversion                                                       */
/* ------------------------------------------------------------------ */

/* keyConversion.cu:10-16 / radixsort_key_conversion.h:117-122 */
uint32_t ggko_key_f32(uint32_t u) {
  uint32_t mask = (u & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
  return u ^ mask;
}
/* keyConversion.cu:29-33 / radixsort_key_conversion.h:126-131 */
uint32_t ggko_unkey_f32(uint32_t k) {
  uint32_t mask = (k & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu;
  return k ^ mask;
}
/* keyConversion.cu:18-24 / radixsort_key_conversion.h:147-152 */
uint64_t ggko_key_f64(uint64_t u) {
  uint64_t mask = (u & 0x8000000000000000ull) ? 0xFFFFFFFFFFFFFFFFull
                                              : 0x8000000000000000ull;
  return u ^ mask;
}
/* keyConversion.cu:37-42 / radixsort_key_conversion.h:156-160 */
uint64_t ggko_unkey_f64(uint64_t k) {
  uint64_t mask = (k & 0x8000000000000000ull) ? 0x8000000000000000ull
                                              : 0xFFFFFFFFFFFFFFFFull;
  return k ^ mask;
}
/* radixsort_key_conversion.h:252-256 (signed int: flip the sign bit) */
uint32_t ggko_key_i32(uint32_t u) { return u ^ 0x80000000u; }
uint32_t ggko_unkey_i32(uint32_t k) { return k ^ 0x80000000u; }

static int wide(int dtype) { return dtype == GGKO_F64; }
static size_t esize(int dtype) { return wide(dtype) ? 8 : 4; }
static int valid_dtype(int dtype) { return dtype >= 0 && dtype <= 3; }

void ggko_keys(int dtype, const void *in, void *out, int64_t n) {
  if (wide(dtype)) {
    const uint64_t *a = (const uint64_t *)in;
    uint64_t *b = (uint64_t *)out;
    for (int64_t i = 0; i < n; ++i) b[i] = ggko_key_f64(a[i]);
    return;
  }
  const uint32_t *a = (const uint32_t *)in;
  uint32_t *b = (uint32_t *)out;
  for (int64_t i = 0; i < n; ++i)
    b[i] = dtype == GGKO_F32 ? ggko_key_f32(a[i])
         : dtype == GGKO_I32 ? ggko_key_i32(a[i]) : a[i];
}

void ggko_unkeys(int dtype, const void *in, void *out, int64_t n) {
  if (wide(dtype)) {
    const uint64_t *a = (const uint64_t *)in;
    uint64_t *b = (uint64_t *)out;
    for (int64_t i = 0; i < n; ++i) b[i] = ggko_unkey_f64(a[i]);
    return;
  }
  const uint32_t *a = (const uint32_t *)in;
  uint32_t *b = (uint32_t *)out;
  for (int64_t i = 0; i < n; ++i)
    b[i] = dtype == GGKO_F32 ? ggko_unkey_f32(a[i])
         : dtype == GGKO_I32 ? ggko_unkey_i32(a[i]) : a[i];
}

/* ------------------------------------------------------------------ */
/* LSD radix sort on converted keys (8-bit digits).  This is the order   */
/* the reference's thrust::sort gives for primitives: b40c radix sort on */
/* PreprocessKeyFunctor-converted bits (sort.inl:259-275).               */
/* ------------------------------------------------------------------ */

static void radix_sort_u32(uint32_t *a, int64_t n) {
  uint32_t *tmp = (uint32_t *)malloc((size_t)n * 4);
  uint32_t *src = a, *dst = tmp;
  for (int shift = 0; shift < 32; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof cnt);
    for (int64_t i = 0; i < n; ++i) cnt[((src[i] >> shift) & 255u) + 1]++;
    if (cnt[((src[0] >> shift) & 255u) + 1] == n) continue; /* one digit */
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 255u]++] = src[i];
    uint32_t *t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, (size_t)n * 4);
  free(tmp);
}

static void radix_sort_u64(uint64_t *a, int64_t n) {
  uint64_t *tmp = (uint64_t *)malloc((size_t)n * 8);
  uint64_t *src = a, *dst = tmp;
  for (int shift = 0; shift < 64; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof cnt);
    for (int64_t i = 0; i < n; ++i) cnt[((src[i] >> shift) & 255u) + 1]++;
    if (cnt[((src[0] >> shift) & 255u) + 1] == n) continue;
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 255u]++] = src[i];
    uint64_t *t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, (size_t)n * 8);
  free(tmp);
}

/* Sorted converted keys of vec (caller frees). */
static void *sorted_keys(int dtype, const void *vec, int64_t n) {
  void *keys = malloc((size_t)n * esize(dtype));
  if (!keys) return NULL;
  ggko_keys(dtype, vec, keys, n);
  if (wide(dtype)) radix_sort_u64((uint64_t *)keys, n);
  else radix_sort_u32((uint32_t *)keys, n);
  return keys;
}

int ggko_sort(int dtype, const void *vec, int64_t n, void *sorted_out) {
  if (!valid_dtype(dtype) || n <= 0 || !vec || !sorted_out) return -1;
  void *keys = sorted_keys(dtype, vec, n);
  if (!keys) return -1;
  ggko_unkeys(dtype, keys, sorted_out, n);
  free(keys);
  return 0;
}

static int check_ks(int64_t n, const uint32_t *k, int64_t nk) {
  if (n <= 0 || nk < 0 || (nk > 0 && !k)) return -1;
  for (int64_t i = 0; i < nk; ++i)
    if (k[i] < 1 || (int64_t)k[i] > n) return -1;
  return 0;
}

/* multiselectTimingFunctions.cu:41-48 (copyInChunk) and :50-95:
   sort ascending, then out[i] = v[numElements - kList[i]]. */
int ggko_sort_choose(int dtype, const void *vec, int64_t n, const uint32_t *k,
                     int64_t nk, void *out) {
  if (!valid_dtype(dtype) || !vec || check_ks(n, k, nk)) return -1;
  void *keys = sorted_keys(dtype, vec, n);
  if (!keys) return -1;
  if (wide(dtype)) {
    uint64_t *s = (uint64_t *)keys, *o = (uint64_t *)out;
    for (int64_t i = 0; i < nk; ++i) o[i] = ggko_unkey_f64(s[n - k[i]]);
  } else {
    uint32_t *s = (uint32_t *)keys, *o = (uint32_t *)out;
    for (int64_t i = 0; i < nk; ++i) {
      uint32_t key = s[n - k[i]];
      o[i] = dtype == GGKO_F32 ? ggko_unkey_f32(key)
           : dtype == GGKO_I32 ? ggko_unkey_i32(key) : key;
    }
  }
  free(keys);
  return 0;
}

/* multiselectTimingFunctions.cu:155-177: thrust::sort(greater), copy first K.
   Descending order of converted keys is the reverse of the ascending sort. */
int ggko_topk(int dtype, const void *vec, int64_t n, int64_t K, void *out) {
  if (!valid_dtype(dtype) || !vec || n <= 0 || K < 1 || K > n) return -1;
  void *keys = sorted_keys(dtype, vec, n);
  if (!keys) return -1;
  size_t w = esize(dtype);
  char *rev = (char *)malloc((size_t)K * w);
  for (int64_t j = 0; j < K; ++j)
    memcpy(rev + j * w, (char *)keys + (size_t)(n - 1 - j) * w, w);
  ggko_unkeys(dtype, rev, out, K);
  free(rev);
  free(keys);
  return 0;
}

/* ------------------------------------------------------------------ */
/* MSD radix select, 4-bit digits.                                      */
/* merrillSelect.cu:62-129 drives one pass per digit (pass 0 = bits      */
/* 28..31 for 32-bit keys, RadixSelect_api.h:757-823; 60..63 for 64-bit, */
/* :887-1024).  Each pass histograms the digit over the current range,   */
/* determineRadixOfInterest (RadixSelect_api.h:450-500) picks the digit  */
/* whose cumulative count covers index (n - k) (inplaceRadixSelect.cu:30) */
/* and narrows the range; the answer's key is the digit sequence          */
/* (inplaceRadixSelect.cu:61-77), converted back by postProcess.          */
/* ------------------------------------------------------------------ */
int ggko_radix_select(int dtype, const void *vec, int64_t n, uint32_t k,
                      void *out, uint8_t *digits_out) {
  if (!valid_dtype(dtype) || !vec || !out || n <= 0 || k < 1 || (int64_t)k > n)
    return -1;
  int bits = wide(dtype) ? 64 : 32;
  int passes = bits / 4;
  void *keys = malloc((size_t)n * esize(dtype));
  ggko_keys(dtype, vec, keys, n);
  int64_t idx = n - (int64_t)k;  /* 0-based ascending index (indexOfK) */
  uint64_t prefix = 0;            /* digits chosen so far */
  int64_t live = n;               /* elements matching the prefix */
  uint64_t *k64 = (uint64_t *)keys;
  uint32_t *k32 = (uint32_t *)keys;
  for (int pass = 0; pass < passes; ++pass) {
    int shift = bits - 4 * (pass + 1);
    int64_t hist[16] = {0};
    int64_t w = 0;
    for (int64_t i = 0; i < live; ++i) {
      uint64_t key = wide(dtype) ? k64[i] : k32[i];
      hist[(key >> shift) & 15u]++;
    }
    int d = 0;
    int64_t below = 0;
    while (below + hist[d] <= idx) below += hist[d++];
    idx -= below;
    prefix = (prefix << 4) | (uint64_t)d;
    if (digits_out) digits_out[pass] = (uint8_t)d;
    /* keep only the digit of interest (the reference scatters it,
       ScanScatterDigits, or clamps the rest, compare_radix) */
    for (int64_t i = 0; i < live; ++i) {
      uint64_t key = wide(dtype) ? k64[i] : k32[i];
      if (((key >> shift) & 15u) == (uint64_t)d) {
        if (wide(dtype)) k64[w++] = key; else k32[w++] = (uint32_t)key;
      }
    }
    live = w;
  }
  if (wide(dtype)) {
    uint64_t v = ggko_unkey_f64(prefix);
    memcpy(out, &v, 8);
  } else {
    uint32_t key = (uint32_t)prefix;
    uint32_t v = dtype == GGKO_F32 ? ggko_unkey_f32(key)
               : dtype == GGKO_I32 ? ggko_unkey_i32(key) : key;
    memcpy(out, &v, 4);
  }
  free(keys);
  return 0;
}

/* ------------------------------------------------------------------ */
/* seeded generator for the pivot/splitter samples (the reference seeds  */
/* with gettimeofday, bucketMultiselect.cu:377-389; here SplitMix64).    */
/* ------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static float uniform01f(uint64_t *s) { /* 24-bit grid in [0,1) */
  return (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------ */
/* Bucket multiselect restatement, bucketMultiselect.cu:521-871.        */
/* Typed on T as in the reference: native comparisons against pivots,   */
/* double interpolation for the sub-bucket.                             */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t key; uint32_t k; uint32_t idx; } kpair;
static int cmp_kpair(const void *a, const void *b) {
  const kpair *x = (const kpair *)a, *y = (const kpair *)b;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx); /* stable as sort_by_key */
}
typedef struct { uint64_t key; uint32_t bucket; } kbpair;
static int cmp_kbpair(const void *a, const void *b) {
  const kbpair *x = (const kbpair *)a, *y = (const kbpair *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return 0;
}

#define DEFINE_BUCKET_MULTISELECT(NAME, T, DT)                                   \
  static int NAME(const T *v, int64_t n, const uint32_t *kori, int64_t nk,       \
                  T *out, uint64_t seed, ggko_bucket_stats *st) {                 \
    const int numBuckets = 8192, numPivots = 17; /* :841, :868 */                 \
    const int sampleSize = 1024, endOffset = 22;  /* :569, :467 */                \
    const int pivotOffset = (sampleSize - endOffset * 2) / (numPivots - 3);       \
    const int numSmallBuckets = numBuckets / (numPivots - 1);                     \
    if (st) { memset(st, 0, sizeof *st); st->num_buckets = numBuckets;            \
              st->monotone = 1; }                                                \
    if (nk == 0) return 0;                                                       \
    /* wrapper :845-846 -- turn k-th largest into k-th smallest */                \
    kpair *kp = (kpair *)malloc((size_t)nk * sizeof(kpair));                     \
    for (int64_t i = 0; i < nk; ++i) {                                           \
      kp[i].k = (uint32_t)(n - kori[i] + 1); kp[i].idx = (uint32_t)i; }          \
    /* step 1 :531-547 -- min / max, flat-vector exit */                          \
    T mn = v[0], mx = v[0];                                                      \
    for (int64_t i = 1; i < n; ++i) { if (v[i] < mn) mn = v[i];                  \
                                      if (mx < v[i]) mx = v[i]; }                 \
    if (mx == mn) {                                                              \
      for (int64_t i = 0; i < nk; ++i) out[i] = mn;                              \
      if (st) { st->flat = 1; st->total_count = n; }                              \
      free(kp); return 0; }                                                      \
    /* step 3 :615-646 -- sort k with original indices, strip k'=n / k'=1 */      \
    qsort(kp, (size_t)nk, sizeof(kpair), cmp_kpair);                             \
    int64_t lo = 0, hi = nk; /* live k range [lo, hi) */                          \
    while (hi > lo && (int64_t)kp[hi - 1].k == n) { out[kp[hi - 1].idx] = mx; hi--; } \
    while (hi > lo && kp[lo].k == 1) { out[kp[lo].idx] = mn; lo++; }             \
    if (hi == lo) { if (st) st->total_count = n; free(kp); return 0; }           \
    /* step 4 generatePivots :465-508 (seeded sample instead of time) */           \
    T *sample = (T *)malloc((size_t)sampleSize * sizeof(T));                    \
    uint64_t rs = seed;                                                          \
    for (int i = 0; i < sampleSize; ++i) {                                       \
      int64_t pos = (int64_t)(uniform01f(&rs) * (float)n);  /* :396-405 */       \
      if (pos >= n) pos = n - 1;                                                 \
      sample[i] = v[pos]; }                                                      \
    ggko_sort(DT, sample, sampleSize, sample);                                   \
    T pivots[17]; double slopes[16];                                             \
    pivots[0] = mn; pivots[numPivots - 1] = mx;                                  \
    pivots[1] = sample[endOffset - 1];                                           \
    pivots[numPivots - 2] = sample[sampleSize - endOffset - 1];                  \
    slopes[0] = numSmallBuckets / ((double)pivots[1] - (double)pivots[0]);       \
    for (int i = 2; i < numPivots - 2; ++i) {                                    \
      pivots[i] = sample[pivotOffset * (i - 1) + endOffset - 1];                 \
      slopes[i - 1] = numSmallBuckets / ((double)pivots[i] - (double)pivots[i - 1]); } \
    slopes[numPivots - 3] = numSmallBuckets /                                    \
        ((double)pivots[numPivots - 2] - (double)pivots[numPivots - 3]);         \
    slopes[numPivots - 2] = numSmallBuckets /                                    \
        ((double)pivots[numPivots - 1] - (double)pivots[numPivots - 2]);         \
    for (int i = 0; i < numPivots - 1; ++i) if (isinf(slopes[i])) slopes[i] = 0; /* :665-667 */ \
    free(sample);                                                                \
    /* step 5 assignSmartBucket :165-193 */                                       \
    uint32_t *bucketOf = (uint32_t *)malloc((size_t)n * 4);                     \
    int64_t *count = (int64_t *)calloc(numBuckets, sizeof(int64_t));            \
    for (int64_t i = 0; i < n; ++i) {                                            \
      T num = v[i];                                                              \
      int minP = 0, maxP = numPivots - 1, mid;                                   \
      for (int j = 1; j < numPivots - 1; j *= 2) {                               \
        mid = (maxP + minP) / 2;                                                 \
        if (num >= pivots[mid]) minP = mid; else maxP = mid; }                   \
      uint32_t b = (uint32_t)(minP * numSmallBuckets +                           \
          (int)(((double)num - (double)pivots[minP]) * slopes[minP]));           \
      if (b == (uint32_t)numBuckets) b = numBuckets - 1;                         \
      bucketOf[i] = b; count[b]++; }                                             \
    if (st) { int64_t t = 0; for (int b = 0; b < numBuckets; ++b) t += count[b]; \
              st->total_count = t;                                               \
              kbpair *kb = (kbpair *)malloc((size_t)n * sizeof(kbpair));          \
              for (int64_t i = 0; i < n; ++i) { uint64_t key;                    \
                if (DT == GGKO_F64) { uint64_t u; memcpy(&u, &v[i], 8); key = ggko_key_f64(u); } \
                else { uint32_t u; memcpy(&u, &v[i], 4);                          \
                  key = DT == GGKO_F32 ? ggko_key_f32(u) : DT == GGKO_I32 ? ggko_key_i32(u) : u; } \
                kb[i].key = key; kb[i].bucket = bucketOf[i]; }                    \
              qsort(kb, (size_t)n, sizeof(kbpair), cmp_kbpair);                   \
              /* equal keys may sit in either order; check per distinct key */    \
              uint32_t prevMax = 0;                                               \
              for (int64_t i = 0; i < n;) { int64_t j = i; uint32_t bmin = kb[i].bucket, bmax = bmin; \
                while (j < n && kb[j].key == kb[i].key) { if (kb[j].bucket < bmin) bmin = kb[j].bucket; \
                  if (kb[j].bucket > bmax) bmax = kb[j].bucket; j++; }            \
                if (bmin < prevMax) st->monotone = 0; prevMax = bmax; i = j; }    \
              free(kb); }                                                        \
    /* findKBuckets :87-111 */                                                    \
    int64_t numKs = hi - lo;                                                     \
    uint32_t *marked = (uint32_t *)malloc((size_t)numKs * 4);                   \
    int64_t *sums = (int64_t *)malloc((size_t)numKs * 8);                       \
    { int kBucket = 0; int64_t sum = count[0];                                   \
      for (int64_t i = 0; i < numKs; ++i) { int64_t kk = kp[lo + i].k;           \
        while (sum < kk && kBucket < numBuckets - 1) { kBucket++; sum += count[kBucket]; } \
        marked[i] = (uint32_t)kBucket; sums[i] = sum - count[kBucket]; } }       \
    /* host reindex :714-730 */                                                   \
    uint32_t *uniq = (uint32_t *)malloc((size_t)numKs * 4);                     \
    int64_t *reidx = (int64_t *)malloc((size_t)numKs * 8);                      \
    int64_t *newk = (int64_t *)malloc((size_t)numKs * 8);                       \
    int64_t nu = 1; uniq[0] = marked[0]; reidx[0] = 0;                           \
    newk[0] = kp[lo].k - sums[0];                                                \
    for (int64_t i = 1; i < numKs; ++i) {                                        \
      if (marked[i] != marked[i - 1]) { uniq[nu] = marked[i];                    \
        reidx[nu] = reidx[nu - 1] + count[marked[i - 1]]; nu++; }                \
      newk[i] = reidx[nu - 1] + kp[lo + i].k - sums[i]; }                        \
    int64_t newLen = reidx[nu - 1] + count[marked[numKs - 1]];                   \
    /* copyElements :249-308 -- bucket segments at the reindexed offsets */        \
    int64_t *cursor = (int64_t *)malloc((size_t)numBuckets * 8);                \
    for (int b = 0; b < numBuckets; ++b) cursor[b] = -1;                         \
    for (int64_t u = 0; u < nu; ++u) cursor[uniq[u]] = reidx[u];                 \
    T *newInput = (T *)malloc((size_t)(newLen > 0 ? newLen : 1) * sizeof(T));   \
    for (int64_t i = 0; i < n; ++i) { int64_t c = cursor[bucketOf[i]];           \
      if (c >= 0) newInput[cursor[bucketOf[i]]++] = v[i]; }                      \
    /* step 8 :799-800 sort, :813-814 copyValuesInChunk */                        \
    ggko_sort(DT, newInput, newLen, newInput);                                   \
    for (int64_t i = 0; i < numKs; ++i) out[kp[lo + i].idx] = newInput[newk[i] - 1]; \
    if (st) { st->reduced_length = newLen; st->unique_buckets = nu; }            \
    free(newInput); free(cursor); free(newk); free(reidx); free(uniq);          \
    free(sums); free(marked); free(count); free(bucketOf); free(kp);            \
    return 0;                                                                    \
  }

DEFINE_BUCKET_MULTISELECT(bucket_ms_f32, float, GGKO_F32)
DEFINE_BUCKET_MULTISELECT(bucket_ms_f64, double, GGKO_F64)
DEFINE_BUCKET_MULTISELECT(bucket_ms_u32, uint32_t, GGKO_U32)
DEFINE_BUCKET_MULTISELECT(bucket_ms_i32, int32_t, GGKO_I32)

int ggko_bucket_multiselect(int dtype, const void *vec, int64_t n,
                            const uint32_t *k, int64_t nk, void *out,
                            uint64_t seed, ggko_bucket_stats *stats) {
  if (!valid_dtype(dtype) || !vec || check_ks(n, k, nk) || (nk > 0 && !out))
    return -1;
  switch (dtype) {
    case GGKO_F32: return bucket_ms_f32((const float *)vec, n, k, nk, (float *)out, seed, stats);
    case GGKO_F64: return bucket_ms_f64((const double *)vec, n, k, nk, (double *)out, seed, stats);
    case GGKO_U32: return bucket_ms_u32((const uint32_t *)vec, n, k, nk, (uint32_t *)out, seed, stats);
    default: return bucket_ms_i32((const int32_t *)vec, n, k, nk, (int32_t *)out, seed, stats);
  }
}

/* ------------------------------------------------------------------ */
/* Randomized top-k restatement, randomizedTopkSelect.cu:128-213 and    */
/* randomizedSelectGenerateSplitters.cu:74-126, on converted keys.      */
/* Repairs: (i) the window used for the answer extends to the upper      */
/* sentinel, so no element above pivot1 is lost; (ii) rightBucket is     */
/* clamped to numSplitters (the reference can index one past the array, */
/* :36-38); (iii) sentinels are the smallest / largest keys for every   */
/* type (the reference writes 0xFFFFFFFF into a T, splitters.cu:84-85).  */
/* ------------------------------------------------------------------ */
static double binom_pdf(int k, double p, int N) { /* gsl_ran_binomial_pdf */
  if (k < 0 || k > N) return 0.0;
  if (p <= 0.0) return k == 0 ? 1.0 : 0.0;
  if (p >= 1.0) return k == N ? 1.0 : 0.0;
  double l = lgamma(N + 1.0) - lgamma(k + 1.0) - lgamma(N - k + 1.0) +
             k * log(p) + (N - k) * log1p(-p);
  return exp(l);
}

/* choosePivots, randomizedTopkSelect.cu:22-41 (with the clamp). */
static void choose_pivots(int64_t length, int64_t k, int numSplitters,
                          double desired, int *left, int *right) {
  double p = (double)k / (double)length;
  int most = (int)((double)k / ((double)length / (numSplitters + 1)));
  if (most > numSplitters) most = numSplitters;
  double q = binom_pdf(most, p, numSplitters);
  *left = most; *right = most;
  int guard = 0;
  while (q < desired && guard++ < 4 * numSplitters + 8) {
    if (*left > 0) { (*left)--; q += binom_pdf(*left, p, numSplitters); }
    if (*right < numSplitters) { (*right)++; q += binom_pdf(*right, p, numSplitters); }
    if (*left == 0 && *right == numSplitters) break;
  }
}

int ggko_randomized_topk(int dtype, const void *vec, int64_t n, int64_t K,
                         void *out, uint64_t seed) {
  if (!valid_dtype(dtype) || !vec || !out || n <= 0 || K < 1 || K > n) return -1;
  int w8 = wide(dtype);
  uint64_t *keys = (uint64_t *)malloc((size_t)n * 8);
  if (w8) memcpy(keys, vec, (size_t)n * 8), ggko_keys(dtype, keys, keys, n);
  else {
    uint32_t *k32 = (uint32_t *)malloc((size_t)n * 4);
    ggko_keys(dtype, vec, k32, n);
    for (int64_t i = 0; i < n; ++i) keys[i] = k32[i];
    free(k32);
  }
  const uint64_t KMAX = w8 ? ~0ull : 0xFFFFFFFFull;
  int64_t k = n - K + 1; /* :134 -- k-th smallest */
  int numSplitters = (int)(8 * sqrt((double)n)); /* :147 */
  if (numSplitters < 1) numSplitters = 1;
  /* generateSplitters + copyTheSplitters: sentinels at both ends */
  uint64_t *spl = (uint64_t *)malloc((size_t)(numSplitters + 2) * 8);
  uint64_t rs = seed;
  spl[0] = 0;
  for (int i = 0; i < numSplitters; ++i)
    spl[1 + i] = keys[splitmix64(&rs) % (uint64_t)n];
  spl[numSplitters + 1] = KMAX;
  radix_sort_u64(spl, numSplitters + 2); /* :158-160 */
  uint64_t *shifted = spl + 1;
  int left, right;
  choose_pivots(n, k, numSplitters, 0.90, &left, &right);
  uint64_t p0 = 0, p1 = KMAX;
  for (int iter = 0; iter < 4 * numSplitters + 8; ++iter) { /* :164-179 */
    p0 = shifted[left - 1];   /* left == 0 reads the lower sentinel */
    p1 = shifted[right];
    int64_t first = 0, middle = 0;
    for (int64_t i = 0; i < n; ++i) { /* queryCount :44-65 */
      if (keys[i] < p0) first++;
      else if (keys[i] <= p1) middle++;
    }
    int where = first >= k ? -1 : (first + middle >= k ? 0 : 1); /* :96-106 */
    if (where == 0) break;
    int nb = right - left + 1; /* shiftBuckets :109-125 */
    if (nb == 1) nb = 2;
    if (where == -1) { right = left; left = left - nb + 1 < 0 ? 0 : left - nb + 1; }
    else { left = right; right = right + nb - 1 > numSplitters ? numSplitters : right + nb - 1; }
  }
  { /* the window [p0, top] must hold at least K elements; otherwise (the
       retry budget ran out) take everything */
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) c += keys[i] >= p0;
    if (c < K) p0 = 0;
  }
  /* partition :67-84 with the window extended to the top, then sort
     descending (:188-189) and take the first K (:191-192). */
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) if (keys[i] >= p0) keys[m++] = keys[i];
  radix_sort_u64(keys, m);
  if (w8) {
    uint64_t *o = (uint64_t *)out;
    for (int64_t j = 0; j < K; ++j) o[j] = ggko_unkey_f64(keys[m - 1 - j]);
  } else {
    uint32_t *o = (uint32_t *)out;
    for (int64_t j = 0; j < K; ++j) {
      uint32_t key = (uint32_t)keys[m - 1 - j];
      o[j] = dtype == GGKO_F32 ? ggko_unkey_f32(key)
           : dtype == GGKO_I32 ? ggko_unkey_i32(key) : key;
    }
  }
  free(spl);
  free(keys);
  return 0;
}
