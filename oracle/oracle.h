/*
 * oracle.h -- CPU restatement of GGKS's selection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libggks.so) never links, loads or calls it.
 *
 * dtype codes: 0 = float32, 1 = float64, 2 = uint32, 3 = int32.
 * k values are 1-based ranks of the k-th LARGEST element (k = 1 is the maximum),
 * as in the reference (bucketMultiselect.cu:845-846,
 * multiselectTimingFunctions.cu:41-48).  The total order is the order of the
 * converted keys (keyConversion.cu:8-42), i.e. the order the reference's radix
 * sorts produce; it differs from IEEE comparison only in ordering -0.0 before
 * +0.0.
 *
 * Return codes: 0 ok, -1 invalid argument (n <= 0, k out of [1, n], null).
 */
#ifndef GGKS_ORACLE_H
#define GGKS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { GGKO_F32 = 0, GGKO_F64 = 1, GGKO_U32 = 2, GGKO_I32 = 3 };

/* Key conversion, keyConversion.cu:10-24 (preProcess) and :26-42 (postProcess);
   b40c radixsort_key_conversion.h:112-130 (float), :142-159 (double), :249-265 (int). */
uint32_t ggko_key_f32(uint32_t bits);
uint32_t ggko_unkey_f32(uint32_t key);
uint64_t ggko_key_f64(uint64_t bits);
uint64_t ggko_unkey_f64(uint64_t key);
uint32_t ggko_key_i32(uint32_t bits);
uint32_t ggko_unkey_i32(uint32_t key);
/* bulk conversion; width = 4 or 8 bytes per element by dtype */
void ggko_keys(int dtype, const void *in, void *out, int64_t n);
void ggko_unkeys(int dtype, const void *in, void *out, int64_t n);

/* Ascending sort of values in converted-key order (thrust::sort on primitives ->
   b40c radix sort on converted keys, thrust/detail/device/cuda/sort.inl:259-275). */
int ggko_sort(int dtype, const void *vec, int64_t n, void *sorted_out);

/* Sort-and-choose multiselect, multiselectTimingFunctions.cu:41-95:
   out[i] = sorted_ascending[n - k[i]]. */
int ggko_sort_choose(int dtype, const void *vec, int64_t n, const uint32_t *k,
                     int64_t nk, void *out);

/* Sort-and-choose top-K, multiselectTimingFunctions.cu:155-177:
   out[j] = sorted_descending[j], j < K. */
int ggko_topk(int dtype, const void *vec, int64_t n, int64_t K, void *out);

/* Single-k MSD radix select with 4-bit digits, merrillSelect.cu:62-129 and
   inplaceRadixSelect.cu:24-80 (RadixSelect_api.h:450-500 picks the digit).
   digits_out (may be NULL) receives the 8 (f32/u32/i32) or 16 (f64) digits,
   most significant first. */
int ggko_radix_select(int dtype, const void *vec, int64_t n, uint32_t k,
                      void *out, uint8_t *digits_out);

/* Statistics from the bucket multiselect restatement (intermediates the
   reference never exposes; used to check invariants). */
typedef struct {
  int64_t num_buckets;      /* 8192 */
  int64_t total_count;      /* sum of bucket counts, must equal n */
  int64_t reduced_length;   /* newInputLength */
  int64_t unique_buckets;   /* numUniqueBuckets */
  int32_t flat;             /* 1 if max == min (early exit) */
  int32_t monotone;         /* 1 if bucket(x) is non-decreasing in key order */
} ggko_bucket_stats;

/* CPU restatement of bucketMultiselectWrapper + phaseOne,
   bucketMultiselect.cu:521-871, with the pivot sample drawn from a seeded
   generator instead of gettimeofday (:377-389). */
int ggko_bucket_multiselect(int dtype, const void *vec, int64_t n,
                            const uint32_t *k, int64_t nk, void *out,
                            uint64_t seed, ggko_bucket_stats *stats);

/* CPU restatement of randomizedTopkSelect, randomizedTopkSelect.cu:128-213,
   with its three hazards repaired (upper sentinel is the largest key, bucket
   indices are clamped, the final window extends to the upper sentinel). */
int ggko_randomized_topk(int dtype, const void *vec, int64_t n, int64_t K,
                         void *out, uint64_t seed);

/* Independent second oracle: std::nth_element over the distinct k values in
   ascending order, narrowing the range each step (nth_select.cpp).  Works on a
   private copy; returns the same values as ggko_sort_choose. */
int ggko_nth_multiselect(int dtype, const void *vec, int64_t n,
                         const uint32_t *k, int64_t nk, void *out);
/* top-K via std::partial_sort on converted keys (descending). */
int ggko_nth_topk(int dtype, const void *vec, int64_t n, int64_t K, void *out);

/* Sort-and-choose with an OpenMP parallel sort (libstdc++ parallel mode),
   standing in for Thrust's OMP backend.  threads <= 0 means all cores. */
int ggko_parallel_sort_choose(int dtype, const void *vec, int64_t n,
                              const uint32_t *k, int64_t nk, void *out,
                              int threads);

#ifdef __cplusplus
}
#endif
#endif
