// nth_select.cpp -- second, independent CPU oracle and the CPU baselines.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Works on converted keys
// (keyConversion.cu:8-42) so the order matches the reference's radix order.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>
#include <parallel/algorithm>
#include <omp.h>

#include "oracle.h"

namespace {

template <typename K>
void to_keys(int dtype, const void* vec, int64_t n, std::vector<K>& keys) {
  keys.resize((size_t)n);
  ggko_keys(dtype, vec, keys.data(), n);
}

template <typename K>
void emit(int dtype, const K* keys, int64_t cnt, void* out) {
  ggko_unkeys(dtype, keys, out, cnt);
}

bool bad_ks(int64_t n, const uint32_t* k, int64_t nk) {
  if (n <= 0 || nk < 0 || (nk > 0 && !k)) return true;
  for (int64_t i = 0; i < nk; ++i)
    if (k[i] < 1 || (int64_t)k[i] > n) return true;
  return false;
}

// std::nth_element over the distinct k values, smallest rank first; each step
// narrows the range to the part right of the previous pick.  This is the
// "nth_element multiselect" CPU baseline: one core.
template <typename K>
int nth_multi(int dtype, const void* vec, int64_t n, const uint32_t* k,
              int64_t nk, void* out) {
  std::vector<K> keys;
  to_keys(dtype, vec, n, keys);
  std::vector<int64_t> idx((size_t)nk);  // ascending 0-based index of each k
  for (int64_t i = 0; i < nk; ++i) idx[i] = n - (int64_t)k[i];
  std::vector<int64_t> order(idx);
  std::sort(order.begin(), order.end());
  order.erase(std::unique(order.begin(), order.end()), order.end());
  std::vector<K> picked(order.size());
  int64_t from = 0;
  for (size_t j = 0; j < order.size(); ++j) {
    std::nth_element(keys.begin() + from, keys.begin() + order[j], keys.end());
    picked[j] = keys[order[j]];
    from = order[j] + 1;
  }
  std::vector<K> res((size_t)nk);
  for (int64_t i = 0; i < nk; ++i) {
    size_t j = std::lower_bound(order.begin(), order.end(), idx[i]) - order.begin();
    res[i] = picked[j];
  }
  emit(dtype, res.data(), nk, out);
  return 0;
}

template <typename K>
int nth_topk(int dtype, const void* vec, int64_t n, int64_t K_, void* out) {
  std::vector<K> keys;
  to_keys(dtype, vec, n, keys);
  std::partial_sort(keys.begin(), keys.begin() + K_, keys.end(), std::greater<K>());
  emit(dtype, keys.data(), K_, out);
  return 0;
}

template <typename K>
int par_sort_choose(int dtype, const void* vec, int64_t n, const uint32_t* k,
                    int64_t nk, void* out, int threads) {
  std::vector<K> keys;
  to_keys(dtype, vec, n, keys);
  int t = threads > 0 ? threads : omp_get_num_procs();
  omp_set_num_threads(t);
  __gnu_parallel::sort(keys.begin(), keys.end(), __gnu_parallel::multiway_mergesort_tag(t));
  std::vector<K> res((size_t)nk);
  for (int64_t i = 0; i < nk; ++i) res[i] = keys[n - (int64_t)k[i]];
  emit(dtype, res.data(), nk, out);
  return 0;
}

}  // namespace

extern "C" int ggko_nth_multiselect(int dtype, const void* vec, int64_t n,
                                    const uint32_t* k, int64_t nk, void* out) {
  if (dtype < 0 || dtype > 3 || !vec || bad_ks(n, k, nk)) return -1;
  return dtype == GGKO_F64 ? nth_multi<uint64_t>(dtype, vec, n, k, nk, out)
                           : nth_multi<uint32_t>(dtype, vec, n, k, nk, out);
}

extern "C" int ggko_nth_topk(int dtype, const void* vec, int64_t n, int64_t K,
                             void* out) {
  if (dtype < 0 || dtype > 3 || !vec || !out || n <= 0 || K < 1 || K > n) return -1;
  return dtype == GGKO_F64 ? nth_topk<uint64_t>(dtype, vec, n, K, out)
                           : nth_topk<uint32_t>(dtype, vec, n, K, out);
}

extern "C" int ggko_parallel_sort_choose(int dtype, const void* vec, int64_t n,
                                         const uint32_t* k, int64_t nk,
                                         void* out, int threads) {
  if (dtype < 0 || dtype > 3 || !vec || bad_ks(n, k, nk)) return -1;
  return dtype == GGKO_F64
             ? par_sort_choose<uint64_t>(dtype, vec, n, k, nk, out, threads)
             : par_sort_choose<uint32_t>(dtype, vec, n, k, nk, out, threads);
}
