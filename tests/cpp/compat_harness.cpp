// Differential harness in the shape of the reference's compareMultiselect /
// compareTopkselect (compareMultiselect.hpp:164-185, compareTopkselect.hpp:
// 164-185): for each problem, run the algorithm through the reference-named
// C++ wrappers (include/ggks_compat.hpp) and compare with sort-and-choose
// computed here with std::sort on a host copy (multiselectTimingFunctions.cu:
// 41-95, :155-177).  Inputs come from the library's seeded device generator.
// Prints one line per problem and exits non-zero on any mismatch.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ggks_compat.hpp"

template <typename T>
static bool same(const T& a, const T& b) {
  return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// 101 percentile k values of a length-n vector, k = round(p (n - 1)) + 1
// (p = 0 is the maximum)
static std::vector<unsigned int> percentiles(int n) {
  std::vector<unsigned int> k;
  for (int p = 0; p <= 100; ++p) k.push_back((unsigned)((p * (double)(n - 1)) / 100.0 + 0.5) + 1);
  return k;
}

// typed C entry points of ggks.h per key type
static int typed_radix(float* d, int64_t n, const uint32_t* k, int32_t nk, float* o) {
  return ggks_radix_multiselect_f32(d, n, k, nk, o, nullptr);
}
static int typed_radix(double* d, int64_t n, const uint32_t* k, int32_t nk, double* o) {
  return ggks_radix_multiselect_f64(d, n, k, nk, o, nullptr);
}
static int typed_radix(unsigned int* d, int64_t n, const uint32_t* k, int32_t nk, unsigned int* o) {
  return ggks_radix_multiselect_u32(d, n, k, nk, o, nullptr);
}
static int typed_radix(int* d, int64_t n, const uint32_t* k, int32_t nk, int* o) {
  return ggks_radix_multiselect_i32(d, n, k, nk, o, nullptr);
}
static int typed_topk(float* d, int64_t n, int32_t K, float* o) {
  return ggks_topk_f32(d, n, K, o, nullptr);
}
static int typed_topk(double* d, int64_t n, int32_t K, double* o) {
  return ggks_topk_f64(d, n, K, o, nullptr);
}
static int typed_topk(unsigned int* d, int64_t n, int32_t K, unsigned int* o) {
  return ggks_topk_u32(d, n, K, o, nullptr);
}
static int typed_topk(int* d, int64_t n, int32_t K, int* o) {
  return ggks_topk_i32(d, n, K, o, nullptr);
}

// InplaceRadix::inplaceRadixSelectWrapper has uint, float and double
// overloads only (inplaceRadixSelect.cu:83-108)
template <typename T>
static int inplace_check(T* d, int n, unsigned int k, const std::vector<T>& h) {
  const T v = InplaceRadix::inplaceRadixSelectWrapper(d, (unsigned)n, k);
  return !same(v, h[(size_t)n - k]);
}
template <>
int inplace_check<int>(int*, int, unsigned int, const std::vector<int>&) {
  return 0;
}

template <typename T>
static int problem(const char* name, int family, uint64_t seed, int n, int64_t lo, int64_t hi,
                   const std::vector<unsigned int>& ks, int K) {
  T* d = nullptr;
  ggks_compat::check(ggks_dev_alloc(reinterpret_cast<void**>(&d), sizeof(T) * (size_t)n));
  ggks_compat::check(ggks_generate(d, ggks_compat::dtype_of<T>::value, family, seed, 0, n, lo, hi));
  std::vector<T> h((size_t)n);
  ggks_compat::check(ggks_d2h(h.data(), d, sizeof(T) * (size_t)n));
  std::sort(h.begin(), h.end());  // NaN-free, no -0.0: the radix order
  int bad = 0;
  std::vector<unsigned int> k = ks;
  std::vector<T> out(k.size());
  // the reference's wrapper returns T(1) (bucketMultiselect.cu:837-838, :870)
  const T one = BucketMultiselect::bucketMultiselectWrapper<T>(d, n, k.data(), (int)k.size(),
                                                              out.data(), 0, 0);
  bad += !(one == (T)1);
  for (size_t i = 0; i < k.size(); ++i) bad += !same(out[i], h[(size_t)n - k[i]]);
  const T r = RadixSelect::RadixSelectWrapper<T>(d, (unsigned)n, k[k.size() / 2]);
  bad += !same(r, h[(size_t)n - k[k.size() / 2]]);
  T* top = randomizedTopkSelectWrapper<T>(d, n, K);
  for (int j = 0; j < K; ++j) bad += !same(top[j], h[(size_t)n - 1 - j]);
  std::free(top);
  // the typed C entry points (radix multiselect over every k, top-K)
  std::vector<T> rout(k.size()), tout((size_t)K);
  ggks_compat::check(typed_radix(d, n, k.data(), (int32_t)k.size(), rout.data()));
  for (size_t i = 0; i < k.size(); ++i) bad += !same(rout[i], h[(size_t)n - k[i]]);
  ggks_compat::check(typed_topk(d, n, K, tout.data()));
  for (int j = 0; j < K; ++j) bad += !same(tout[j], h[(size_t)n - 1 - j]);
  bad += inplace_check<T>(d, n, k[0], h);
  // the input must be intact (the reference's radix paths overwrite it)
  std::vector<T> again((size_t)n);
  ggks_compat::check(ggks_d2h(again.data(), d, sizeof(T) * (size_t)n));
  std::sort(again.begin(), again.end());
  bad += !std::equal(again.begin(), again.end(), h.begin(), same<T>);
  ggks_compat::check(ggks_dev_free(d));
  std::printf("%-28s n=%-9d k=%zu K=%d  %s\n", name, n, k.size(), K, bad ? "MISMATCH" : "OK");
  return bad;
}

int main() {
  int bad = 0;
  try {
    const int n1 = 1 << 20;
    bad += problem<float>("uniform f32", 0, 0, n1, 0, 0, percentiles(n1), 64);
    bad += problem<double>("normal f64", 1, 2, n1 + 3, 0, 0, percentiles(n1 + 3), 1000);
    bad += problem<unsigned int>("u32 in [0,1023]", 0, 3, 3 << 20, 0, 1023, percentiles(3 << 20),
                                 4096);
    bad += problem<int>("i32 full range", 0, 4, 777777, -2147483648ll, 2147483647ll,
                        percentiles(777777), 17);
    bad += problem<float>("small f32 (one work-group)", 1, 5, 3000, 0, 0, {1, 2, 1500, 2999, 3000}, 3000);
  } catch (const ggks_compat::error& e) {
    std::printf("error %d: %s\n", e.code, e.what());
    return 2;
  }
  std::printf("%s\n", bad ? "FAILED" : "ALL OK");
  return bad ? 1 : 0;
}
