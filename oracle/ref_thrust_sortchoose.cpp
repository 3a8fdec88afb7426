// ref_thrust_sortchoose.cpp -- the reference's sort-and-choose oracle
// (multiselectTimingFunctions.cu:50-95, top-k :155-177) driven through the
// reference's own vendored Thrust 1.4.0, built host-only with its OpenMP device
// backend (THRUST_DEVICE_BACKEND=2, thrust/detail/config/device_backend.h:24-29).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile from the headers under
// /root/reference into oracle/_ref/ (never committed); used as a checker and as
// the "reference" CPU baseline.  Only this driver file is ours.
#include <thrust/device_vector.h>
#include <thrust/functional.h>
#include <thrust/sort.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <omp.h>

namespace {

template <typename T>
int sort_choose(const T* h, int64_t n, const uint32_t* k, int64_t nk, T* out,
                double* sort_ms) {
  thrust::device_vector<T> d(h, h + n);  // setupForTiming copy-in (:17-26)
  auto t0 = std::chrono::steady_clock::now();
  thrust::sort(d.begin(), d.end());  // :61
  for (int64_t i = 0; i < nk; ++i) out[i] = d[n - k[i]];  // copyInChunk :41-48
  auto t1 = std::chrono::steady_clock::now();
  if (sort_ms) *sort_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

template <typename T>
int topk(const T* h, int64_t n, int64_t K, T* out, double* sort_ms) {
  thrust::device_vector<T> d(h, h + n);
  auto t0 = std::chrono::steady_clock::now();
  thrust::sort(d.begin(), d.end(), thrust::greater<T>());  // :166
  thrust::copy(d.begin(), d.begin() + K, out);              // :168
  auto t1 = std::chrono::steady_clock::now();
  if (sort_ms) *sort_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return 0;
}

}  // namespace

extern "C" {
// dtype: 0 f32, 1 f64, 2 u32, 3 i32.  threads <= 0: all cores.
int ggkr_sort_choose(int dtype, const void* h, int64_t n, const uint32_t* k,
                     int64_t nk, void* out, int threads, double* sort_ms) {
  if (n <= 0) return -1;
  for (int64_t i = 0; i < nk; ++i) if (k[i] < 1 || (int64_t)k[i] > n) return -1;
  omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
  switch (dtype) {
    case 0: return sort_choose((const float*)h, n, k, nk, (float*)out, sort_ms);
    case 1: return sort_choose((const double*)h, n, k, nk, (double*)out, sort_ms);
    case 2: return sort_choose((const unsigned*)h, n, k, nk, (unsigned*)out, sort_ms);
    case 3: return sort_choose((const int*)h, n, k, nk, (int*)out, sort_ms);
  }
  return -1;
}
int ggkr_topk(int dtype, const void* h, int64_t n, int64_t K, void* out,
              int threads, double* sort_ms) {
  if (n <= 0 || K < 1 || K > n) return -1;
  omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
  switch (dtype) {
    case 0: return topk((const float*)h, n, K, (float*)out, sort_ms);
    case 1: return topk((const double*)h, n, K, (double*)out, sort_ms);
    case 2: return topk((const unsigned*)h, n, K, (unsigned*)out, sort_ms);
    case 3: return topk((const int*)h, n, K, (int*)out, sort_ms);
  }
  return -1;
}
}
