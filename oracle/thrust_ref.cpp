// Driver around the reference's own Thrust OpenMP stable sort (TEST INFRASTRUCTURE:
// compiled by oracle/Makefile against the Thrust/CUB/libcudacxx headers under
// /root/reference into oracle/_ref/libthrust_ref.so; used by tests to pin the oracle
// and by the benchmark as the CPU baseline).  The reference code that runs is
// thrust/thrust/system/omp/detail/sort.h (stable_sort / stable_sort_by_key on the OMP
// system: per-thread sequential sort, then pairwise stable merges).
//
// Comparator modes:
//   0  Thrust's default (less / greater): primitive radix path per thread
//      (thrust/thrust/system/detail/sequential/stable_radix_sort.h)
//   1  a plain functor a < b / a > b: forces the stable merge-sort path; +-0.0 compare
//      equal, as in CUB's radix sort
//   2  a functor comparing the [begin_bit, end_bit) window of the order-preserving bit
//      transform: CUB's bit-window semantics
#include <thrust/functional.h>
#include <thrust/sort.h>
#include <thrust/system/omp/execution_policy.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <omp.h>
#include <type_traits>

namespace {

template <class K>
struct Less
{
  bool operator()(const K& a, const K& b) const
  {
    return a < b;
  }
};

template <class K>
struct Greater
{
  bool operator()(const K& a, const K& b) const
  {
    return a > b;
  }
};

template <class K>
struct WindowCompare
{
  int begin_bit;
  int end_bit;
  bool descending;

  using U = std::make_unsigned_t<std::conditional_t<std::is_floating_point_v<K>,
                                                    std::conditional_t<sizeof(K) == 4, int32_t, int64_t>,
                                                    K>>;
  U striped(K k) const
  {
    U u;
    std::memcpy(&u, &k, sizeof(K));
    constexpr int bits = int(sizeof(K) * 8);
    constexpr U high   = U(U(1) << (bits - 1));
    if constexpr (std::is_floating_point_v<K>)
    {
      if (u == high)
      {
        u = 0;
      }
      u = (u & high) ? U(~u) : U(u ^ high);
    }
    else if constexpr (std::is_signed_v<K>)
    {
      u = U(u ^ high);
    }
    const int w = end_bit - begin_bit;
    if (begin_bit > 0 || end_bit < bits)
    {
      const U m = w >= bits ? U(~U(0)) : U((U(1) << w) - 1);
      u &= U(m << begin_bit);
    }
    return u;
  }
  bool operator()(const K& a, const K& b) const
  {
    return descending ? striped(a) > striped(b) : striped(a) < striped(b);
  }
};

template <class K, class V>
void run(K* k, V* v, uint64_t n, int descending, int begin_bit, int end_bit, int mode)
{
  auto pol = thrust::omp::par;
  if (mode == 0)
  {
    if (descending)
    {
      if constexpr (std::is_void_v<V>)
        thrust::stable_sort(pol, k, k + n, thrust::greater<K>());
      else
        thrust::stable_sort_by_key(pol, k, k + n, v, thrust::greater<K>());
    }
    else
    {
      if constexpr (std::is_void_v<V>)
        thrust::stable_sort(pol, k, k + n);
      else
        thrust::stable_sort_by_key(pol, k, k + n, v);
    }
  }
  else if (mode == 1)
  {
    if (descending)
    {
      if constexpr (std::is_void_v<V>)
        thrust::stable_sort(pol, k, k + n, Greater<K>());
      else
        thrust::stable_sort_by_key(pol, k, k + n, v, Greater<K>());
    }
    else
    {
      if constexpr (std::is_void_v<V>)
        thrust::stable_sort(pol, k, k + n, Less<K>());
      else
        thrust::stable_sort_by_key(pol, k, k + n, v, Less<K>());
    }
  }
  else
  {
    WindowCompare<K> c{begin_bit, end_bit, descending != 0};
    if constexpr (std::is_void_v<V>)
      thrust::stable_sort(pol, k, k + n, c);
    else
      thrust::stable_sort_by_key(pol, k, k + n, v, c);
  }
}

template <class K>
int run_values(void* keys, void* vals, uint64_t n, int val_bytes, int descending, int begin_bit, int end_bit, int mode)
{
  K* k = static_cast<K*>(keys);
  switch (val_bytes)
  {
    case 0:
      run<K, void>(k, nullptr, n, descending, begin_bit, end_bit, mode);
      return 0;
    case 4:
      run<K, uint32_t>(k, static_cast<uint32_t*>(vals), n, descending, begin_bit, end_bit, mode);
      return 0;
    case 8:
      run<K, uint64_t>(k, static_cast<uint64_t*>(vals), n, descending, begin_bit, end_bit, mode);
      return 0;
    default:
      return 1;
  }
}

} // namespace

extern "C" {

// Sorts host arrays in place with the reference's Thrust OMP backend on `threads`
// threads.  key_type uses cccl_type_enum values (0..3 int8..int64, 4..7 uint8..uint64,
// 9 float32, 10 float64).  Returns the wall time of the sort call in seconds, or -1.
double ref_stable_sort(
  int key_type, void* keys, void* vals, uint64_t n, int val_bytes, int descending, int begin_bit, int end_bit, int mode, int threads)
{
  if (threads > 0)
  {
    omp_set_num_threads(threads);
  }
  const auto t0 = std::chrono::steady_clock::now();
  int rc        = 1;
  switch (key_type)
  {
    case 0:
      rc = run_values<int8_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 1:
      rc = run_values<int16_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 2:
      rc = run_values<int32_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 3:
      rc = run_values<int64_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 4:
      rc = run_values<uint8_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 5:
      rc = run_values<uint16_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 6:
      rc = run_values<uint32_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 7:
      rc = run_values<uint64_t>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 9:
      rc = run_values<float>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    case 10:
      rc = run_values<double>(keys, vals, n, val_bytes, descending, begin_bit, end_bit, mode);
      break;
    default:
      rc = 1;
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (rc != 0)
  {
    return -1.0;
  }
  return std::chrono::duration<double>(t1 - t0).count();
}

int ref_max_threads(void)
{
  return omp_get_max_threads();
}

} // extern "C"
