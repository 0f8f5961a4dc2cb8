// thrust::sort / stable_sort / sort_by_key / stable_sort_by_key on device memory, routed
// to the MI355X radix sort the way Thrust's CUDA backend routes them to CUB:
// thrust/thrust/system/cuda/detail/sort.h:142-266 (the radix route: a DoubleBuffer whose
// alternate half is temporary storage, the result copied back when it ends there) and
// :288-335 (taken when cub/cub/device/device_radix_sort.cuh:48-63 __can_use_radix_sort
// holds: arithmetic or __half keys compared with std::less / std::greater).
//
// The calls sort in place, on `stream` (default: the null stream), and return without
// synchronising: temporary storage comes from hipMallocAsync and goes back with
// hipFreeAsync on the same stream.  Comparators other than std::less / std::greater
// (Thrust's merge-sort route) are rejected at compile time.  Both sorts are stable,
// as the radix route is.  Returns hipError_t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <type_traits>

#include "device_radix_sort.hpp"

namespace cccl_amd::thrust_compat {

template <typename K, typename Compare>
inline constexpr bool can_use_radix_sort =
  (std::is_arithmetic_v<K> || std::is_same_v<K, __half>)
  && (std::is_same_v<std::remove_cv_t<Compare>, std::less<>> || std::is_same_v<std::remove_cv_t<Compare>, std::less<K>>
      || std::is_same_v<std::remove_cv_t<Compare>, std::greater<>>
      || std::is_same_v<std::remove_cv_t<Compare>, std::greater<K>>);

namespace detail {

template <typename Compare, typename K>
constexpr bool descending()
{
  using C = std::remove_cv_t<Compare>;
  return std::is_same_v<C, std::greater<>> || std::is_same_v<C, std::greater<K>>;
}

constexpr size_t round_up(size_t x, size_t a)
{
  return (x + a - 1) / a * a;
}

// thrust/thrust/system/cuda/detail/sort.h:219-266.
template <bool Pairs, bool Desc, typename K, typename V>
hipError_t radix_sort_in_place(K* keys, V* items, size_t count, hipStream_t stream)
{
  if (count == 0)
  {
    return hipSuccess;
  }
  DoubleBuffer<K> kb(keys, nullptr);
  DoubleBuffer<V> ib(items, nullptr);
  auto run = [&](void* t, size_t& bytes) -> hipError_t {
    if constexpr (Pairs)
    {
      return Desc ? DeviceRadixSort::SortPairsDescending(t, bytes, kb, ib, count, 0, int(sizeof(K) * 8), stream)
                  : DeviceRadixSort::SortPairs(t, bytes, kb, ib, count, 0, int(sizeof(K) * 8), stream);
    }
    else
    {
      return Desc ? DeviceRadixSort::SortKeysDescending(t, bytes, kb, count, 0, int(sizeof(K) * 8), stream)
                  : DeviceRadixSort::SortKeys(t, bytes, kb, count, 0, int(sizeof(K) * 8), stream);
    }
  };
  size_t temp_bytes = 0;
  if (hipError_t e = run(nullptr, temp_bytes); e != hipSuccess)
  {
    return e;
  }
  const size_t keys_bytes  = round_up(sizeof(K) * count, 256);
  const size_t items_bytes = Pairs ? round_up(sizeof(V) * count, 256) : 0;
  char* tmp                = nullptr;
  if (hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&tmp), keys_bytes + items_bytes + temp_bytes, stream);
      e != hipSuccess)
  {
    return e;
  }
  kb.d_buffers[1] = reinterpret_cast<K*>(tmp);
  ib.d_buffers[1] = reinterpret_cast<V*>(tmp + keys_bytes);
  hipError_t e    = run(tmp + keys_bytes + items_bytes, temp_bytes);
  if (e == hipSuccess && kb.selector != 0)
  {
    e = hipMemcpyAsync(keys, kb.d_buffers[1], sizeof(K) * count, hipMemcpyDeviceToDevice, stream);
  }
  if constexpr (Pairs)
  {
    if (e == hipSuccess && ib.selector != 0)
    {
      e = hipMemcpyAsync(items, ib.d_buffers[1], sizeof(V) * count, hipMemcpyDeviceToDevice, stream);
    }
  }
  const hipError_t f = hipFreeAsync(tmp, stream);
  return e != hipSuccess ? e : f;
}

} // namespace detail

template <typename K, typename Compare = std::less<K>>
hipError_t sort(K* first, K* last, Compare = Compare{}, hipStream_t stream = nullptr)
{
  static_assert(can_use_radix_sort<K, Compare>,
                "only arithmetic keys compared with std::less / std::greater take the radix route");
  return detail::radix_sort_in_place<false, detail::descending<Compare, K>(), K, cccl_amd::detail::NullValue>(
    first, nullptr, size_t(last - first), stream);
}

template <typename K, typename Compare = std::less<K>>
hipError_t stable_sort(K* first, K* last, Compare comp = Compare{}, hipStream_t stream = nullptr)
{
  return sort(first, last, comp, stream);
}

template <typename K, typename V, typename Compare = std::less<K>>
hipError_t sort_by_key(K* keys_first, K* keys_last, V* values_first, Compare = Compare{}, hipStream_t stream = nullptr)
{
  static_assert(can_use_radix_sort<K, Compare>,
                "only arithmetic keys compared with std::less / std::greater take the radix route");
  return detail::radix_sort_in_place<true, detail::descending<Compare, K>(), K, V>(
    keys_first, values_first, size_t(keys_last - keys_first), stream);
}

template <typename K, typename V, typename Compare = std::less<K>>
hipError_t stable_sort_by_key(K* keys_first, K* keys_last, V* values_first, Compare comp = Compare{},
                              hipStream_t stream = nullptr)
{
  return sort_by_key(keys_first, keys_last, values_first, comp, stream);
}

} // namespace cccl_amd::thrust_compat
