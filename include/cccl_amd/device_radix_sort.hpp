// C++ host API of the MI355X sorts in the shape of CUB's device-wide algorithms:
// cccl_amd::DeviceRadixSort, DeviceSegmentedRadixSort, DeviceSegmentedSort, DeviceTopK
// and DoubleBuffer, each a thin inline layer over the typed C entry points of
// radix_sort_amd.h / topk.h (libcccl_amd.so).  Argument lists and defaults follow
//   cub/cub/device/device_radix_sort.cuh:412-425 (SortPairs), :1127-1137 (DoubleBuffer
//     form), :1776 (SortPairsDescending), :3034 (SortKeys), :4211 (SortKeysDescending);
//   cub/cub/device/device_segmented_radix_sort.cuh:234-247, :417, :887, :1531, :2139;
//   cub/cub/device/device_segmented_sort.cuh:361 (SortKeys), :862 (DoubleBuffer form),
//     :1361 (StableSortKeys), :2468 (SortPairs), :3638 (StableSortPairs);
//   cub/cub/device/device_topk.cuh:297 (MaxPairs), :538 (MaxKeys), :775 (MinPairs),
// with a hipStream_t where CUB takes a cudaStream_t (DeviceTopK takes the stream in
// place of CUB's execution environment).  Every call is the two-phase protocol:
// d_temp_storage == nullptr only writes temp_storage_bytes.  Returns hipError_t.
//
// Keys: bool, [u]int8/16/32/64, __half (or any 2-byte floating type with
// is_floating_point), float, double.  Values: any trivially copyable type of 1, 2, 4, 8 or
// 16 bytes.  Segment offsets: pointers to int32 / int64 (or their unsigned variants).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "radix_sort_amd.h"
#include "topk.h"

namespace cccl_amd {

// cub::DoubleBuffer (cub/cub/util_type.cuh:750-778).
template <typename T>
struct DoubleBuffer
{
  T* d_buffers[2] = {nullptr, nullptr};
  int selector    = 0;

  DoubleBuffer() = default;
  DoubleBuffer(T* d_current, T* d_alternate)
      : d_buffers{d_current, d_alternate}
      , selector(0)
  {}
  T* Current() const
  {
    return d_buffers[selector];
  }
  T* Alternate() const
  {
    return d_buffers[selector ^ 1];
  }
};

namespace detail {

struct NullValue
{};

template <typename K>
constexpr cccl_type_enum key_type()
{
  using T = std::remove_cv_t<K>;
  if constexpr (std::is_same_v<T, bool>)
  {
    return CCCL_BOOLEAN;
  }
  else if constexpr (std::is_floating_point_v<T> || std::is_same_v<T, __half>)
  {
    static_assert(sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "unsupported floating-point key");
    return sizeof(T) == 2 ? CCCL_FLOAT16 : (sizeof(T) == 4 ? CCCL_FLOAT32 : CCCL_FLOAT64);
  }
  else
  {
    static_assert(std::is_integral_v<T>, "keys are arithmetic types");
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "unsupported key width");
    if constexpr (std::is_signed_v<T>)
    {
      return sizeof(T) == 1 ? CCCL_INT8 : sizeof(T) == 2 ? CCCL_INT16 : sizeof(T) == 4 ? CCCL_INT32 : CCCL_INT64;
    }
    else
    {
      return sizeof(T) == 1 ? CCCL_UINT8 : sizeof(T) == 2 ? CCCL_UINT16 : sizeof(T) == 4 ? CCCL_UINT32 : CCCL_UINT64;
    }
  }
}

template <typename V>
constexpr size_t value_bytes()
{
  if constexpr (std::is_same_v<V, NullValue>)
  {
    return 0;
  }
  else
  {
    static_assert(std::is_trivially_copyable_v<V>, "values are copied as bytes");
    static_assert(sizeof(V) == 1 || sizeof(V) == 2 || sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16,
                  "value types of 1, 2, 4, 8 or 16 bytes");
    return sizeof(V);
  }
}

template <typename OffsetPtr>
constexpr int offset_bytes()
{
  using O = std::remove_cv_t<std::remove_pointer_t<OffsetPtr>>;
  static_assert(std::is_pointer_v<OffsetPtr> && std::is_integral_v<O> && (sizeof(O) == 4 || sizeof(O) == 8),
                "segment offsets are pointers to 32- or 64-bit integers");
  return int(sizeof(O));
}

inline CUstream as_cu(hipStream_t s)
{
  return reinterpret_cast<CUstream>(s);
}

template <typename K, typename V>
hipError_t radix(void* t, size_t& b, K* k0, K* k1, V* v0, V* v1, int& sel, uint64_t n, int bb, int eb, bool desc,
                 bool overwrite, hipStream_t s)
{
  return hipError_t(cccl_amd_radix_sort(t, &b, (void*) k0, (void*) k1, (void*) v0, (void*) v1, &sel, key_type<K>(),
                                        value_bytes<V>(), n, bb, eb, desc, overwrite, nullptr, as_cu(s)));
}

template <typename K, typename V, typename B, typename E>
hipError_t seg_radix(void* t, size_t& b, K* k0, K* k1, V* v0, V* v1, int& sel, int64_t n, int64_t nseg, B beg, E end,
                     int bb, int eb, bool desc, bool overwrite, hipStream_t s)
{
  static_assert(offset_bytes<B>() == offset_bytes<E>(), "begin and end offsets of one width");
  return hipError_t(cccl_amd_segmented_radix_sort(
    t, &b, (void*) k0, (void*) k1, (void*) v0, (void*) v1, &sel, key_type<K>(), value_bytes<V>(), uint64_t(n), nseg,
    (const void*) beg, (const void*) end, offset_bytes<B>(), bb, eb, desc, overwrite, nullptr, as_cu(s)));
}

template <typename K, typename V, typename B, typename E>
hipError_t seg_sort(void* t, size_t& b, K* k0, K* k1, V* v0, V* v1, int& sel, int64_t n, int64_t nseg, B beg, E end,
                    bool desc, bool overwrite, hipStream_t s)
{
  static_assert(offset_bytes<B>() == offset_bytes<E>(), "begin and end offsets of one width");
  return hipError_t(cccl_amd_segmented_sort(
    t, &b, (void*) k0, (void*) k1, (void*) v0, (void*) v1, &sel, key_type<K>(), value_bytes<V>(), uint64_t(n), nseg,
    (const void*) beg, (const void*) end, offset_bytes<B>(), desc, overwrite, nullptr, as_cu(s)));
}

} // namespace detail

// cub::DeviceRadixSort.  Pointer forms leave the input untouched and write the result to
// the output arrays; DoubleBuffer forms may use both buffers and leave the result in
// d_keys.Current() / d_values.Current().
struct DeviceRadixSort
{
  template <typename KeyT, typename ValueT, typename NumItemsT>
  static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                              KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out, NumItemsT num_items,
                              int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::radix(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), d_keys_out,
                         const_cast<ValueT*>(d_values_in), d_values_out, sel, uint64_t(num_items), begin_bit, end_bit,
                         false, false, stream);
  }

  template <typename KeyT, typename ValueT, typename NumItemsT>
  static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                              DoubleBuffer<ValueT>& d_values, NumItemsT num_items, int begin_bit = 0,
                              int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    return db_pairs(d_temp_storage, temp_storage_bytes, d_keys, d_values, num_items, begin_bit, end_bit, false, stream);
  }

  template <typename KeyT, typename ValueT, typename NumItemsT>
  static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                        KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out,
                                        NumItemsT num_items, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                                        hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::radix(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), d_keys_out,
                         const_cast<ValueT*>(d_values_in), d_values_out, sel, uint64_t(num_items), begin_bit, end_bit,
                         true, false, stream);
  }

  template <typename KeyT, typename ValueT, typename NumItemsT>
  static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                        DoubleBuffer<ValueT>& d_values, NumItemsT num_items, int begin_bit = 0,
                                        int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    return db_pairs(d_temp_storage, temp_storage_bytes, d_keys, d_values, num_items, begin_bit, end_bit, true, stream);
  }

  template <typename KeyT, typename NumItemsT>
  static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,
                             NumItemsT num_items, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                             hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::radix<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in),
                                                  d_keys_out, nullptr, nullptr, sel, uint64_t(num_items), begin_bit,
                                                  end_bit, false, false, stream);
  }

  template <typename KeyT, typename NumItemsT>
  static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                             NumItemsT num_items, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                             hipStream_t stream = nullptr)
  {
    return db_keys(d_temp_storage, temp_storage_bytes, d_keys, num_items, begin_bit, end_bit, false, stream);
  }

  template <typename KeyT, typename NumItemsT>
  static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                       KeyT* d_keys_out, NumItemsT num_items, int begin_bit = 0,
                                       int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::radix<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in),
                                                  d_keys_out, nullptr, nullptr, sel, uint64_t(num_items), begin_bit,
                                                  end_bit, true, false, stream);
  }

  template <typename KeyT, typename NumItemsT>
  static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                       NumItemsT num_items, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                                       hipStream_t stream = nullptr)
  {
    return db_keys(d_temp_storage, temp_storage_bytes, d_keys, num_items, begin_bit, end_bit, true, stream);
  }

private:
  template <typename KeyT, typename ValueT, typename NumItemsT>
  static hipError_t db_pairs(void* t, size_t& b, DoubleBuffer<KeyT>& k, DoubleBuffer<ValueT>& v, NumItemsT n, int bb,
                             int eb, bool desc, hipStream_t s)
  {
    // the C entry takes one selector for both buffers, so hand it the values' pair in
    // the keys' orientation
    ValueT* v0  = k.selector == v.selector ? v.d_buffers[0] : v.d_buffers[1];
    ValueT* v1  = k.selector == v.selector ? v.d_buffers[1] : v.d_buffers[0];
    int sel     = k.selector;
    hipError_t e = detail::radix(t, b, k.d_buffers[0], k.d_buffers[1], v0, v1, sel, uint64_t(n), bb, eb, desc, true, s);
    if (e == hipSuccess && t != nullptr)
    {
      v.selector = (v.selector + (sel - k.selector) + 2) & 1;
      k.selector = sel;
    }
    return e;
  }

  template <typename KeyT, typename NumItemsT>
  static hipError_t db_keys(void* t, size_t& b, DoubleBuffer<KeyT>& k, NumItemsT n, int bb, int eb, bool desc,
                            hipStream_t s)
  {
    int sel      = k.selector;
    hipError_t e = detail::radix<KeyT, detail::NullValue>(t, b, k.d_buffers[0], k.d_buffers[1], nullptr, nullptr, sel,
                                                          uint64_t(n), bb, eb, desc, true, s);
    if (e == hipSuccess && t != nullptr)
    {
      k.selector = sel;
    }
    return e;
  }
};

// cub::DeviceSegmentedRadixSort: one stable sort per [d_begin_offsets[i], d_end_offsets[i]).
struct DeviceSegmentedRadixSort
{
  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                              KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out, int64_t num_items,
                              int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,
                              EndOffsetIteratorT d_end_offsets, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                              hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::seg_radix(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), d_keys_out,
                             const_cast<ValueT*>(d_values_in), d_values_out, sel, num_items, num_segments,
                             d_begin_offsets, d_end_offsets, begin_bit, end_bit, false, false, stream);
  }

  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                        KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out,
                                        int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,
                                        EndOffsetIteratorT d_end_offsets, int begin_bit = 0,
                                        int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::seg_radix(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), d_keys_out,
                             const_cast<ValueT*>(d_values_in), d_values_out, sel, num_items, num_segments,
                             d_begin_offsets, d_end_offsets, begin_bit, end_bit, true, false, stream);
  }

  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,
                             int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,
                             EndOffsetIteratorT d_end_offsets, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                             hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::seg_radix<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in),
                                                      d_keys_out, nullptr, nullptr, sel, num_items, num_segments,
                                                      d_begin_offsets, d_end_offsets, begin_bit, end_bit, false, false,
                                                      stream);
  }

  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                                       KeyT* d_keys_out, int64_t num_items, int64_t num_segments,
                                       BeginOffsetIteratorT d_begin_offsets, EndOffsetIteratorT d_end_offsets,
                                       int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    int sel = 0;
    return detail::seg_radix<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in),
                                                      d_keys_out, nullptr, nullptr, sel, num_items, num_segments,
                                                      d_begin_offsets, d_end_offsets, begin_bit, end_bit, true, false,
                                                      stream);
  }

  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortPairs(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                              DoubleBuffer<ValueT>& d_values, int64_t num_items, int64_t num_segments,
                              BeginOffsetIteratorT d_begin_offsets, EndOffsetIteratorT d_end_offsets,
                              int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    return db(d_temp_storage, temp_storage_bytes, d_keys, &d_values, num_items, num_segments, d_begin_offsets,
              d_end_offsets, begin_bit, end_bit, false, stream);
  }

  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortPairsDescending(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                        DoubleBuffer<ValueT>& d_values, int64_t num_items, int64_t num_segments,
                                        BeginOffsetIteratorT d_begin_offsets, EndOffsetIteratorT d_end_offsets,
                                        int begin_bit = 0, int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    return db(d_temp_storage, temp_storage_bytes, d_keys, &d_values, num_items, num_segments, d_begin_offsets,
              d_end_offsets, begin_bit, end_bit, true, stream);
  }

  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortKeys(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                             int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,
                             EndOffsetIteratorT d_end_offsets, int begin_bit = 0, int end_bit = sizeof(KeyT) * 8,
                             hipStream_t stream = nullptr)
  {
    return db(d_temp_storage, temp_storage_bytes, d_keys, static_cast<DoubleBuffer<detail::NullValue>*>(nullptr),
              num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit, end_bit, false, stream);
  }

  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>
  static hipError_t SortKeysDescending(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,
                                       int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,
                                       EndOffsetIteratorT d_end_offsets, int begin_bit = 0,
                                       int end_bit = sizeof(KeyT) * 8, hipStream_t stream = nullptr)
  {
    return db(d_temp_storage, temp_storage_bytes, d_keys, static_cast<DoubleBuffer<detail::NullValue>*>(nullptr),
              num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit, end_bit, true, stream);
  }

private:
  template <typename KeyT, typename ValueT, typename B, typename E>
  static hipError_t db(void* t, size_t& b, DoubleBuffer<KeyT>& k, DoubleBuffer<ValueT>* v, int64_t n, int64_t nseg,
                       B beg, E end, int bb, int eb, bool desc, hipStream_t s)
  {
    ValueT* v0 = nullptr;
    ValueT* v1 = nullptr;
    if (v != nullptr)
    {
      v0 = k.selector == v->selector ? v->d_buffers[0] : v->d_buffers[1];
      v1 = k.selector == v->selector ? v->d_buffers[1] : v->d_buffers[0];
    }
    int sel      = k.selector;
    hipError_t e = detail::seg_radix(t, b, k.d_buffers[0], k.d_buffers[1], v0, v1, sel, n, nseg, beg, end, bb, eb, desc,
                                     true, s);
    if (e == hipSuccess && t != nullptr)
    {
      if (v != nullptr)
      {
        v->selector = (v->selector + (sel - k.selector) + 2) & 1;
      }
      k.selector = sel;
    }
    return e;
  }
};

// cub::DeviceSegmentedSort: every segment sorted whole by key.  The sort is stable, so
// SortX and StableSortX give the same (stable) result.
struct DeviceSegmentedSort
{
#define CCCL_AMD_SEGSORT_POINTER_PAIRS(NAME, DESC)                                                                     \
  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>                \
  static hipError_t NAME(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,    \
                         const ValueT* d_values_in, ValueT* d_values_out, int64_t num_items, int64_t num_segments,     \
                         BeginOffsetIteratorT d_begin_offsets, EndOffsetIteratorT d_end_offsets,                       \
                         hipStream_t stream = nullptr)                                                                 \
  {                                                                                                                    \
    int sel = 0;                                                                                                       \
    return detail::seg_sort(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), d_keys_out,              \
                            const_cast<ValueT*>(d_values_in), d_values_out, sel, num_items, num_segments,              \
                            d_begin_offsets, d_end_offsets, DESC, false, stream);                                      \
  }
#define CCCL_AMD_SEGSORT_POINTER_KEYS(NAME, DESC)                                                                      \
  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>                                 \
  static hipError_t NAME(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,    \
                         int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,                \
                         EndOffsetIteratorT d_end_offsets, hipStream_t stream = nullptr)                               \
  {                                                                                                                    \
    int sel = 0;                                                                                                       \
    return detail::seg_sort<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, const_cast<KeyT*>(d_keys_in), \
                                                     d_keys_out, nullptr, nullptr, sel, num_items, num_segments,       \
                                                     d_begin_offsets, d_end_offsets, DESC, false, stream);             \
  }
#define CCCL_AMD_SEGSORT_DB_PAIRS(NAME, DESC)                                                                          \
  template <typename KeyT, typename ValueT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>                \
  static hipError_t NAME(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,                 \
                         DoubleBuffer<ValueT>& d_values, int64_t num_items, int64_t num_segments,                      \
                         BeginOffsetIteratorT d_begin_offsets, EndOffsetIteratorT d_end_offsets,                       \
                         hipStream_t stream = nullptr)                                                                 \
  {                                                                                                                    \
    return db(d_temp_storage, temp_storage_bytes, d_keys, &d_values, num_items, num_segments, d_begin_offsets,         \
              d_end_offsets, DESC, stream);                                                                            \
  }
#define CCCL_AMD_SEGSORT_DB_KEYS(NAME, DESC)                                                                           \
  template <typename KeyT, typename BeginOffsetIteratorT, typename EndOffsetIteratorT>                                 \
  static hipError_t NAME(void* d_temp_storage, size_t& temp_storage_bytes, DoubleBuffer<KeyT>& d_keys,                 \
                         int64_t num_items, int64_t num_segments, BeginOffsetIteratorT d_begin_offsets,                \
                         EndOffsetIteratorT d_end_offsets, hipStream_t stream = nullptr)                               \
  {                                                                                                                    \
    return db(d_temp_storage, temp_storage_bytes, d_keys, static_cast<DoubleBuffer<detail::NullValue>*>(nullptr),      \
              num_items, num_segments, d_begin_offsets, d_end_offsets, DESC, stream);                                  \
  }

  CCCL_AMD_SEGSORT_POINTER_PAIRS(SortPairs, false)
  CCCL_AMD_SEGSORT_POINTER_PAIRS(SortPairsDescending, true)
  CCCL_AMD_SEGSORT_POINTER_PAIRS(StableSortPairs, false)
  CCCL_AMD_SEGSORT_POINTER_PAIRS(StableSortPairsDescending, true)
  CCCL_AMD_SEGSORT_POINTER_KEYS(SortKeys, false)
  CCCL_AMD_SEGSORT_POINTER_KEYS(SortKeysDescending, true)
  CCCL_AMD_SEGSORT_POINTER_KEYS(StableSortKeys, false)
  CCCL_AMD_SEGSORT_POINTER_KEYS(StableSortKeysDescending, true)
  CCCL_AMD_SEGSORT_DB_PAIRS(SortPairs, false)
  CCCL_AMD_SEGSORT_DB_PAIRS(SortPairsDescending, true)
  CCCL_AMD_SEGSORT_DB_PAIRS(StableSortPairs, false)
  CCCL_AMD_SEGSORT_DB_PAIRS(StableSortPairsDescending, true)
  CCCL_AMD_SEGSORT_DB_KEYS(SortKeys, false)
  CCCL_AMD_SEGSORT_DB_KEYS(SortKeysDescending, true)
  CCCL_AMD_SEGSORT_DB_KEYS(StableSortKeys, false)
  CCCL_AMD_SEGSORT_DB_KEYS(StableSortKeysDescending, true)
#undef CCCL_AMD_SEGSORT_POINTER_PAIRS
#undef CCCL_AMD_SEGSORT_POINTER_KEYS
#undef CCCL_AMD_SEGSORT_DB_PAIRS
#undef CCCL_AMD_SEGSORT_DB_KEYS

private:
  template <typename KeyT, typename ValueT, typename B, typename E>
  static hipError_t db(void* t, size_t& b, DoubleBuffer<KeyT>& k, DoubleBuffer<ValueT>* v, int64_t n, int64_t nseg,
                       B beg, E end, bool desc, hipStream_t s)
  {
    ValueT* v0 = nullptr;
    ValueT* v1 = nullptr;
    if (v != nullptr)
    {
      v0 = k.selector == v->selector ? v->d_buffers[0] : v->d_buffers[1];
      v1 = k.selector == v->selector ? v->d_buffers[1] : v->d_buffers[0];
    }
    int sel      = k.selector;
    hipError_t e = detail::seg_sort(t, b, k.d_buffers[0], k.d_buffers[1], v0, v1, sel, n, nseg, beg, end, desc, true, s);
    if (e == hipSuccess && t != nullptr)
    {
      if (v != nullptr)
      {
        v->selector = (v->selector + (sel - k.selector) + 2) & 1;
      }
      k.selector = sel;
    }
    return e;
  }
};

// cub::DeviceTopK: the k largest (Max*) or smallest (Min*) keys, unordered; k is capped
// to num_items; ties at the k-th key are broken arbitrarily.
struct DeviceTopK
{
  template <typename KeyT, typename ValueT, typename NumItemsT, typename NumOutItemsT>
  static hipError_t MaxPairs(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                             KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out, NumItemsT num_items,
                             NumOutItemsT k, hipStream_t stream = nullptr)
  {
    return run<KeyT, ValueT>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out,
                             uint64_t(num_items), uint64_t(k), true, stream);
  }

  template <typename KeyT, typename ValueT, typename NumItemsT, typename NumOutItemsT>
  static hipError_t MinPairs(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in,
                             KeyT* d_keys_out, const ValueT* d_values_in, ValueT* d_values_out, NumItemsT num_items,
                             NumOutItemsT k, hipStream_t stream = nullptr)
  {
    return run<KeyT, ValueT>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out,
                             uint64_t(num_items), uint64_t(k), false, stream);
  }

  template <typename KeyT, typename NumItemsT, typename NumOutItemsT>
  static hipError_t MaxKeys(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,
                            NumItemsT num_items, NumOutItemsT k, hipStream_t stream = nullptr)
  {
    return run<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, nullptr, nullptr,
                                         uint64_t(num_items), uint64_t(k), true, stream);
  }

  template <typename KeyT, typename NumItemsT, typename NumOutItemsT>
  static hipError_t MinKeys(void* d_temp_storage, size_t& temp_storage_bytes, const KeyT* d_keys_in, KeyT* d_keys_out,
                            NumItemsT num_items, NumOutItemsT k, hipStream_t stream = nullptr)
  {
    return run<KeyT, detail::NullValue>(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, nullptr, nullptr,
                                         uint64_t(num_items), uint64_t(k), false, stream);
  }

private:
  template <typename KeyT, typename ValueT>
  static hipError_t run(void* t, size_t& b, const KeyT* ki, KeyT* ko, const ValueT* vi, ValueT* vo, uint64_t n,
                        uint64_t k, bool select_max, hipStream_t s)
  {
    return hipError_t(cccl_amd_topk(t, &b, ki, ko, vi, vo, detail::key_type<KeyT>(), detail::value_bytes<ValueT>(), n,
                                    k, select_max ? 1 : 0, detail::as_cu(s)));
  }
};

} // namespace cccl_amd
