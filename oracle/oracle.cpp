// CPU oracle of the radix sort.  TEST INFRASTRUCTURE ONLY: imported by tests/, by the
// smoke check and by the benchmark's CPU-baseline leg as the checker; no product code
// path links or calls it.
//
// The sort is the closed form of cub/test/catch2_radix_sort_helper.cuh:174-312: each key
// maps to its "striped" bit-ordered form (get_striped_keys, :174-212: -0.0 -> +0.0 for
// floating point, then the order-preserving transform of cub/cub/util_type.cuh:846-965,
// then the [begin_bit, end_bit) window mask), and the output is the stable order of the
// items by that form, ascending with `<` or descending with `>` (:214-268), keys and
// values gathered through the permutation (:270-312).  Stability makes the permutation
// unique, so it is computed here as the sort of (striped key, input index) pairs.
//
// Parity pins (tests/test_oracle.py): the reference's own Thrust OpenMP
// stable_sort_by_key built from /root/reference (oracle/_ref), brute-force
// permutation enumeration on tiny inputs, std::stable_sort with an independent
// comparator, the radix-operation special cases of
// cub/test/catch2_test_radix_operations.cu:582-603 (+-0 share a digit), and
// closed-form histogram invariants.
#include <parallel/algorithm>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

enum Kind
{
  kUnsigned = 0,
  kSigned   = 1,
  kFloat    = 2
};

// Key of key_bytes at p as a zero-extended unsigned 64-bit pattern.
inline uint64_t load_bits(const void* p, uint64_t i, int key_bytes)
{
  switch (key_bytes)
  {
    case 1:
      return static_cast<const uint8_t*>(p)[i];
    case 2:
      return static_cast<const uint16_t*>(p)[i];
    case 4:
      return static_cast<const uint32_t*>(p)[i];
    default:
      return static_cast<const uint64_t*>(p)[i];
  }
}

inline uint64_t width_mask(int bits)
{
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
}

// cub::Traits<T>::TwiddleIn on a bits-wide pattern (util_type.cuh:846-965), after the
// -0.0 -> +0.0 replacement of get_striped_keys.
inline uint64_t bit_ordered(uint64_t k, int key_bytes, int kind)
{
  const int bits      = key_bytes * 8;
  const uint64_t high = 1ull << (bits - 1);
  const uint64_t all  = width_mask(bits);
  if (kind == kFloat)
  {
    if (k == high)
    {
      k = 0;
    }
    return (k & high) ? (~k & all) : (k ^ high);
  }
  if (kind == kSigned)
  {
    return k ^ high;
  }
  return k;
}

inline uint64_t striped(uint64_t k, int key_bytes, int kind, int begin_bit, int end_bit)
{
  uint64_t s = bit_ordered(k, key_bytes, kind);
  if (begin_bit > 0 || end_bit < key_bytes * 8)
  {
    s &= width_mask(end_bit - begin_bit) << begin_bit;
  }
  return s;
}

struct Item
{
  uint64_t key;
  uint64_t idx;
};

void sort_range(std::vector<Item>& v, size_t b, size_t e, bool descending, bool parallel)
{
  auto asc  = [](const Item& x, const Item& y) { return x.key < y.key || (x.key == y.key && x.idx < y.idx); };
  auto desc = [](const Item& x, const Item& y) { return x.key > y.key || (x.key == y.key && x.idx < y.idx); };
  if (parallel && e - b > (1u << 16))
  {
    if (descending)
    {
      __gnu_parallel::sort(v.begin() + b, v.begin() + e, desc);
    }
    else
    {
      __gnu_parallel::sort(v.begin() + b, v.begin() + e, asc);
    }
  }
  else
  {
    if (descending)
    {
      std::sort(v.begin() + b, v.begin() + e, desc);
    }
    else
    {
      std::sort(v.begin() + b, v.begin() + e, asc);
    }
  }
}

void gather(const void* src, void* dst, const std::vector<Item>& v, size_t b, size_t e, int bytes)
{
  const auto* s = static_cast<const unsigned char*>(src);
  auto* d       = static_cast<unsigned char*>(dst);
  for (size_t j = b; j < e; ++j)
  {
    std::memcpy(d + j * bytes, s + v[j].idx * bytes, bytes);
  }
}

bool args_ok(int key_bytes, int kind, int begin_bit, int end_bit)
{
  return (key_bytes == 1 || key_bytes == 2 || key_bytes == 4 || key_bytes == 8) && kind >= 0 && kind <= 2
      && begin_bit >= 0 && begin_bit <= end_bit && end_bit <= key_bytes * 8;
}

} // namespace

extern "C" {

// Striped (bit-ordered, window-masked) form of each key.
int oracle_striped_keys(
  const void* keys, uint64_t n, int key_bytes, int kind, int begin_bit, int end_bit, uint64_t* out)
{
  if (!args_ok(key_bytes, kind, begin_bit, end_bit))
  {
    return 1;
  }
  for (uint64_t i = 0; i < n; ++i)
  {
    out[i] = striped(load_bits(keys, i, key_bytes), key_bytes, kind, begin_bit, end_bit);
  }
  return 0;
}

// Stable sort permutation: perm[j] = input index of output item j.
int oracle_sort_permutation(
  const void* keys, uint64_t n, int key_bytes, int kind, int descending, int begin_bit, int end_bit, uint64_t* perm)
{
  if (!args_ok(key_bytes, kind, begin_bit, end_bit))
  {
    return 1;
  }
  std::vector<Item> v(n);
  for (uint64_t i = 0; i < n; ++i)
  {
    v[i] = {striped(load_bits(keys, i, key_bytes), key_bytes, kind, begin_bit, end_bit), i};
  }
  sort_range(v, 0, n, descending != 0, true);
  for (uint64_t j = 0; j < n; ++j)
  {
    perm[j] = v[j].idx;
  }
  return 0;
}

// Device-wide sort: keys_out/vals_out receive the sorted keys and values
// (vals == nullptr or val_bytes == 0 for keys only).
int oracle_sort(
  const void* keys,
  const void* vals,
  uint64_t n,
  int key_bytes,
  int kind,
  int val_bytes,
  int descending,
  int begin_bit,
  int end_bit,
  void* keys_out,
  void* vals_out)
{
  if (!args_ok(key_bytes, kind, begin_bit, end_bit))
  {
    return 1;
  }
  std::vector<Item> v(n);
  for (uint64_t i = 0; i < n; ++i)
  {
    v[i] = {striped(load_bits(keys, i, key_bytes), key_bytes, kind, begin_bit, end_bit), i};
  }
  sort_range(v, 0, n, descending != 0, true);
  gather(keys, keys_out, v, 0, n, key_bytes);
  if (vals != nullptr && val_bytes > 0)
  {
    gather(vals, vals_out, v, 0, n, val_bytes);
  }
  return 0;
}

// Segmented sort (catch2_radix_sort_helper.cuh:238-268): each [begin[s], end[s]) is
// sorted on its own; items outside every segment keep what keys_out/vals_out held.
int oracle_segmented_sort(
  const void* keys,
  const void* vals,
  uint64_t n,
  int key_bytes,
  int kind,
  int val_bytes,
  int descending,
  int begin_bit,
  int end_bit,
  int64_t num_segments,
  const int64_t* seg_begin,
  const int64_t* seg_end,
  void* keys_out,
  void* vals_out)
{
  if (!args_ok(key_bytes, kind, begin_bit, end_bit))
  {
    return 1;
  }
  std::vector<Item> v(n);
  for (uint64_t i = 0; i < n; ++i)
  {
    v[i] = {striped(load_bits(keys, i, key_bytes), key_bytes, kind, begin_bit, end_bit), i};
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < num_segments; ++s)
  {
    const int64_t b = seg_begin[s], e = seg_end[s];
    if (e <= b || b < 0 || uint64_t(e) > n)
    {
      continue;
    }
    sort_range(v, size_t(b), size_t(e), descending != 0, false);
    gather(keys, keys_out, v, size_t(b), size_t(e), key_bytes);
    if (vals != nullptr && val_bytes > 0)
    {
      gather(vals, vals_out, v, size_t(b), size_t(e), val_bytes);
    }
  }
  return 0;
}

// Digit histogram of every 8-bit pass of [begin_bit, end_bit) over the keys in the
// order the sort consumes them (bit-ordered, inverted for descending).
int oracle_histograms(
  const void* keys, uint64_t n, int key_bytes, int kind, int descending, int begin_bit, int end_bit, uint64_t* bins)
{
  if (!args_ok(key_bytes, kind, begin_bit, end_bit))
  {
    return 1;
  }
  const int passes = (end_bit - begin_bit + 7) / 8;
  std::memset(bins, 0, sizeof(uint64_t) * 256 * size_t(passes));
  const uint64_t all = width_mask(key_bytes * 8);
  for (uint64_t i = 0; i < n; ++i)
  {
    uint64_t o = bit_ordered(load_bits(keys, i, key_bytes), key_bytes, kind);
    if (descending)
    {
      o = ~o & all;
    }
    for (int p = 0; p < passes; ++p)
    {
      const int shift = begin_bit + 8 * p;
      const int bits  = std::min(8, end_bit - shift);
      ++bins[p * 256 + ((o >> shift) & width_mask(bits))];
    }
  }
  return 0;
}

} // extern "C"
