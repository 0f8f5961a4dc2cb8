// Drives the C++ API of include/cccl_amd/device_radix_sort.hpp the way CUB's own tests
// drive cub::DeviceRadixSort & co. (cub/test/catch2_test_device_radix_sort_pairs.cu,
// catch2_test_device_segmented_radix_sort_keys.cu, catch2_test_device_segmented_sort_*,
// catch2_test_device_topk_*): two-phase temp storage, pointer and DoubleBuffer forms,
// checked against std::stable_sort on the host (keys compared as the radix order:
// -0.0 == +0.0, bit windows on the order image).
//
//   test_device_api --query-only   size queries only (no device needed)
//   test_device_api                every case on the current device; exit status 0 = all passed
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <cccl_amd/device_radix_sort.hpp>
#include <cccl_amd/thrust_sort.hpp>

namespace {

int g_failures = 0;

void report(const char* name, bool ok)
{
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
  std::fflush(stdout);
  g_failures += ok ? 0 : 1;
}

#define HIP_CHECK(x)                                                                       \
  do                                                                                       \
  {                                                                                        \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
    {                                                                                      \
      std::printf("HIP error %d (%s) at %s:%d\n", int(e_), hipGetErrorString(e_), __FILE__, __LINE__); \
      std::exit(2);                                                                        \
    }                                                                                      \
  } while (0)

uint64_t splitmix(uint64_t& s)
{
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z          = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z          = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Order image of a key as an unsigned 64-bit value (unsigned order = key order, the two
// zeros equal), restricted to bits [b, e).
template <typename K>
uint64_t order_image(K k, int b, int e)
{
  constexpr int bits = int(sizeof(K) * 8);
  uint64_t u         = 0;
  std::memcpy(&u, &k, sizeof(K));
  const uint64_t high = uint64_t(1) << (bits - 1);
  const uint64_t mask = bits == 64 ? ~uint64_t(0) : ((uint64_t(1) << bits) - 1);
  if constexpr (std::is_floating_point_v<K>)
  {
    if (u == high)
    {
      u = 0; // -0.0 sorts as +0.0
    }
    u = (u & high) ? (~u & mask) : (u | high);
  }
  else if constexpr (std::is_signed_v<K>)
  {
    u ^= high;
  }
  const int w = e - b;
  return w >= 64 ? (u >> b) : ((u >> b) & ((uint64_t(1) << w) - 1));
}

template <typename T>
struct Dev
{
  T* p = nullptr;
  explicit Dev(size_t n)
  {
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  }
  Dev(const std::vector<T>& h)
      : Dev(h.size())
  {
    HIP_CHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  ~Dev()
  {
    (void) hipFree(p);
  }
  std::vector<T> host(size_t n) const
  {
    std::vector<T> h(n);
    HIP_CHECK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

struct Temp
{
  void* p = nullptr;
  explicit Temp(size_t n)
  {
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1)));
  }
  ~Temp()
  {
    (void) hipFree(p);
  }
};

// Expected permutation of one range: stable by the (windowed) order image.
template <typename K>
void stable_perm(const std::vector<K>& keys, size_t b, size_t e, bool desc, int bb, int eb, std::vector<uint32_t>& perm)
{
  std::stable_sort(perm.begin() + b, perm.begin() + e, [&](uint32_t x, uint32_t y) {
    const uint64_t ox = order_image(keys[x], bb, eb), oy = order_image(keys[y], bb, eb);
    return desc ? ox > oy : ox < oy;
  });
}

template <typename K>
bool same_bits(const std::vector<K>& a, const std::vector<K>& b)
{
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(K)) == 0;
}

template <typename K>
std::vector<K> random_keys(size_t n, uint64_t seed)
{
  std::vector<K> k(n);
  for (auto& x : k)
  {
    const uint64_t r = splitmix(seed);
    if constexpr (std::is_floating_point_v<K>)
    {
      x = K((int64_t(r % 2001) - 1000) / 8.0); // many ties, both signs
      if (r % 97 == 0)
      {
        x = (r & 256) ? K(-0.0) : K(0.0);
      }
    }
    else
    {
      std::memcpy(&x, &r, sizeof(K));
      if (r % 3 == 0)
      {
        x = K(x & K(0x0f)); // ties
      }
    }
  }
  return k;
}

void radix_pairs_u32()
{
  const size_t n = (1 << 20) + 123;
  auto keys      = random_keys<uint32_t>(n, 1);
  std::vector<uint32_t> vals(n);
  std::iota(vals.begin(), vals.end(), 0u);
  Dev<uint32_t> kin(keys), kout(n), vin(vals), vout(n);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortPairs(nullptr, bytes, kin.p, kout.p, vin.p, vout.p, n));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortPairs(t.p, bytes, kin.p, kout.p, vin.p, vout.p, n));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  stable_perm(keys, 0, n, false, 0, 32, perm);
  std::vector<uint32_t> ek(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
  }
  report("DeviceRadixSort::SortPairs u32 pointer", same_bits(kout.host(n), ek) && vout.host(n) == perm
                                                     && kin.host(n) == keys);
}

void radix_pairs_desc_f32_doublebuffer()
{
  const size_t n = 300000;
  auto keys      = random_keys<float>(n, 2);
  std::vector<uint64_t> vals(n);
  std::iota(vals.begin(), vals.end(), uint64_t(0));
  Dev<float> k0(keys), k1(n);
  Dev<uint64_t> v0(vals), v1(n);
  cccl_amd::DoubleBuffer<float> dk(k0.p, k1.p);
  cccl_amd::DoubleBuffer<uint64_t> dv(v0.p, v1.p);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortPairsDescending(nullptr, bytes, dk, dv, n));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortPairsDescending(t.p, bytes, dk, dv, n));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  stable_perm(keys, 0, n, true, 0, 32, perm);
  std::vector<float> ek(n);
  std::vector<uint64_t> ev(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
    ev[i] = perm[i];
  }
  const bool sel_ok = dk.selector == 0 && dv.selector == 0; // 4 passes: back in the first buffer
  auto hk           = std::vector<float>(n);
  HIP_CHECK(hipMemcpy(hk.data(), dk.Current(), n * sizeof(float), hipMemcpyDeviceToHost));
  auto hv = std::vector<uint64_t>(n);
  HIP_CHECK(hipMemcpy(hv.data(), dv.Current(), n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  report("DeviceRadixSort::SortPairsDescending f32 DoubleBuffer", sel_ok && same_bits(hk, ek) && hv == ev);
}

void radix_keys_i64_window()
{
  const size_t n = 200000;
  const int b = 8, e = 40;
  auto keys = random_keys<int64_t>(n, 3);
  Dev<int64_t> kin(keys), kout(n);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortKeys(nullptr, bytes, kin.p, kout.p, n, b, e));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceRadixSort::SortKeys(t.p, bytes, kin.p, kout.p, n, b, e));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  stable_perm(keys, 0, n, false, b, e, perm);
  std::vector<int64_t> ek(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
  }
  report("DeviceRadixSort::SortKeys i64 bits [8,40)", same_bits(kout.host(n), ek));
}

template <typename K, typename Off>
void segmented_case(const char* name, bool segsort, bool desc, bool doublebuffer, size_t nseg, uint64_t seed)
{
  std::vector<Off> beg(nseg), end(nseg);
  uint64_t s = seed;
  size_t n   = 0;
  for (size_t i = 0; i < nseg; ++i)
  {
    const uint64_t r = splitmix(s);
    size_t sz        = (r % 10 == 0) ? 20000 + r % 40000 : r % 700; // a few large segments, some empty
    beg[i]           = Off(n);
    n += sz;
    end[i] = Off(n);
  }
  auto keys = random_keys<K>(n, seed + 1);
  std::vector<uint32_t> vals(n);
  std::iota(vals.begin(), vals.end(), 0u);
  Dev<K> k0(keys), k1(n);
  Dev<uint32_t> v0(vals), v1(n);
  Dev<Off> db(beg), de(end);
  size_t bytes = 0;
  cccl_amd::DoubleBuffer<K> dk(k0.p, k1.p);
  cccl_amd::DoubleBuffer<uint32_t> dv(v0.p, v1.p);
  auto call = [&](void* t) {
    if (segsort)
    {
      if (doublebuffer)
      {
        return desc ? cccl_amd::DeviceSegmentedSort::StableSortPairsDescending(t, bytes, dk, dv, n, nseg, db.p, de.p)
                    : cccl_amd::DeviceSegmentedSort::StableSortPairs(t, bytes, dk, dv, n, nseg, db.p, de.p);
      }
      return desc ? cccl_amd::DeviceSegmentedSort::SortPairsDescending(t, bytes, k0.p, k1.p, v0.p, v1.p, n, nseg, db.p,
                                                                         de.p)
                  : cccl_amd::DeviceSegmentedSort::SortPairs(t, bytes, k0.p, k1.p, v0.p, v1.p, n, nseg, db.p, de.p);
    }
    if (doublebuffer)
    {
      return desc ? cccl_amd::DeviceSegmentedRadixSort::SortPairsDescending(t, bytes, dk, dv, n, nseg, db.p, de.p)
                  : cccl_amd::DeviceSegmentedRadixSort::SortPairs(t, bytes, dk, dv, n, nseg, db.p, de.p);
    }
    return desc ? cccl_amd::DeviceSegmentedRadixSort::SortPairsDescending(t, bytes, k0.p, k1.p, v0.p, v1.p, n, nseg,
                                                                            db.p, de.p)
                : cccl_amd::DeviceSegmentedRadixSort::SortPairs(t, bytes, k0.p, k1.p, v0.p, v1.p, n, nseg, db.p, de.p);
  };
  HIP_CHECK(call(nullptr));
  Temp t(bytes);
  HIP_CHECK(call(t.p));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  for (size_t i = 0; i < nseg; ++i)
  {
    stable_perm(keys, size_t(beg[i]), size_t(end[i]), desc, 0, int(sizeof(K) * 8), perm);
  }
  std::vector<K> ek(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
  }
  K* rk        = doublebuffer ? dk.Current() : k1.p;
  uint32_t* rv = doublebuffer ? dv.Current() : v1.p;
  bool sel_ok  = true;
  if (doublebuffer && segsort)
  {
    const int bits  = int(sizeof(K) * 8);
    const int p_cub = (bits + (sizeof(K) > 1 ? 6 : 4) - 1) / (sizeof(K) > 1 ? 6 : 4);
    sel_ok          = dk.selector == (p_cub & 1) && dv.selector == dk.selector;
  }
  std::vector<K> hk(n);
  std::vector<uint32_t> hv(n);
  HIP_CHECK(hipMemcpy(hk.data(), rk, n * sizeof(K), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(hv.data(), rv, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  report(name, sel_ok && same_bits(hk, ek) && hv == perm);
}

void topk_case()
{
  const size_t n = 1 << 20;
  const size_t k = 5000;
  auto keys      = random_keys<int32_t>(n, 9);
  std::vector<uint32_t> vals(n);
  std::iota(vals.begin(), vals.end(), 0u);
  Dev<int32_t> kin(keys), kout(k);
  Dev<uint32_t> vin(vals), vout(k);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceTopK::MaxPairs(nullptr, bytes, kin.p, kout.p, vin.p, vout.p, n, k));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceTopK::MaxPairs(t.p, bytes, kin.p, kout.p, vin.p, vout.p, n, k));
  auto hk = kout.host(k);
  auto hv = vout.host(k);
  std::vector<int32_t> sorted = keys;
  std::sort(sorted.begin(), sorted.end(), std::greater<int32_t>());
  std::vector<int32_t> expect(sorted.begin(), sorted.begin() + k), got = hk;
  std::sort(got.begin(), got.end(), std::greater<int32_t>());
  bool ok = got == expect;
  std::vector<uint32_t> idx = hv;
  std::sort(idx.begin(), idx.end());
  ok &= std::adjacent_find(idx.begin(), idx.end()) == idx.end();
  for (size_t i = 0; i < k && ok; ++i)
  {
    ok &= hv[i] < n && keys[hv[i]] == hk[i];
  }
  report("DeviceTopK::MaxPairs i32", ok);
}

// thrust::sort / stable_sort_by_key through the radix route (in place, copy back when the
// result ends in the temporary half: one-byte keys take one pass).
template <typename K, typename Compare>
void thrust_keys_case(const char* name, size_t n, uint64_t seed, Compare comp)
{
  auto keys = random_keys<K>(n, seed);
  Dev<K> d(keys);
  HIP_CHECK(cccl_amd::thrust_compat::sort(d.p, d.p + n, comp));
  HIP_CHECK(hipDeviceSynchronize());
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  stable_perm(keys, 0, n, std::is_same_v<Compare, std::greater<K>> || std::is_same_v<Compare, std::greater<>>, 0,
              int(sizeof(K) * 8), perm);
  std::vector<K> ek(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
  }
  report(name, same_bits(d.host(n), ek));
}

void thrust_pairs_case()
{
  const size_t n = 250000;
  auto keys      = random_keys<double>(n, 21);
  std::vector<uint64_t> vals(n);
  std::iota(vals.begin(), vals.end(), uint64_t(0));
  Dev<double> dk(keys);
  Dev<uint64_t> dv(vals);
  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreate(&s));
  HIP_CHECK(cccl_amd::thrust_compat::stable_sort_by_key(dk.p, dk.p + n, dv.p, std::greater<double>(), s));
  HIP_CHECK(hipStreamSynchronize(s));
  HIP_CHECK(hipStreamDestroy(s));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  stable_perm(keys, 0, n, true, 0, 64, perm);
  std::vector<double> ek(n);
  std::vector<uint64_t> ev(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
    ev[i] = perm[i];
  }
  report("thrust_compat::stable_sort_by_key f64 greater, own stream", same_bits(dk.host(n), ek) && dv.host(n) == ev);
}

void segmented_keys_doublebuffer_case()
{
  const size_t nseg = 150;
  std::vector<int64_t> beg(nseg), end(nseg);
  uint64_t sd = 31;
  size_t n    = 0;
  for (size_t i = 0; i < nseg; ++i)
  {
    const uint64_t r = splitmix(sd);
    beg[i]           = int64_t(n);
    n += (r % 7 == 0) ? 30000 + r % 20000 : r % 900;
    end[i] = int64_t(n);
  }
  auto keys = random_keys<int16_t>(n, 32);
  Dev<int16_t> k0(keys), k1(n);
  Dev<int64_t> db(beg), de(end);
  cccl_amd::DoubleBuffer<int16_t> dk(k0.p, k1.p);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceSegmentedRadixSort::SortKeysDescending(nullptr, bytes, dk, n, nseg, db.p, de.p, 3, 14));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceSegmentedRadixSort::SortKeysDescending(t.p, bytes, dk, n, nseg, db.p, de.p, 3, 14));
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  for (size_t i = 0; i < nseg; ++i)
  {
    stable_perm(keys, size_t(beg[i]), size_t(end[i]), true, 3, 14, perm);
  }
  std::vector<int16_t> ek(n), hk(n);
  for (size_t i = 0; i < n; ++i)
  {
    ek[i] = keys[perm[i]];
  }
  HIP_CHECK(hipMemcpy(hk.data(), dk.Current(), n * sizeof(int16_t), hipMemcpyDeviceToHost));
  report("DeviceSegmentedRadixSort::SortKeysDescending i16 DoubleBuffer bits [3,14)", same_bits(hk, ek));
}

void topk_min_keys_case()
{
  const size_t n = 3 << 20, k = 100000;
  auto keys      = random_keys<double>(n, 33);
  Dev<double> kin(keys), kout(k);
  size_t bytes = 0;
  HIP_CHECK(cccl_amd::DeviceTopK::MinKeys(nullptr, bytes, kin.p, kout.p, n, k));
  Temp t(bytes);
  HIP_CHECK(cccl_amd::DeviceTopK::MinKeys(t.p, bytes, kin.p, kout.p, n, k));
  auto got = kout.host(k);
  // the k smallest by the radix image (+0.0 above -0.0), as a multiset of images
  std::vector<uint64_t> img(n), out(k);
  auto image = [](double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | (uint64_t(1) << 63));
  };
  for (size_t i = 0; i < n; ++i)
  {
    img[i] = image(keys[i]);
  }
  for (size_t i = 0; i < k; ++i)
  {
    out[i] = image(got[i]);
  }
  std::sort(img.begin(), img.end());
  std::sort(out.begin(), out.end());
  report("DeviceTopK::MinKeys f64 (signed zeros, ties)", std::equal(out.begin(), out.end(), img.begin()));
}

bool query_only()
{
  size_t b1 = 0, b2 = 0, b3 = 0, b4 = 0;
  const uint32_t* k = nullptr;
  uint32_t* ko      = nullptr;
  const int32_t* off = nullptr;
  bool ok = cccl_amd::DeviceRadixSort::SortKeys(nullptr, b1, k, ko, 1 << 20) == hipSuccess && b1 > 0;
  ok &= cccl_amd::DeviceSegmentedRadixSort::SortKeys(nullptr, b2, k, ko, 1 << 20, 100, off, off) == hipSuccess && b2 > 0;
  ok &= cccl_amd::DeviceSegmentedSort::StableSortKeys(nullptr, b3, k, ko, 1 << 20, 100, off, off) == hipSuccess && b3 > 0;
  ok &= cccl_amd::DeviceTopK::MinKeys(nullptr, b4, k, ko, 1 << 20, 10) == hipSuccess && b4 > 0;
  std::printf("query bytes: radix %zu segmented-radix %zu segmented-sort %zu topk %zu\n", b1, b2, b3, b4);
  return ok;
}

} // namespace

int main(int argc, char** argv)
{
  const bool queries = argc > 1 && std::string(argv[1]) == "--query-only";
  report("temp-storage queries", query_only());
  if (!queries)
  {
    radix_pairs_u32();
    radix_pairs_desc_f32_doublebuffer();
    radix_keys_i64_window();
    segmented_case<uint16_t, int32_t>("DeviceSegmentedRadixSort::SortPairs u16 pointer", false, false, false, 300, 11);
    segmented_case<double, int64_t>("DeviceSegmentedRadixSort::SortPairsDescending f64 DoubleBuffer", false, true, true,
                                    200, 12);
    segmented_case<int32_t, int32_t>("DeviceSegmentedSort::SortPairs i32 pointer", true, false, false, 400, 13);
    segmented_case<uint8_t, int64_t>("DeviceSegmentedSort::StableSortPairs u8 DoubleBuffer", true, false, true, 250, 14);
    segmented_case<uint64_t, int32_t>("DeviceSegmentedSort::StableSortPairsDescending u64 DoubleBuffer", true, true,
                                      true, 250, 15);
    segmented_case<float, int32_t>("DeviceSegmentedSort::StableSortPairs f32 DoubleBuffer", true, false, true, 250, 16);
    topk_case();
    thrust_keys_case<uint32_t>("thrust_compat::sort u32 less", (1 << 20) + 5, 22, std::less<uint32_t>());
    thrust_keys_case<uint8_t>("thrust_compat::sort u8 greater<> (copy back)", 100000, 23, std::greater<>());
    thrust_pairs_case();
    segmented_keys_doublebuffer_case();
    topk_min_keys_case();
  }
  std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
  return g_failures ? 1 : 0;
}
