// cub::DeviceTopK::{Max,Min}{Keys,Pairs} (cub/cub/device/device_topk.cuh:297, :406, :538,
// :666, :775, ...) as an MSB-first radix select over the twiddled key bits.
//
// Order space: o = twiddle(key) exactly as cub/cub/device/dispatch/dispatch_topk.cuh:57-82
// (unsigned: identity; signed: flip the sign bit; floating point: flip all bits of
// negatives, the sign bit of the rest -- no -0.0 / +0.0 merging, so +0.0 ranks above
// -0.0 and NaNs rank by bit pattern), then NOT for Max.  The call returns the k items of
// smallest o, in no particular order; among items whose o equals the k-th smallest,
// which ones are returned is unspecified (the reference's only supported mode,
// device_topk.cuh:86-92: determinism not guaranteed, output unsorted).
//
// Launch sequence (stream-ordered, no host synchronisation, no memset nodes):
//   topk_init    the selection state: prefix 0, k_rem = k, counters and bins 0
//   per 8-bit digit, most significant first:
//     topk_hist    bins of this digit over the keys whose higher digits equal the prefix
//                  (LDS bins per workgroup, one global atomic per non-zero bin)
//     topk_pick    one workgroup: the digit d whose bin holds the k_rem-th candidate;
//                  prefix |= d, k_rem -= candidates below d; when the bin is taken whole,
//                  the selection is complete and later passes return at once
//   topk_output  keys with o below the threshold take slots [0, k - k_rem) and keys equal
//                to it the first k_rem slots after that: each workgroup stages the items
//                it selects in LDS and flushes a batch with one global atomic (its base
//                slot) and coalesced stores.
// Each histogram pass reads every key once (K bytes per item), so a call moves at most
// (P + 1) * N * K + k * (K + V) bytes for P digit passes.
#include <algorithm>

#include "../../include/cccl_amd/topk.h"
#include "dispatch.hpp"
#include "radix_common.hpp"

namespace cccl_amd {

namespace {

constexpr int kTopkThreads = 256;
constexpr int kTopkIpt     = 16; // items per thread per loop iteration of the streaming kernels

struct TopkState
{
  uint64_t prefix; // threshold digits chosen so far (order space)
  uint64_t k_rem; // items still to take among the keys that match the prefix
  uint32_t done; // 1: the bin at done_shift is taken whole, the selection is complete
  uint32_t done_shift;
  uint32_t count_lt; // output slots taken by keys below the threshold
  uint32_t count_eq; // output slots taken by keys equal to the threshold
  uint32_t bins[kRadix];
};

template <class UK>
__device__ __forceinline__ UK topk_order(UK k, int key_kind, bool select_max)
{
  constexpr int kBits = int(sizeof(UK) * 8);
  constexpr UK kHigh  = UK(UK(1) << (kBits - 1));
  if (key_kind == kFloatKey)
  {
    k = (k & kHigh) ? UK(~k) : UK(k | kHigh);
  }
  else if (key_kind == kSignedKey)
  {
    k = UK(k ^ kHigh);
  }
  return select_max ? UK(~k) : k;
}

struct TopkInitParams
{
  TopkState* st;
  uint64_t k;
};

__global__ void __launch_bounds__(kTopkThreads) topk_init_kernel(TopkInitParams p)
{
  p.st->bins[threadIdx.x] = 0u;
  if (threadIdx.x == 0)
  {
    p.st->prefix     = 0;
    p.st->k_rem      = p.k;
    p.st->done       = 0u;
    p.st->done_shift = 0u;
    p.st->count_lt   = 0u;
    p.st->count_eq   = 0u;
  }
}

struct TopkPassParams
{
  const void* keys;
  uint64_t num_items;
  TopkState* st;
  int key_kind;
  int select_max;
  int shift; // lowest bit of this pass's digit
};

template <class UK>
__global__ void __launch_bounds__(kTopkThreads) topk_hist_kernel(TopkPassParams p)
{
  __shared__ uint32_t bins[kRadix];
  if (p.st->done != 0u)
  {
    return;
  }
  constexpr int kBits   = int(sizeof(UK) * 8);
  const int hi          = p.shift + kRadixBits;
  const UK prefix_hi    = hi >= kBits ? UK(0) : UK(UK(p.st->prefix) >> hi);
  const UK* keys        = static_cast<const UK*>(p.keys);
  const bool select_max = p.select_max != 0;
  bins[threadIdx.x]     = 0u;
  __syncthreads();
  const uint64_t chunk = uint64_t(kTopkThreads) * kTopkIpt;
  for (uint64_t base = uint64_t(blockIdx.x) * chunk; base < p.num_items; base += uint64_t(gridDim.x) * chunk)
  {
    UK k[kTopkIpt];
#pragma unroll
    for (int j = 0; j < kTopkIpt; ++j)
    {
      const uint64_t i = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      k[j]             = i < p.num_items ? keys[i] : UK(0);
    }
#pragma unroll
    for (int j = 0; j < kTopkIpt; ++j)
    {
      const uint64_t i = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      const UK o       = topk_order(k[j], p.key_kind, select_max);
      const bool match = i < p.num_items && (hi >= kBits || UK(o >> hi) == prefix_hi);
      if (match)
      {
        lds_count(bins, extract_digit(o, p.shift, kRadix - 1));
      }
    }
  }
  __syncthreads();
  const uint32_t c = bins[threadIdx.x];
  if (c != 0u)
  {
    __hip_atomic_fetch_add(&p.st->bins[threadIdx.x], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void __launch_bounds__(kTopkThreads) topk_pick_kernel(TopkPassParams p)
{
  __shared__ uint32_t s_wave[kTopkThreads / 64];
  TopkState* st = p.st;
  if (st->done != 0u)
  {
    return;
  }
  const uint32_t t     = threadIdx.x;
  const uint32_t c     = st->bins[t];
  const uint64_t k_rem = st->k_rem;
  uint32_t total       = 0;
  const uint32_t excl  = radix_exclusive_scan<uint32_t>(c, s_wave, total);
  __syncthreads(); // every thread has read k_rem and its bin before they are rewritten
  st->bins[t] = 0u; // clean for the next pass
  // The candidates below digit t number excl, so t holds the k_rem-th candidate when
  // excl < k_rem <= excl + c; exactly one digit satisfies that because the candidates
  // number at least k_rem.
  if (uint64_t(excl) < k_rem && k_rem <= uint64_t(excl) + c)
  {
    st->prefix |= uint64_t(t) << p.shift;
    st->k_rem = k_rem - excl;
    if (k_rem - excl == c)
    {
      st->done       = 1u;
      st->done_shift = uint32_t(p.shift);
    }
  }
}

template <class V>
__device__ __forceinline__ void copy_value(void* out, const void* in, uint64_t dst, uint64_t src)
{
  if constexpr (!std::is_same<V, NoValue>::value)
  {
    static_cast<V*>(out)[dst] = static_cast<const V*>(in)[src];
  }
}

// Slot of this lane among the lanes of m, after adding popcount(m) to the LDS counter
// *n with one atomic per wave (issued by the first lane in m).
__device__ __forceinline__ uint32_t wave_lds_slot(uint32_t* n, uint64_t m)
{
  const int leader = __builtin_ctzll(m);
  uint32_t base    = 0;
  if (int(lane_id()) == leader)
  {
    base = __hip_atomic_fetch_add(n, uint32_t(__builtin_popcountll(m)), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  base = __builtin_amdgcn_readlane(base, leader);
  return base + mask_prefix_count(m);
}

struct TopkOutputParams
{
  const void* keys_in;
  const void* vals_in;
  void* keys_out;
  void* vals_out;
  uint64_t num_items;
  uint64_t k;
  TopkState* st;
  int key_kind;
  int select_max;
};

// LDS staging of the output: each workgroup collects the items it selects in two LDS
// buffers of CAP items (below / equal to the threshold) and flushes a buffer -- one
// global atomic for its base slot, then coalesced stores -- once it holds more than
// CAP / 2 items, so one pass of CAP / 2 items always fits.  CAP keeps both buffers of
// keys and values within 32 KiB (four workgroups per CU).
template <class UK, class V>
struct TopkStage
{
  static constexpr bool kHasValue = !std::is_same<V, NoValue>::value;
  static constexpr int kItemBytes = int(sizeof(UK)) + (kHasValue ? int(sizeof(V)) : 0);
  static constexpr int kCap       = kItemBytes <= 8 ? 2048 : (kItemBytes <= 16 ? 1024 : 512);
  static constexpr int kIpt       = kCap / (2 * kTopkThreads); // one pass: kCap / 2 items
  static constexpr int kValCap    = kHasValue ? kCap : 1;
  using VS                        = typename std::conditional<kHasValue, V, uint8_t>::type;
};

// Writes the buffered items [0, count) to the output from slot `base` on, keeping only
// slots below `limit` (relative to `base`'s origin); every thread of the workgroup calls it.
template <class UK, class V>
__device__ __forceinline__ void topk_flush(const UK* s_keys, const typename TopkStage<UK, V>::VS* s_vals,
                                           uint32_t count, uint64_t base, uint64_t limit, uint64_t out0,
                                           UK* keys_out, void* vals_out)
{
  for (uint32_t i = threadIdx.x; i < count; i += kTopkThreads)
  {
    if (base + i < limit)
    {
      keys_out[out0 + base + i] = s_keys[i];
      if constexpr (TopkStage<UK, V>::kHasValue)
      {
        static_cast<V*>(vals_out)[out0 + base + i] = s_vals[i];
      }
    }
  }
}

template <class UK, class V>
__global__ void __launch_bounds__(kTopkThreads) topk_output_kernel(TopkOutputParams p)
{
  using S                  = TopkStage<UK, V>;
  constexpr int kCap       = S::kCap;
  constexpr int kIpt       = S::kIpt;
  constexpr uint32_t kFull = uint32_t(kCap / 2); // flush above this many items
  __shared__ UK s_klt[kCap];
  __shared__ UK s_keq[kCap];
  __shared__ typename S::VS s_vlt[S::kValCap];
  __shared__ typename S::VS s_veq[S::kValCap];
  __shared__ uint32_t s_n[2]; // items buffered: below, equal
  __shared__ uint64_t s_base;
  __shared__ uint32_t s_eq_full;

  TopkState* st         = p.st;
  const int s           = st->done != 0u ? int(st->done_shift) : 0;
  const UK thr          = UK(UK(st->prefix) >> s);
  const uint64_t k_rem  = st->k_rem;
  const uint64_t eq0    = p.k - k_rem; // keys below the threshold fill [0, eq0), equal ones [eq0, k)
  const UK* keys        = static_cast<const UK*>(p.keys_in);
  UK* keys_out          = static_cast<UK*>(p.keys_out);
  const bool select_max = p.select_max != 0;
  if (threadIdx.x == 0)
  {
    s_n[0] = s_n[1] = 0u;
    s_eq_full       = 0u;
  }
  __syncthreads();

  const uint64_t chunk = uint64_t(kTopkThreads) * kIpt;
  for (uint64_t base = uint64_t(blockIdx.x) * chunk; base < p.num_items; base += uint64_t(gridDim.x) * chunk)
  {
    UK k[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j)
    {
      const uint64_t i = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      k[j]             = i < p.num_items ? keys[i] : UK(0);
    }
    const bool take_eq = s_eq_full == 0u;
#pragma unroll
    for (int j = 0; j < kIpt; ++j)
    {
      const uint64_t i   = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      const bool in      = i < p.num_items;
      const UK o         = UK(topk_order(k[j], p.key_kind, select_max) >> s);
      const bool lt      = in && o < thr;
      const bool eq      = in && take_eq && o == thr;
      const uint64_t mlt = __ballot(lt);
      const uint64_t meq = __ballot(eq);
      if (lt)
      {
        const uint32_t slot = wave_lds_slot(&s_n[0], mlt);
        s_klt[slot]         = k[j];
        if constexpr (S::kHasValue)
        {
          s_vlt[slot] = static_cast<const V*>(p.vals_in)[i];
        }
      }
      if (eq)
      {
        const uint32_t slot = wave_lds_slot(&s_n[1], meq);
        s_keq[slot]         = k[j];
        if constexpr (S::kHasValue)
        {
          s_veq[slot] = static_cast<const V*>(p.vals_in)[i];
        }
      }
    }
    __syncthreads();
    const bool last = base + uint64_t(gridDim.x) * chunk >= p.num_items;
    const uint32_t nlt = s_n[0];
    const uint32_t neq = s_n[1];
    if (nlt > kFull || (last && nlt > 0u))
    {
      if (threadIdx.x == 0)
      {
        s_base = __hip_atomic_fetch_add(&st->count_lt, nlt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      topk_flush<UK, V>(s_klt, s_vlt, nlt, s_base, eq0, 0, keys_out, p.vals_out); // below eq0: always
      __syncthreads();
      if (threadIdx.x == 0)
      {
        s_n[0] = 0u;
      }
    }
    if (neq > kFull || (last && neq > 0u))
    {
      if (threadIdx.x == 0)
      {
        s_base = __hip_atomic_fetch_add(&st->count_eq, neq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_base + neq >= k_rem)
        {
          s_eq_full = 1u; // the tied slots are all taken: stop collecting equal keys
        }
      }
      __syncthreads();
      topk_flush<UK, V>(s_keq, s_veq, neq, s_base, k_rem, eq0, keys_out, p.vals_out);
      __syncthreads();
      if (threadIdx.x == 0)
      {
        s_n[1] = 0u;
      }
    }
    __syncthreads();
  }
}

template <class UK>
const void* output_kernel_for(int val_bytes)
{
  switch (val_bytes)
  {
    case 0:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, NoValue>);
    case 1:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, uint8_t>);
    case 2:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, uint16_t>);
    case 4:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, uint32_t>);
    case 8:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, uint64_t>);
    case 16:
      return reinterpret_cast<const void*>(&topk_output_kernel<UK, Value16>);
    default:
      return nullptr;
  }
}

struct TopkKernels
{
  const void* hist;
  const void* output;
};

TopkKernels topk_kernels(int key_bytes, int val_bytes)
{
  switch (key_bytes)
  {
    case 1:
      return {reinterpret_cast<const void*>(&topk_hist_kernel<uint8_t>), output_kernel_for<uint8_t>(val_bytes)};
    case 2:
      return {reinterpret_cast<const void*>(&topk_hist_kernel<uint16_t>), output_kernel_for<uint16_t>(val_bytes)};
    case 4:
      return {reinterpret_cast<const void*>(&topk_hist_kernel<uint32_t>), output_kernel_for<uint32_t>(val_bytes)};
    case 8:
      return {reinterpret_cast<const void*>(&topk_hist_kernel<uint64_t>), output_kernel_for<uint64_t>(val_bytes)};
    default:
      return {nullptr, nullptr};
  }
}

} // namespace

hipError_t topk(
  void* d_temp_storage,
  size_t& temp_storage_bytes,
  const void* d_keys_in,
  void* d_keys_out,
  const void* d_values_in,
  void* d_values_out,
  int key_bytes,
  int key_kind,
  int val_bytes,
  uint64_t num_items,
  uint64_t k,
  bool select_max,
  hipStream_t stream)
{
  const TopkKernels kern = topk_kernels(key_bytes, val_bytes);
  if (kern.hist == nullptr || kern.output == nullptr)
  {
    set_last_error("top-k: unsupported key width %d or value width %d", key_bytes, val_bytes);
    return hipErrorInvalidValue;
  }
  if (num_items > 0xffffffffull)
  {
    set_last_error("top-k: at most 2^32 - 1 items (bins and slots are 32-bit counters)");
    return hipErrorInvalidValue;
  }
  size_t sizes[1] = {sizeof(TopkState)};
  void* parts[1]  = {};
  if (hipError_t e = alias_temporaries(d_temp_storage, temp_storage_bytes, parts, sizes, 1); e != hipSuccess)
  {
    return e;
  }
  if (d_temp_storage == nullptr)
  {
    return hipSuccess;
  }
  k = std::min(k, num_items); // device_topk.cuh: k is capped to num_items
  if (k == 0)
  {
    return hipSuccess;
  }
  const bool pairs = val_bytes > 0;
  if (k == num_items)
  {
    // every item is selected; any order is a valid answer
    hipError_t e = hipMemcpyAsync(d_keys_out, d_keys_in, size_t(num_items) * key_bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess && pairs)
    {
      e = hipMemcpyAsync(d_values_out, d_values_in, size_t(num_items) * val_bytes, hipMemcpyDeviceToDevice, stream);
    }
    return e;
  }
  TopkState* st = static_cast<TopkState*>(parts[0]);
  {
    TopkInitParams ip{st, k};
    if (hipError_t e = launch_kernel(reinterpret_cast<const void*>(&topk_init_kernel), 1, kTopkThreads, 0, stream,
                                     &ip, 11);
        e != hipSuccess)
    {
      return e;
    }
  }
  const uint64_t chunks = (num_items + uint64_t(kTopkThreads) * kTopkIpt - 1) / (uint64_t(kTopkThreads) * kTopkIpt);
  const uint64_t grid   = std::min<uint64_t>(chunks, uint64_t(device_cu_count()) * 4);
  for (int pass = 0; pass < key_bytes; ++pass)
  {
    TopkPassParams pp{d_keys_in, num_items, st, key_kind, select_max ? 1 : 0, (key_bytes - 1 - pass) * kRadixBits};
    if (hipError_t e = launch_kernel(kern.hist, grid, kTopkThreads, 0, stream, &pp, 11); e != hipSuccess)
    {
      return e;
    }
    if (hipError_t e =
          launch_kernel(reinterpret_cast<const void*>(&topk_pick_kernel), 1, kTopkThreads, 0, stream, &pp, 11);
        e != hipSuccess)
    {
      return e;
    }
  }
  TopkOutputParams op{d_keys_in, d_values_in, d_keys_out, d_values_out, num_items, k, st, key_kind, select_max ? 1 : 0};
  return launch_kernel(kern.output, grid, kTopkThreads, 0, stream, &op, 11);
}

} // namespace cccl_amd

extern "C" CCCL_C_API int cccl_amd_topk(
  void* d_temp_storage,
  size_t* temp_storage_bytes,
  const void* d_keys_in,
  void* d_keys_out,
  const void* d_values_in,
  void* d_values_out,
  cccl_type_enum key_type,
  size_t value_bytes,
  uint64_t num_items,
  uint64_t k,
  int select_max,
  CUstream stream)
{
  if (temp_storage_bytes == nullptr)
  {
    return hipErrorInvalidValue;
  }
  int key_bytes = 0, key_kind = 0;
  if (!cccl_amd::key_traits_from_enum(int(key_type), key_bytes, key_kind)
      || !(value_bytes == 0 || value_bytes == 1 || value_bytes == 2 || value_bytes == 4 || value_bytes == 8
           || value_bytes == 16))
  {
    return hipErrorNotSupported;
  }
  if (d_temp_storage != nullptr && num_items > 0 && k > 0
      && (d_keys_in == nullptr || d_keys_out == nullptr
          || (value_bytes > 0 && (d_values_in == nullptr || d_values_out == nullptr))))
  {
    cccl_amd::set_last_error("top-k: a key or value buffer is null");
    return hipErrorInvalidValue;
  }
  return int(cccl_amd::topk(
    d_temp_storage,
    *temp_storage_bytes,
    d_keys_in,
    d_keys_out,
    d_values_in,
    d_values_out,
    key_bytes,
    key_kind,
    int(value_bytes),
    num_items,
    k,
    select_max != 0,
    reinterpret_cast<hipStream_t>(stream)));
}
