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
//                  (LDS bins per workgroup, one global atomic per non-zero bin); the
//                  second digit's pass is topk_filter instead (below)
//     topk_pick    one workgroup: the digit d whose bin holds the k_rem-th candidate;
//                  prefix |= d, k_rem -= candidates below d; when the bin is taken whole,
//                  the selection is complete and later passes return at once
//   topk_output  keys with o below the threshold take slots [0, k - k_rem) and keys equal
//                to it the first k_rem slots after that: each workgroup stages the items
//                it selects in LDS and flushes a batch with one global atomic (its base
//                slot) and coalesced stores.
// topk_filter (the second digit pass) also writes the keys below the top digit straight to
// the output and the keys on it -- the candidates -- with their input indices to a
// buffer of up to N / 16 items; the later passes and topk_output read that buffer instead
// of the input when the candidates fit (for spread keys about N / 256 of them), so only
// the first two passes read every key: about 2 * N * K bytes plus the candidates and the
// k outputs, against (P + 1) * N * K when every pass reads the input.
#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "../../include/cccl_amd/topk.h"
#include "dispatch.hpp"
#include "radix_common.hpp"

namespace cccl_amd {

namespace {

constexpr int kTopkThreads = 256;
constexpr int kTopkIpt     = 16; // items per thread per loop iteration of the streaming kernels
// Histogram workgroups per CU: twice what fits at once, which read 2^28 4-byte keys
// 7% faster on MI355X than one resident round.
constexpr int kTopkHistWgPerCu   = 16;

// Keys of at most 4 bytes use the tuned streaming set-up (16 histogram workgroups per CU,
// 1,024-item LDS staging, the filter pass prefetching its next chunk, as many filter
// workgroups as fit); with 8-byte keys the prefetch doubles the filter's registers (103
// VGPRs, 4 waves per SIMD) and that set-up was 18% slower over 2^27 keys, so they keep 4
// workgroups per CU for both passes, 2,048-item staging and no prefetch.
template <class UK>
struct TopkTuning
{
  static constexpr bool kNarrow     = sizeof(UK) <= 4;
  static constexpr bool kPrefetch   = kNarrow;
  static constexpr int kStageItems  = kNarrow ? 1024 : 2048; // LDS staging for items of at most 8 bytes
};

struct TopkState
{
  uint64_t prefix; // threshold digits chosen so far (order space)
  uint64_t k_rem; // items still to take among the keys that match the prefix
  uint32_t done; // 1: the bin at done_shift is taken whole, the selection is complete
  uint32_t done_shift;
  uint32_t count_lt; // output slots taken by keys below the threshold
  uint32_t count_eq; // output slots taken by keys equal to the threshold
  uint32_t filtered; // 1: the filter pass ran (keys below the top digit are out already)
  uint32_t cand_count; // candidates the filter pass found (kept when at most cand_cap)
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
    p.st->filtered   = 0u;
    p.st->cand_count = 0u;
  }
}

// Where a pass reads its keys: the whole input, or -- once the filter pass has run and
// its candidates fit -- the candidate buffer (keys and their input indices).
struct TopkSource
{
  const void* keys;
  const void* cand_keys;
  const uint32_t* cand_idx;
  uint64_t num_items;
  uint32_t cand_cap;
};

struct TopkSourceView
{
  bool cand;
  uint64_t count;
};

__device__ __forceinline__ TopkSourceView topk_source(const TopkState* st, const TopkSource& src)
{
  const bool cand = st->filtered != 0u && st->cand_count <= src.cand_cap;
  return {cand, cand ? uint64_t(st->cand_count) : src.num_items};
}

struct TopkPassParams
{
  TopkSource src;
  TopkState* st;
  int key_kind;
  int select_max;
  int shift; // lowest bit of this pass's digit
};

// One loop step of the streaming kernels: a workgroup's chunk of kTopkThreads * kTopkIpt
// keys from `base`, every key read once (non-temporal), all of a thread's loads in
// flight together.  A whole chunk of a 16-byte aligned array is read with 16-byte loads
// (a wave's load is 1 KiB of consecutive keys): key j of a thread is item
// (j / kPer) * kTopkThreads * kPer + threadIdx.x * kPer + j % kPer of the chunk.  A
// partial chunk is read one key per load, key j being item j * kTopkThreads + threadIdx.x,
// and keys past `count` read as 0.
template <class UK>
struct TopkChunk
{
  static constexpr int kPer = int(16 / sizeof(UK));
  static_assert(kTopkIpt % kPer == 0, "a thread's keys are whole 16-byte vectors");
  UK k[kTopkIpt];
  bool vec;

  __device__ __forceinline__ void load(const UK* keys, uint64_t base, uint64_t count)
  {
    using NV              = unsigned int __attribute__((ext_vector_type(4)));
    constexpr uint64_t kN = uint64_t(kTopkThreads) * kTopkIpt;
    vec                   = base + kN <= count && (reinterpret_cast<uintptr_t>(keys) & 15u) == 0u;
    if (vec)
    {
      const NV* src = reinterpret_cast<const NV*>(keys + base);
      NV x[kTopkIpt / kPer];
#pragma unroll
      for (int u = 0; u < kTopkIpt / kPer; ++u)
      {
        x[u] = __builtin_nontemporal_load(&src[u * kTopkThreads + threadIdx.x]);
      }
#pragma unroll
      for (int u = 0; u < kTopkIpt / kPer; ++u)
      {
        __builtin_memcpy(&k[u * kPer], &x[u], 16);
      }
    }
    else
    {
#pragma unroll
      for (int j = 0; j < kTopkIpt; ++j)
      {
        const uint64_t i = base + uint64_t(j) * kTopkThreads + threadIdx.x;
        k[j]             = i < count ? __builtin_nontemporal_load(&keys[i]) : UK(0);
      }
    }
  }

  // Input position of key j (past `count` for a padding key of a partial chunk).
  __device__ __forceinline__ uint64_t item(uint64_t base, int j) const
  {
    return vec ? base + uint64_t(j / kPer) * (kTopkThreads * kPer) + threadIdx.x * kPer + (j % kPer)
               : base + uint64_t(j) * kTopkThreads + threadIdx.x;
  }
};

template <class UK>
__global__ void __launch_bounds__(kTopkThreads) topk_hist_kernel(TopkPassParams p)
{
  __shared__ uint32_t bins[kRadix];
  if (p.st->done != 0u)
  {
    return;
  }
  constexpr int kBits       = int(sizeof(UK) * 8);
  const int hi              = p.shift + kRadixBits;
  const UK prefix_hi        = hi >= kBits ? UK(0) : UK(UK(p.st->prefix) >> hi);
  const TopkSourceView view = topk_source(p.st, p.src);
  const UK* keys            = static_cast<const UK*>(view.cand ? p.src.cand_keys : p.src.keys);
  const bool select_max     = p.select_max != 0;
  bins[threadIdx.x]         = 0u;
  __syncthreads();
  const uint64_t chunk = uint64_t(kTopkThreads) * kTopkIpt;
  for (uint64_t base = uint64_t(blockIdx.x) * chunk; base < view.count; base += uint64_t(gridDim.x) * chunk)
  {
    TopkChunk<UK> cur;
    cur.load(keys, base, view.count);
#pragma unroll
    for (int j = 0; j < kTopkIpt; ++j)
    {
      const uint64_t i = cur.item(base, j);
      const UK o       = topk_order(cur.k[j], p.key_kind, select_max);
      const bool match = i < view.count && (hi >= kBits || UK(o >> hi) == prefix_hi);
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

// LDS staging of selected items: a workgroup collects the items it selects in LDS
// buffers of CAP items and flushes a buffer -- one global atomic for its base slot, then
// coalesced stores -- once it holds more than CAP / 2 items, so one loop step of CAP / 2
// items always fits.  For 4-byte keys and values CAP keeps two buffers of keys and values
// within 16 KiB, so 7 filter workgroups fit on a CU (at 32 KiB only 4 did, and the pass
// over 2^28 keys took 0.32 ms instead of 0.29 ms).
template <class UK, class V>
struct TopkStage
{
  static constexpr bool kHasValue = !std::is_same<V, NoValue>::value;
  static constexpr int kItemBytes = int(sizeof(UK)) + (kHasValue ? int(sizeof(V)) : 0);
  static constexpr int kStage     = TopkTuning<UK>::kStageItems;
  static constexpr int kCapRaw    = kItemBytes <= 8 ? kStage : (kItemBytes <= 16 ? kStage / 2 : kStage / 4);
  static constexpr int kCap       = kCapRaw < 2 * kTopkThreads ? 2 * kTopkThreads : kCapRaw;
  static constexpr int kIpt       = kCap / (2 * kTopkThreads); // one loop step: kCap / 2 items
  static constexpr int kValCap    = kHasValue ? kCap : 1;
  using VS                        = typename std::conditional<kHasValue, V, uint8_t>::type;
};

// Writes the buffered items [0, count) from slot `base` on (slots below `limit` only,
// counted from `out0`); every thread of the workgroup calls it.
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

struct TopkFilterParams
{
  TopkSource src;
  const void* vals_in;
  void* keys_out;
  void* vals_out;
  uint64_t k;
  TopkState* st;
  void* cand_keys; // writable views of src.cand_keys / src.cand_idx
  uint32_t* cand_idx;
  int key_kind;
  int select_max;
};

// The second digit pass, run once the top digit d0 is picked: over the whole input, keys
// whose top digit is below d0 are selected for certain and go straight to the output;
// keys whose top digit is d0 are the candidates: their second digit is counted, and they
// (with their input index) go to the candidate buffer, which the later passes and the
// output pass read instead of the input when the candidates fit.
template <class UK, class V>
__global__ void __launch_bounds__(kTopkThreads) topk_filter_kernel(TopkFilterParams p)
{
  using S                  = TopkStage<UK, V>;
  constexpr int kCap       = S::kCap;
  constexpr int kIpt       = S::kIpt;
  constexpr uint32_t kFull = uint32_t(kCap / 2);
  constexpr int kBits      = int(sizeof(UK) * 8);
  constexpr int kTop       = kBits - kRadixBits;
  __shared__ uint32_t bins[kRadix];
  __shared__ UK s_klt[kCap];
  __shared__ typename S::VS s_vlt[S::kValCap];
  __shared__ UK s_kc[kCap];
  __shared__ uint32_t s_ic[kCap];
  __shared__ uint32_t s_n[2]; // items buffered: selected, candidates
  __shared__ uint64_t s_base;

  TopkState* st = p.st;
  if (st->done != 0u)
  {
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    st->filtered = 1u;
  }
  const UK d0           = UK(UK(st->prefix) >> kTop);
  const UK* keys        = static_cast<const UK*>(p.src.keys);
  UK* keys_out          = static_cast<UK*>(p.keys_out);
  UK* cand_keys         = static_cast<UK*>(p.cand_keys);
  const bool select_max = p.select_max != 0;
  const uint64_t n      = p.src.num_items;
  bins[threadIdx.x]     = 0u;
  if (threadIdx.x == 0)
  {
    s_n[0] = s_n[1] = 0u;
  }
  __syncthreads();

  // Each step stages a chunk of kTopkIpt keys per thread in groups of kIpt, with a flush
  // check after each group so a group always fits; with kPrefetch the next chunk's loads
  // are in flight meanwhile (2^28 4-byte keys and values: 0.29 ms against 0.35 ms
  // without).  The barriers order LDS accesses only (lds_barrier), so they do not wait
  // for those loads; the flushes' global stores have no reader in this kernel.
  constexpr bool kPrefetch = TopkTuning<UK>::kPrefetch;
  constexpr int kGroups    = kTopkIpt / kIpt;
  const uint64_t chunk     = uint64_t(kTopkThreads) * kTopkIpt;
  const uint64_t stride    = uint64_t(gridDim.x) * chunk;
  uint64_t base            = uint64_t(blockIdx.x) * chunk;
  TopkChunk<UK> ch;
  if (kPrefetch && base < n)
  {
    ch.load(keys, base, n);
  }
  while (base < n)
  {
    const uint64_t next  = base + stride;
    const bool last_step = next >= n;
    TopkChunk<UK> nxt;
    if constexpr (kPrefetch)
    {
      if (!last_step)
      {
        nxt.load(keys, next, n);
      }
    }
    else
    {
      ch.load(keys, base, n);
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g)
    {
#pragma unroll
      for (int jj = 0; jj < kIpt; ++jj)
      {
        const int j          = g * kIpt + jj;
        const uint64_t i     = ch.item(base, j);
        const bool in        = i < n;
        const UK o           = topk_order(ch.k[j], p.key_kind, select_max);
        const UK top         = UK(o >> kTop);
        const bool lt        = in && top < d0;
        const bool cand      = in && top == d0;
        const uint64_t mlt   = __ballot(lt);
        const uint64_t mcnd  = __ballot(cand);
        if (lt)
        {
          const uint32_t slot = wave_lds_slot(&s_n[0], mlt);
          s_klt[slot]         = ch.k[j];
          if constexpr (S::kHasValue)
          {
            s_vlt[slot] = static_cast<const V*>(p.vals_in)[i];
          }
        }
        if (cand)
        {
          lds_count(bins, extract_digit(o, kTop - kRadixBits, kRadix - 1));
          const uint32_t slot = wave_lds_slot(&s_n[1], mcnd);
          s_kc[slot]          = ch.k[j];
          s_ic[slot]          = uint32_t(i);
        }
      }
      lds_barrier();
      const bool last    = last_step && g == kGroups - 1;
      const uint32_t nlt = s_n[0];
      const uint32_t nc  = s_n[1];
      if (nlt > kFull || (last && nlt > 0u))
      {
        if (threadIdx.x == 0)
        {
          s_base = __hip_atomic_fetch_add(&st->count_lt, nlt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        lds_barrier();
        topk_flush<UK, V>(s_klt, s_vlt, nlt, s_base, p.k, 0, keys_out, p.vals_out); // always below k
        lds_barrier();
        if (threadIdx.x == 0)
        {
          s_n[0] = 0u;
        }
      }
      if (nc > kFull || (last && nc > 0u))
      {
        if (threadIdx.x == 0)
        {
          s_base = __hip_atomic_fetch_add(&st->cand_count, nc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        lds_barrier();
        for (uint32_t c = threadIdx.x; c < nc; c += kTopkThreads)
        {
          if (s_base + c < p.src.cand_cap) // beyond: the later passes read the input instead
          {
            cand_keys[s_base + c]  = s_kc[c];
            p.cand_idx[s_base + c] = s_ic[c];
          }
        }
        lds_barrier();
        if (threadIdx.x == 0)
        {
          s_n[1] = 0u;
        }
      }
      lds_barrier();
    }
    if constexpr (kPrefetch)
    {
      if (!last_step)
      {
        ch = nxt;
      }
    }
    base = next;
  }
  const uint32_t c = bins[threadIdx.x];
  if (c != 0u)
  {
    __hip_atomic_fetch_add(&st->bins[threadIdx.x], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

struct TopkOutputParams
{
  TopkSource src;
  const void* vals_in;
  void* keys_out;
  void* vals_out;
  uint64_t k;
  TopkState* st;
  int key_kind;
  int select_max;
};

// The last pass: keys below the threshold fill slots [count_lt, k - k_rem) (the filter
// pass filled the first ones), keys equal to it the k_rem slots after that, both staged
// in LDS per workgroup.  After a filter pass only keys with the top digit d0 are
// considered (the others are out already or not selected).
template <class UK, class V>
__global__ void __launch_bounds__(kTopkThreads) topk_output_kernel(TopkOutputParams p)
{
  using S                  = TopkStage<UK, V>;
  constexpr int kCap       = S::kCap;
  constexpr int kIpt       = S::kIpt;
  constexpr uint32_t kFull = uint32_t(kCap / 2); // flush above this many items
  constexpr int kBits      = int(sizeof(UK) * 8);
  constexpr int kTop       = kBits - kRadixBits;
  __shared__ UK s_klt[kCap];
  __shared__ UK s_keq[kCap];
  __shared__ typename S::VS s_vlt[S::kValCap];
  __shared__ typename S::VS s_veq[S::kValCap];
  __shared__ uint32_t s_n[2]; // items buffered: below, equal
  __shared__ uint64_t s_base;
  __shared__ uint32_t s_eq_full;

  TopkState* st             = p.st;
  const int s               = st->done != 0u ? int(st->done_shift) : 0;
  const UK thr              = UK(UK(st->prefix) >> s);
  const bool filtered       = st->filtered != 0u;
  const UK d0               = UK(UK(st->prefix) >> kTop);
  const uint64_t k_rem      = st->k_rem;
  const uint64_t eq0        = p.k - k_rem; // keys below the threshold fill [0, eq0), equal ones [eq0, k)
  const TopkSourceView view = topk_source(st, p.src);
  const UK* keys            = static_cast<const UK*>(view.cand ? p.src.cand_keys : p.src.keys);
  UK* keys_out              = static_cast<UK*>(p.keys_out);
  const bool select_max     = p.select_max != 0;
  if (threadIdx.x == 0)
  {
    s_n[0] = s_n[1] = 0u;
    s_eq_full       = 0u;
  }
  __syncthreads();

  const uint64_t chunk = uint64_t(kTopkThreads) * kIpt;
  for (uint64_t base = uint64_t(blockIdx.x) * chunk; base < view.count; base += uint64_t(gridDim.x) * chunk)
  {
    UK k[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j)
    {
      const uint64_t i = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      k[j]             = i < view.count ? keys[i] : UK(0);
    }
    const bool take_eq = s_eq_full == 0u;
#pragma unroll
    for (int j = 0; j < kIpt; ++j)
    {
      const uint64_t i   = base + uint64_t(j) * kTopkThreads + threadIdx.x;
      const UK full      = topk_order(k[j], p.key_kind, select_max);
      const bool in      = i < view.count && (!filtered || UK(full >> kTop) == d0);
      const UK o         = UK(full >> s);
      const bool lt      = in && o < thr;
      const bool eq      = in && take_eq && o == thr;
      const uint64_t mlt = __ballot(lt);
      const uint64_t meq = __ballot(eq);
      if (lt || eq)
      {
        const uint64_t src_i = view.cand ? uint64_t(p.src.cand_idx[i]) : i;
        if (lt)
        {
          const uint32_t slot = wave_lds_slot(&s_n[0], mlt);
          s_klt[slot]         = k[j];
          if constexpr (S::kHasValue)
          {
            s_vlt[slot] = static_cast<const V*>(p.vals_in)[src_i];
          }
        }
        else
        {
          const uint32_t slot = wave_lds_slot(&s_n[1], meq);
          s_keq[slot]         = k[j];
          if constexpr (S::kHasValue)
          {
            s_veq[slot] = static_cast<const V*>(p.vals_in)[src_i];
          }
        }
      }
    }
    __syncthreads();
    const bool last    = base + uint64_t(gridDim.x) * chunk >= view.count;
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

struct TopkKernels
{
  const void* hist;
  const void* filter;
  const void* output;
};

// Workgroups of kTopkThreads threads of kernel fn (static LDS only) that fit on one CU at
// once (at least 1), cached per kernel.
int topk_resident_workgroups(const void* fn)
{
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(fn);
  if (it != cache.end())
  {
    return it->second;
  }
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kTopkThreads, 0) != hipSuccess || n < 1)
  {
    (void) hipGetLastError();
    n = 1;
  }
  cache.emplace(fn, n);
  return n;
}

template <class UK>
TopkKernels kernels_for_value(int val_bytes)
{
  auto make = [](auto v) -> TopkKernels {
    using V = decltype(v);
    return {reinterpret_cast<const void*>(&topk_hist_kernel<UK>),
            reinterpret_cast<const void*>(&topk_filter_kernel<UK, V>),
            reinterpret_cast<const void*>(&topk_output_kernel<UK, V>)};
  };
  switch (val_bytes)
  {
    case 0:
      return make(NoValue{});
    case 1:
      return make(uint8_t{});
    case 2:
      return make(uint16_t{});
    case 4:
      return make(uint32_t{});
    case 8:
      return make(uint64_t{});
    case 16:
      return make(Value16{});
    default:
      return {nullptr, nullptr, nullptr};
  }
}

TopkKernels topk_kernels(int key_bytes, int val_bytes)
{
  switch (key_bytes)
  {
    case 1:
      return kernels_for_value<uint8_t>(val_bytes);
    case 2:
      return kernels_for_value<uint16_t>(val_bytes);
    case 4:
      return kernels_for_value<uint32_t>(val_bytes);
    case 8:
      return kernels_for_value<uint64_t>(val_bytes);
    default:
      return {nullptr, nullptr, nullptr};
  }
}

// Candidate-buffer capacity: a sixteenth of the input (at least 2^16 items, at most all).
uint64_t topk_cand_cap(uint64_t num_items)
{
  return std::min<uint64_t>(num_items, std::max<uint64_t>(num_items / 16, uint64_t(1) << 16));
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
  if (kern.hist == nullptr)
  {
    set_last_error("top-k: unsupported key width %d or value width %d", key_bytes, val_bytes);
    return hipErrorInvalidValue;
  }
  if (num_items > 0xffffffffull)
  {
    set_last_error("top-k: at most 2^32 - 1 items (bins, slots and indices are 32-bit)");
    return hipErrorInvalidValue;
  }
  const bool filter = key_bytes > 1; // one-byte keys have a single digit pass
  const uint64_t cap = filter ? topk_cand_cap(num_items) : 0;
  size_t sizes[3]    = {sizeof(TopkState), size_t(cap) * key_bytes, size_t(cap) * sizeof(uint32_t)};
  void* parts[3]     = {};
  if (hipError_t e = alias_temporaries(d_temp_storage, temp_storage_bytes, parts, sizes, 3); e != hipSuccess)
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
  const TopkSource src{d_keys_in, parts[1], static_cast<const uint32_t*>(parts[2]), num_items, uint32_t(cap)};
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
  // The streaming passes are bound by how many loads are in flight: for keys of at most
  // 4 bytes the histogram runs more workgroups per CU than fit at once, the filter pass
  // as many as fit (its LDS staging and registers decide: 7 for 4-byte keys and values).
  const bool narrow          = key_bytes <= 4;
  const uint64_t cus         = uint64_t(device_cu_count());
  const uint64_t hist_grid   = std::min<uint64_t>(chunks, cus * (narrow ? uint64_t(kTopkHistWgPerCu) : 4u));
  const uint64_t filter_grid = std::min<uint64_t>(
    chunks, cus * (narrow ? uint64_t(topk_resident_workgroups(kern.filter)) : 4u));
  const int sm          = select_max ? 1 : 0;
  for (int pass = 0; pass < key_bytes; ++pass)
  {
    const int shift = (key_bytes - 1 - pass) * kRadixBits;
    TopkPassParams pp{src, st, key_kind, sm, shift};
    if (pass == 1)
    {
      TopkFilterParams fp{src, d_values_in, d_keys_out, d_values_out, k, st, parts[1], static_cast<uint32_t*>(parts[2]),
                          key_kind, sm};
      if (hipError_t e = launch_kernel(kern.filter, filter_grid, kTopkThreads, 0, stream, &fp, 11); e != hipSuccess)
      {
        return e;
      }
    }
    else if (hipError_t e = launch_kernel(kern.hist, hist_grid, kTopkThreads, 0, stream, &pp, 11); e != hipSuccess)
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
  TopkOutputParams op{src, d_values_in, d_keys_out, d_values_out, k, st, key_kind, sm};
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
