/* Device-wide top-k selection on MI355X (gfx950): cub::DeviceTopK.
 *
 * The reference has no C-ABI for top-k (c/parallel/include/cccl/c/ holds none); this
 * typed entry point follows the C++ API it replaces, cub::DeviceTopK::MaxKeys / MinKeys /
 * MaxPairs / MinPairs (cub/cub/device/device_topk.cuh:297, :538, :775 and their keys-only
 * overloads), in the style of cccl_amd_radix_sort. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifndef CCCL_C_EXPERIMENTAL
#  define CCCL_C_EXPERIMENTAL
#endif
#include <cccl/c/types.h>

CCCL_C_EXTERN_C_BEGIN

/* Writes the k largest (select_max != 0) or smallest keys of d_keys_in[0, num_items) to
 * d_keys_out[0, k), and with value_bytes > 0 the values of the same items to
 * d_values_out[0, k).  k is capped to num_items (device_topk.cuh, "Capped to a maximum
 * of num_items").  Keys are ordered by their radix-sort bit image: +0.0 ranks above
 * -0.0 and NaNs rank by bit pattern (dispatch_topk.cuh:57-82).  The output is in no
 * particular order, and when several items tie with the k-th key, which of them are
 * returned is unspecified -- the one mode the reference supports
 * (device_topk.cuh:86-92: determinism not guaranteed, output unsorted).  Input and
 * output ranges must not overlap.  d_temp_storage == NULL writes *temp_storage_bytes
 * and launches nothing.  value_bytes is 0, 1, 2, 4, 8 or 16; num_items < 2^32.
 * Returns a hipError_t value. */
CCCL_C_API int cccl_amd_topk(
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
  CUstream stream);

CCCL_C_EXTERN_C_END
