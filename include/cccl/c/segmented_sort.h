/* Drop-in C-ABI of cub::DeviceSegmentedSort on MI355X (gfx950).
 *
 * Every declaration here replaces the entry of the same name in
 * c/parallel/include/cccl/c/segmented_sort.h (struct :27-53, functions :56-139); the
 * argument order, types and error codes are the reference's.  "build" selects
 * precompiled gfx950 code objects of the batched segmented radix sort (size-binned
 * one-wave / one-workgroup sorts for small segments, batched 8-bit digit passes for
 * large ones) plus a launch configuration; nothing is JIT-compiled.
 *
 * Contract kept from the reference (c/parallel/src/segmented_sort.cu):
 *   - keys-only is d_values_in.type == CCCL_POINTER with state == NULL (:436);
 *   - d_temp_storage == NULL writes *temp_storage_bytes and launches nothing;
 *   - every segment [start_offset_in[i], end_offset_in[i]) is sorted whole by key, `<`
 *     ascending or `>` descending, equal keys (and -0.0 / +0.0) in input order, the
 *     order std::stable_sort gives (c/parallel/test/test_segmented_sort.cpp:425-436);
 *     items outside every segment are not written;
 *   - *selector receives the keys DoubleBuffer selector after the call (:929): without
 *     is_overwrite_okay 1 (the result is in d_keys_out / d_values_out); with it
 *     (0 + P) & 1, where P is CUB's large-segment pass count, ceil(8K / 6) for K-byte
 *     keys with K > 1 and 2 for one-byte keys
 *     (cub/cub/device/dispatch/dispatch_segmented_sort.cuh:1293-1299, :1537-1538;
 *     tuning_segmented_sort.cuh:193); 0 when num_items or num_segments is 0.
 *
 * Differences, each rejected with an error rather than run differently:
 *   - keys, values and offsets must be CCCL_POINTER iterators (offset iterators with
 *     device-side advance / dereference ops need LTO-IR linking): build returns
 *     CUDA_ERROR_NOT_SUPPORTED for other offset kinds, the run CUDA_ERROR_UNKNOWN;
 *   - offsets are 32- or 64-bit integers; unsigned 32-bit offsets need
 *     num_items < 2^31;
 *   - values are 1, 2, 4, 8 or 16 bytes wide;
 *   - more than 2^31 - 1 segments: CUDA_ERROR_INVALID_VALUE (the Python layer of the
 *     reference rejects them too, cuda/compute/algorithms/_sort/_segmented_sort.py:114). */
#pragma once

#ifndef CCCL_C_EXPERIMENTAL
#  error "C exposure is experimental and subject to change. Define CCCL_C_EXPERIMENTAL to acknowledge this notice."
#endif

#include <stdbool.h>
#include <stdint.h>

#include <cccl/c/extern_c.h>
#include <cccl/c/types.h>

CCCL_C_EXTERN_C_BEGIN

/* Field order and types as c/parallel/include/cccl/c/segmented_sort.h:27-53.
 * runtime_policy holds a cccl_amd_segmented_config_t (size classes, large-segment tile,
 * digit-pass strategy); partition_runtime_policy a cccl_amd_segment_partition_policy_t
 * (the item bound of each size class, what the classifier bins by).  load sets all five
 * kernel handles: fallback = the one-workgroup-per-segment digit pass, small = the
 * one-wave / one-workgroup segment sort, large = the batched digit pass,
 * three_way_partition_init = the class-counter reset, three_way_partition = the size
 * classifier.  The two selector ops stay zeroed: the classifier compares sizes itself. */
typedef struct cccl_device_segmented_sort_build_result_t
{
  int cc;
  void* payload;
  size_t payload_size;
  cccl_payload_kind_t payload_kind;
  CUlibrary library;
  cccl_type_info key_type;
  cccl_type_info offset_type;
  cccl_op_t large_segments_selector_op;
  cccl_op_t small_segments_selector_op;
  CUkernel segmented_sort_fallback_kernel;
  CUkernel segmented_sort_kernel_small;
  CUkernel segmented_sort_kernel_large;
  CUkernel three_way_partition_init_kernel;
  CUkernel three_way_partition_kernel;
  void* runtime_policy;
  size_t runtime_policy_size;
  void* partition_runtime_policy;
  size_t partition_runtime_policy_size;
  cccl_sort_order_t order;
  char* segmented_sort_fallback_kernel_lowered_name;
  char* segmented_sort_kernel_small_lowered_name;
  char* segmented_sort_kernel_large_lowered_name;
  char* three_way_partition_init_kernel_lowered_name;
  char* three_way_partition_kernel_lowered_name;
} cccl_device_segmented_sort_build_result_t;

/* Item bounds of the size classes (partition_runtime_policy). */
typedef struct cccl_amd_segment_partition_policy_t
{
  int32_t num_classes; /* one-wave / one-workgroup classes, 0..6 */
  uint32_t class_max_items[6]; /* a segment joins the first class whose bound holds it */
  uint32_t large_tile_items; /* tile of the batched digit passes for larger segments */
} cccl_amd_segment_partition_policy_t;

/* Replaces segmented_sort.h:56-68.  cc_major/cc_minor are the gfx target's major/minor
 * (9, 5 for gfx950; 0, 0 means "current device"); the four paths are accepted and
 * ignored. */
CCCL_C_API CUresult cccl_device_segmented_sort_build(
  cccl_device_segmented_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t d_keys_in,
  cccl_iterator_t d_values_in,
  cccl_iterator_t begin_offset_in,
  cccl_iterator_t end_offset_in,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path);

/* Replaces segmented_sort.h:71-84.  config's extra flags are ignored. */
CCCL_C_API CUresult cccl_device_segmented_sort_build_ex(
  cccl_device_segmented_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t d_keys_in,
  cccl_iterator_t d_values_in,
  cccl_iterator_t begin_offset_in,
  cccl_iterator_t end_offset_in,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path,
  cccl_build_config* config);

/* Replaces segmented_sort.h:86-99: fills payload, policies and names; the library and
 * kernel handles stay NULL until _load. */
CCCL_C_API CUresult cccl_device_segmented_sort_compile(
  cccl_device_segmented_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t d_keys_in,
  cccl_iterator_t d_values_in,
  cccl_iterator_t begin_offset_in,
  cccl_iterator_t end_offset_in,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path,
  cccl_build_config* config);

/* Replaces segmented_sort.h:101: resolves the kernel handles named by the payload. */
CCCL_C_API CUresult cccl_device_segmented_sort_load(cccl_device_segmented_sort_build_result_t* build);

/* Replaces segmented_sort.h:103-117 (c/parallel/src/segmented_sort.cu:949-985). */
CCCL_C_API CUresult cccl_device_segmented_sort(
  cccl_device_segmented_sort_build_result_t build,
  void* d_temp_storage,
  size_t* temp_storage_bytes,
  cccl_iterator_t d_keys_in,
  cccl_iterator_t d_keys_out,
  cccl_iterator_t d_values_in,
  cccl_iterator_t d_values_out,
  uint64_t num_items,
  uint64_t num_segments,
  cccl_iterator_t start_offset_in,
  cccl_iterator_t end_offset_in,
  bool is_overwrite_okay,
  int* selector,
  CUstream stream);

/* Replaces segmented_sort.h:119-123: CUDA_ERROR_NOT_SUPPORTED when num_inputs > 0,
 * CUDA_SUCCESS for an empty list. */
CCCL_C_API CUresult cccl_device_segmented_sort_link_ltoir(
  cccl_device_segmented_sort_build_result_t* build,
  const void** input_blobs,
  const size_t* input_sizes,
  size_t num_inputs);

/* Replaces segmented_sort.h:125-130 (:129).  The blob is this library's own format (magic
 * "CCAMDSS1", key / offset type info, order, payload, both policies, 5 names); free it
 * with cccl_serialization_buffer_free. */
CCCL_C_API CUresult cccl_device_segmented_sort_serialize(
  const cccl_device_segmented_sort_build_result_t* build, void** out_buf, size_t* out_size);

/* Replaces segmented_sort.h:132-137 (:136).  Kernel handles stay NULL until _load; on failure
 * *build is left unchanged. */
CCCL_C_API CUresult
cccl_device_segmented_sort_deserialize(cccl_device_segmented_sort_build_result_t* build, const void* buf, size_t size);

/* Replaces segmented_sort.h:139: frees payload, policies and names. */
CCCL_C_API CUresult cccl_device_segmented_sort_cleanup(cccl_device_segmented_sort_build_result_t* bld_ptr);

CCCL_C_EXTERN_C_END
