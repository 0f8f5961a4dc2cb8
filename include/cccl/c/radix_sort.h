/* Drop-in C-ABI of the device-wide LSD radix sort on MI355X (gfx950).
 *
 * Every declaration here replaces the entry of the same name in
 * c/parallel/include/cccl/c/radix_sort.h (struct :27-64, functions :66-143); the
 * argument order, types and error codes are the reference's.  What differs is what
 * "build" means: the reference JIT-compiles CUB through NVRTC and links a CUBIN; this
 * library ships precompiled gfx950 code objects for every supported
 * (key type, value size, order) and "build" selects them plus a launch
 * configuration (the runtime_policy, a cccl_amd_radix_sort_config_t).
 *
 * Contract kept from the reference:
 *   - keys-only is input_values_it.type == CCCL_POINTER with state == NULL
 *     (c/parallel/src/radix_sort.cu:240);
 *   - all four run-time iterators must be CCCL_POINTER (:599-607), else CUDA_ERROR_UNKNOWN;
 *   - d_temp_storage == NULL writes *temp_storage_bytes and launches nothing;
 *   - *selector receives the keys DoubleBuffer selector: with is_overwrite_okay the
 *     result is in d_keys_in when 0 and in d_keys_out when 1; without it the result
 *     is always in d_keys_out (:643);
 *   - only the identity decomposer (decomposer.name == NULL or empty) is accepted.
 *     A custom decomposer needs LTO-IR linking, which has no gfx950 counterpart here:
 *     build returns CUDA_ERROR_NOT_SUPPORTED for it. */
#pragma once

#ifndef CCCL_C_EXPERIMENTAL
#  error "C exposure is experimental and subject to change. Define CCCL_C_EXPERIMENTAL to acknowledge this notice."
#endif

#include <stdbool.h>
#include <stdint.h>

#include <cccl/c/extern_c.h>
#include <cccl/c/types.h>

CCCL_C_EXTERN_C_BEGIN

/* Field order and types as c/parallel/include/cccl/c/radix_sort.h:27-64.
 * The 11 kernel handles name the reference's kernels.  Each digit pass here is an
 * upsweep, a scan of the bins and a downsweep, so load sets single_tile, upsweep,
 * scan_bins and downsweep; the handles of kernels this design does not have (the
 * alternates, histogram, exclusive_sum, init_bins_and_counters, init_lookback, onesweep)
 * stay NULL. */
typedef struct cccl_device_radix_sort_build_result_t
{
  int cc;
  void* payload;
  size_t payload_size;
  cccl_payload_kind_t payload_kind;
  CUlibrary library;
  cccl_type_info key_type;
  cccl_type_info value_type;
  CUkernel single_tile_kernel;
  CUkernel upsweep_kernel;
  CUkernel alt_upsweep_kernel;
  CUkernel scan_bins_kernel;
  CUkernel downsweep_kernel;
  CUkernel alt_downsweep_kernel;
  CUkernel histogram_kernel;
  CUkernel exclusive_sum_kernel;
  CUkernel init_bins_and_counters_kernel;
  CUkernel init_lookback_kernel;
  CUkernel onesweep_kernel;
  cccl_sort_order_t order;
  void* runtime_policy;
  size_t runtime_policy_size;
  char* single_tile_kernel_lowered_name;
  char* upsweep_kernel_lowered_name;
  char* alt_upsweep_kernel_lowered_name;
  char* scan_bins_kernel_lowered_name;
  char* downsweep_kernel_lowered_name;
  char* alt_downsweep_kernel_lowered_name;
  char* histogram_kernel_lowered_name;
  char* exclusive_sum_kernel_lowered_name;
  char* init_bins_and_counters_kernel_lowered_name;
  char* init_lookback_kernel_lowered_name;
  char* onesweep_kernel_lowered_name;
} cccl_device_radix_sort_build_result_t;

/* Replaces radix_sort.h:66-78 (c/parallel/src/radix_sort.cu:700).  cc_major/cc_minor
 * are the gfx target's major/minor (9, 5 for gfx950; 0 means "current device");
 * cub_path, thrust_path, libcudacxx_path and ctk_path are accepted and ignored. */
CCCL_C_API CUresult cccl_device_radix_sort_build(
  cccl_device_radix_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t input_keys_it,
  cccl_iterator_t input_values_it,
  cccl_op_t decomposer,
  const char* decomposer_return_type,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path);

/* Replaces radix_sort.h:81-94.  config's extra flags are ignored (nothing is compiled). */
CCCL_C_API CUresult cccl_device_radix_sort_build_ex(
  cccl_device_radix_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t input_keys_it,
  cccl_iterator_t input_values_it,
  cccl_op_t decomposer,
  const char* decomposer_return_type,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path,
  cccl_build_config* config);

/* Replaces radix_sort.h:96-109: fills payload/policy/names without resolving kernels. */
CCCL_C_API CUresult cccl_device_radix_sort_compile(
  cccl_device_radix_sort_build_result_t* build,
  cccl_sort_order_t sort_order,
  cccl_iterator_t input_keys_it,
  cccl_iterator_t input_values_it,
  cccl_op_t decomposer,
  const char* decomposer_return_type,
  int cc_major,
  int cc_minor,
  const char* cub_path,
  const char* thrust_path,
  const char* libcudacxx_path,
  const char* ctk_path,
  cccl_build_config* config);

/* Replaces radix_sort.h:111: resolves the kernel handles named by the payload. */
CCCL_C_API CUresult cccl_device_radix_sort_load(cccl_device_radix_sort_build_result_t* build);

/* Replaces radix_sort.h:113-127 (c/parallel/src/radix_sort.cu:663). */
CCCL_C_API CUresult cccl_device_radix_sort(
  cccl_device_radix_sort_build_result_t build,
  void* d_temp_storage,
  size_t* temp_storage_bytes,
  cccl_iterator_t d_keys_in,
  cccl_iterator_t d_keys_out,
  cccl_iterator_t d_values_in,
  cccl_iterator_t d_values_out,
  cccl_op_t decomposer,
  uint64_t num_items,
  int begin_bit,
  int end_bit,
  bool is_overwrite_okay,
  int* selector,
  CUstream stream);

/* Replaces radix_sort.h:129-130.  LTO-IR has no gfx950 meaning: returns
 * CUDA_ERROR_NOT_SUPPORTED when num_inputs > 0, CUDA_SUCCESS for an empty list. */
CCCL_C_API CUresult cccl_device_radix_sort_link_ltoir(
  cccl_device_radix_sort_build_result_t* build, const void** input_blobs, const size_t* input_sizes, size_t num_inputs);

/* Replaces radix_sort.h:136-137.  The blob is this library's own format (magic
 * "CCAMDRS1", key/value type info, order, payload, policy, 11 names); free it with
 * cccl_serialization_buffer_free. */
CCCL_C_API CUresult
cccl_device_radix_sort_serialize(const cccl_device_radix_sort_build_result_t* build, void** out_buf, size_t* out_size);

/* Replaces radix_sort.h:142-143.  Kernel handles stay NULL until _load. */
CCCL_C_API CUresult
cccl_device_radix_sort_deserialize(cccl_device_radix_sort_build_result_t* build, const void* buf, size_t size);

/* Replaces radix_sort.h:145. */
CCCL_C_API CUresult cccl_device_radix_sort_cleanup(cccl_device_radix_sort_build_result_t* bld_ptr);

/* Replaces c/parallel/include/cccl/c/serialization.h:44. */
CCCL_C_API void cccl_serialization_buffer_free(void* buf);

CCCL_C_EXTERN_C_END
