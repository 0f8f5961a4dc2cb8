"""Python host API of cub::DeviceSegmentedSort over the cccl_device_segmented_sort_* C-ABI.

Mirrors cuda.compute's segmented sort
(python/cuda_cccl/cuda/compute/algorithms/_sort/_segmented_sort.py:24-280):
``make_segmented_sort`` builds a reusable sorter through
``cccl_device_segmented_sort_build``; calling it with ``temp_storage=None`` returns the
temp-storage byte count, with a buffer it sorts.  ``segmented_sort`` is the one-shot
form.  Keys (and values) given as ``DoubleBuffer`` mean is_overwrite_okay: after the
call their ``selector`` names the buffer holding the result.  Every segment
``[start_offsets_in[i], end_offsets_in[i])`` is sorted whole by key, equal keys in input
order.  Device arrays are torch tensors on a ROCm device; offsets are int32 or int64
tensors.  Every call goes through libcccl_amd.so: there is no host fallback.
"""

from __future__ import annotations

import ctypes

import torch

from . import _bindings as B
from .radix_sort import DoubleBuffer, SortError, SortOrder, _arrays, _item_bytes, _ptr, _stream_handle, _temp, \
    key_type_of

_OFFSET_TYPES = {torch.int32: B.CCCL_INT32, torch.int64: B.CCCL_INT64}


def _offset_type(t) -> tuple[int, int]:
    try:
        return _OFFSET_TYPES[t.dtype], t.element_size()
    except KeyError:
        raise TypeError(f"unsupported offset dtype {t.dtype}: use int32 or int64") from None


class _SegmentedSort:
    """A built segmented sorter (cccl_device_segmented_sort_build result) for one key /
    value / offset type and order."""

    def __init__(self, d_in_keys, d_out_keys, d_in_values, d_out_values, start_offsets_in, end_offsets_in,
                 order: SortOrder):
        L = B.lib()
        ki, _, vi, _ = _arrays(d_in_keys, d_out_keys, d_in_values, d_out_values)
        self.key_dtype = ki.dtype
        self.key_type = key_type_of(ki.dtype)
        self.value_bytes = 0 if vi is None else _item_bytes(vi)
        self.offset_dtype = start_offsets_in.dtype
        if end_offsets_in.dtype != self.offset_dtype:
            raise TypeError("start and end offsets must have the same dtype")
        off_type, off_size = _offset_type(start_offsets_in)
        self.order = order
        keys_it = B.pointer_iterator(self.key_type, ki.element_size(), None)
        vals_it = B.pointer_iterator(B.CCCL_STORAGE, self.value_bytes, 1 if self.value_bytes else None)
        off_it = B.pointer_iterator(off_type, off_size, None)
        self.build = B.cccl_device_segmented_sort_build_result_t()
        rc = L.cccl_device_segmented_sort_build(ctypes.byref(self.build), order.value, keys_it, vals_it, off_it,
                                                off_it, 9, 5, None, None, None, None)
        if rc != B.CUDA_SUCCESS:
            raise SortError("cccl_device_segmented_sort_build", rc)
        self._off = (off_type, off_size)

    def __del__(self):
        try:
            if getattr(self, "build", None) is not None and self.build.payload:
                B.lib().cccl_device_segmented_sort_cleanup(ctypes.byref(self.build))
        except Exception:
            pass

    def __call__(self, *, temp_storage, d_in_keys, d_out_keys, d_in_values, d_out_values, num_items: int,
                 num_segments: int, start_offsets_in, end_offsets_in, stream=None) -> int:
        L = B.lib()
        if num_segments > 2**31 - 1:
            raise RuntimeError("Segmented sort does not currently support more than 2^31-1 segments.")
        ki, ko, vi, vo = _arrays(d_in_keys, d_out_keys, d_in_values, d_out_values)
        if ki.dtype != self.key_dtype:
            raise TypeError("key dtype differs from the one the sorter was built for")
        if start_offsets_in.dtype != self.offset_dtype or end_offsets_in.dtype != self.offset_dtype:
            raise TypeError("offset dtype differs from the one the sorter was built for")
        is_overwrite_okay = isinstance(d_in_keys, DoubleBuffer)
        vbytes = self.value_bytes
        it = lambda t, ty, sz: B.pointer_iterator(ty, sz, _ptr(t))  # noqa: E731
        k_in, k_out = it(ki, self.key_type, ki.element_size()), it(ko, self.key_type, ki.element_size())
        v_in, v_out = it(vi, B.CCCL_STORAGE, vbytes), it(vo, B.CCCL_STORAGE, vbytes)
        o_begin, o_end = it(start_offsets_in, *self._off), it(end_offsets_in, *self._off)
        if temp_storage is None:
            d_temp, nbytes = None, ctypes.c_size_t(0)
        else:
            d_temp, nbytes = _ptr(temp_storage), ctypes.c_size_t(temp_storage.numel() * temp_storage.element_size())
        selector = ctypes.c_int(-1)
        rc = L.cccl_device_segmented_sort(self.build, d_temp, ctypes.byref(nbytes), k_in, k_out, v_in, v_out,
                                          num_items, num_segments, o_begin, o_end, is_overwrite_okay,
                                          ctypes.byref(selector), _stream_handle(stream))
        if rc != B.CUDA_SUCCESS:
            raise SortError("cccl_device_segmented_sort", rc)
        if is_overwrite_okay and temp_storage is not None:
            assert selector.value in (0, 1)
            d_in_keys.selector = selector.value
            if d_in_values is not None:
                d_in_values.selector = selector.value
        return nbytes.value


def make_segmented_sort(*, d_in_keys, d_out_keys=None, d_in_values=None, d_out_values=None, start_offsets_in,
                        end_offsets_in, order: SortOrder) -> _SegmentedSort:
    """A reusable segmented sorter (cuda.compute.make_segmented_sort)."""
    return _SegmentedSort(d_in_keys, d_out_keys, d_in_values, d_out_values, start_offsets_in, end_offsets_in, order)


def segmented_sort(*, d_in_keys, d_out_keys=None, d_in_values=None, d_out_values=None, num_items: int,
                   num_segments: int, start_offsets_in, end_offsets_in, order: SortOrder, stream=None):
    """One-shot segmented sort (cuda.compute.segmented_sort): query, allocate temp storage, sort."""
    sorter = make_segmented_sort(d_in_keys=d_in_keys, d_out_keys=d_out_keys, d_in_values=d_in_values,
                                 d_out_values=d_out_values, start_offsets_in=start_offsets_in,
                                 end_offsets_in=end_offsets_in, order=order)
    kw = dict(d_in_keys=d_in_keys, d_out_keys=d_out_keys, d_in_values=d_in_values, d_out_values=d_out_values,
              num_items=num_items, num_segments=num_segments, start_offsets_in=start_offsets_in,
              end_offsets_in=end_offsets_in, stream=stream)
    nbytes = sorter(temp_storage=None, **kw)
    ki = _arrays(d_in_keys, d_out_keys, d_in_values, d_out_values)[0]
    sorter(temp_storage=_temp(nbytes, ki), **kw)
