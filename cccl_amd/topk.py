"""Python host API of cub::DeviceTopK over the typed entry point ``cccl_amd_topk``.

Mirrors cub::DeviceTopK::MaxKeys / MinKeys / MaxPairs / MinPairs
(cub/cub/device/device_topk.cuh:297, :538, :775): the k largest (or smallest) keys of
``d_keys_in[:num_items]`` go to ``d_keys_out[:k]`` in no particular order, with their
values for the pair forms; k is capped to num_items.  ``temp_storage=None`` returns the
temp-storage byte count (the two-phase protocol); with ``temp_storage`` omitted entirely
the call allocates it.  Keys are ordered by their radix bit image (+0.0 above -0.0, NaNs
by bit pattern); which of several items tied with the k-th key are returned is
unspecified.  Device arrays are torch tensors; every call goes through libcccl_amd.so.
"""

from __future__ import annotations

import ctypes

from . import _bindings as B
from .radix_sort import SortError, _item_bytes, _ptr, _stream_handle, _temp, key_type_of

_ALLOCATE = object()


def _topk(select_max, d_keys_in, d_keys_out, d_values_in, d_values_out, num_items, k, temp_storage, stream):
    L = B.lib()
    vbytes = 0 if d_values_in is None else _item_bytes(d_values_in)
    if (d_values_in is None) != (d_values_out is None):
        raise ValueError("pass both d_values_in and d_values_out, or neither")
    args = lambda t, n: (_ptr(t), ctypes.byref(n), _ptr(d_keys_in), _ptr(d_keys_out), _ptr(d_values_in),  # noqa: E731
                         _ptr(d_values_out), key_type_of(d_keys_in.dtype), vbytes, num_items, k, int(select_max),
                         _stream_handle(stream))
    if temp_storage is _ALLOCATE:
        nbytes = ctypes.c_size_t(0)
        rc = L.cccl_amd_topk(*args(None, nbytes))
        if rc != 0:
            raise SortError("cccl_amd_topk", rc)
        temp_storage = _temp(nbytes.value, d_keys_in)
    nbytes = ctypes.c_size_t(0 if temp_storage is None else temp_storage.numel() * temp_storage.element_size())
    rc = L.cccl_amd_topk(*args(temp_storage, nbytes))
    if rc != 0:
        raise SortError("cccl_amd_topk", rc)
    return nbytes.value


def max_keys(d_keys_in, d_keys_out, num_items: int, k: int, *, temp_storage=_ALLOCATE, stream=None) -> int:
    """cub::DeviceTopK::MaxKeys."""
    return _topk(True, d_keys_in, d_keys_out, None, None, num_items, k, temp_storage, stream)


def min_keys(d_keys_in, d_keys_out, num_items: int, k: int, *, temp_storage=_ALLOCATE, stream=None) -> int:
    """cub::DeviceTopK::MinKeys."""
    return _topk(False, d_keys_in, d_keys_out, None, None, num_items, k, temp_storage, stream)


def max_pairs(d_keys_in, d_keys_out, d_values_in, d_values_out, num_items: int, k: int, *, temp_storage=_ALLOCATE,
              stream=None) -> int:
    """cub::DeviceTopK::MaxPairs."""
    return _topk(True, d_keys_in, d_keys_out, d_values_in, d_values_out, num_items, k, temp_storage, stream)


def min_pairs(d_keys_in, d_keys_out, d_values_in, d_values_out, num_items: int, k: int, *, temp_storage=_ALLOCATE,
              stream=None) -> int:
    """cub::DeviceTopK::MinPairs."""
    return _topk(False, d_keys_in, d_keys_out, d_values_in, d_values_out, num_items, k, temp_storage, stream)


def temp_storage_bytes(d_keys_in, num_items: int, k: int, value_bytes: int = 0) -> int:
    """Temp-storage bytes of a top-k call (no launch)."""
    nbytes = ctypes.c_size_t(0)
    rc = B.lib().cccl_amd_topk(None, ctypes.byref(nbytes), None, None, None, None, key_type_of(d_keys_in.dtype),
                               value_bytes, num_items, k, 1, None)
    if rc != 0:
        raise SortError("cccl_amd_topk", rc)
    return nbytes.value


__all__ = ["max_keys", "min_keys", "max_pairs", "min_pairs", "temp_storage_bytes"]
