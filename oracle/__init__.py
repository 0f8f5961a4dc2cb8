"""CPU oracle of the radix sort -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package, and only as the checker / the baseline being timed.  The product path
(cccl_amd) never calls it.

* liboracle.so (oracle.cpp): closed-form restatement of CUB's test oracle
  (cub/test/catch2_radix_sort_helper.cuh:174-312).
* topk_order_bits / topk (numpy): the order of cub::DeviceTopK
  (cub/cub/device/dispatch/dispatch_topk.cuh:57-82) and the multiset of order images
  the k selected keys must have.
* _ref/libthrust_ref.so (thrust_ref.cpp): the reference's own Thrust OpenMP
  stable_sort[_by_key] (thrust/thrust/system/omp/detail/sort.h), compiled from
  /root/reference by oracle/Makefile.  Absent where the reference tree is absent
  at build time; the prebuilt file travels with the repository snapshot.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libthrust_ref.so"

KIND_UNSIGNED, KIND_SIGNED, KIND_FLOAT = 0, 1, 2

_c_u64p = ctypes.POINTER(ctypes.c_uint64)
_c_i64p = ctypes.POINTER(ctypes.c_int64)

_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle` (or __graft_entry__.build())")
        lib = ctypes.CDLL(str(ORACLE_SO))
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        lib.oracle_striped_keys.argtypes = [vp, u64, i32, i32, i32, i32, _c_u64p]
        lib.oracle_sort_permutation.argtypes = [vp, u64, i32, i32, i32, i32, i32, _c_u64p]
        lib.oracle_sort.argtypes = [vp, vp, u64, i32, i32, i32, i32, i32, i32, vp, vp]
        lib.oracle_segmented_sort.argtypes = [
            vp, vp, u64, i32, i32, i32, i32, i32, i32, ctypes.c_int64, _c_i64p, _c_i64p, vp, vp]
        lib.oracle_histograms.argtypes = [vp, u64, i32, i32, i32, i32, i32, _c_u64p]
        for f in (lib.oracle_striped_keys, lib.oracle_sort_permutation, lib.oracle_sort,
                  lib.oracle_segmented_sort, lib.oracle_histograms):
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


def ref_available() -> bool:
    return REF_SO.exists()


def _load_ref():
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise RuntimeError(f"{REF_SO} missing: built by `make -C oracle` where /root/reference exists")
        lib = ctypes.CDLL(str(REF_SO))
        lib.ref_stable_sort.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int]
        lib.ref_stable_sort.restype = ctypes.c_double
        lib.ref_max_threads.restype = ctypes.c_int
        _ref = lib
    return _ref


def key_kind(dtype) -> tuple[int, int]:
    """(bytes, kind) of a numpy key dtype."""
    dt = np.dtype(dtype)
    if dt.kind == "u" or dt.kind == "b":
        return dt.itemsize, KIND_UNSIGNED
    if dt.kind == "i":
        return dt.itemsize, KIND_SIGNED
    if dt.kind == "f":
        return dt.itemsize, KIND_FLOAT
    raise TypeError(f"unsupported key dtype {dt}")


def _ptr(a: np.ndarray | None):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _bits(dtype, begin_bit, end_bit):
    nbits = np.dtype(dtype).itemsize * 8
    return (0 if begin_bit is None else begin_bit), (nbits if end_bit is None else end_bit)


def striped(keys: np.ndarray, begin_bit=None, end_bit=None, kind=None) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    kd = kd if kind is None else kind
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    out = np.empty(keys.size, np.uint64)
    rc = _load().oracle_striped_keys(_ptr(keys), keys.size, kb, kd, b, e, out.ctypes.data_as(_c_u64p))
    assert rc == 0
    return out


def permutation(keys: np.ndarray, descending=False, begin_bit=None, end_bit=None, kind=None) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    kd = kd if kind is None else kind
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    out = np.empty(keys.size, np.uint64)
    rc = _load().oracle_sort_permutation(_ptr(keys), keys.size, kb, kd, int(descending), b, e,
                                         out.ctypes.data_as(_c_u64p))
    assert rc == 0
    return out


def sort(keys: np.ndarray, values: np.ndarray | None = None, descending=False, begin_bit=None,
         end_bit=None, kind=None):
    """Stable sort of keys (and values) by the striped key; returns (keys, values)."""
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    kd = kd if kind is None else kind
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    ko = np.empty_like(keys)
    vo = None
    vb = 0
    if values is not None:
        values = np.ascontiguousarray(values)
        assert values.shape[0] == keys.shape[0]
        vo = np.empty_like(values)
        vb = values.dtype.itemsize * (values.size // values.shape[0] if values.shape[0] else 1)
    rc = _load().oracle_sort(_ptr(keys), _ptr(values), keys.size, kb, kd, vb, int(descending), b, e,
                             _ptr(ko), _ptr(vo))
    assert rc == 0
    return ko, vo


def segmented_sort(keys, values, seg_begin, seg_end, keys_init, values_init=None, descending=False,
                   begin_bit=None, end_bit=None, kind=None):
    """Per-segment stable sort; items outside every segment keep keys_init/values_init."""
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    kd = kd if kind is None else kind
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    sb = np.ascontiguousarray(seg_begin, dtype=np.int64)
    se = np.ascontiguousarray(seg_end, dtype=np.int64)
    ko = np.array(keys_init, dtype=keys.dtype, copy=True)
    vo = None
    vb = 0
    if values is not None:
        values = np.ascontiguousarray(values)
        vo = np.array(values_init, dtype=values.dtype, copy=True)
        vb = values.dtype.itemsize
    rc = _load().oracle_segmented_sort(_ptr(keys), _ptr(values), keys.size, kb, kd, vb, int(descending), b, e,
                                       sb.size, sb.ctypes.data_as(_c_i64p), se.ctypes.data_as(_c_i64p),
                                       _ptr(ko), _ptr(vo))
    assert rc == 0
    return ko, vo


def histograms(keys: np.ndarray, descending=False, begin_bit=None, end_bit=None, kind=None) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    kd = kd if kind is None else kind
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    passes = (e - b + 7) // 8
    out = np.zeros((max(passes, 1), 256), np.uint64)
    rc = _load().oracle_histograms(_ptr(keys), keys.size, kb, kd, int(descending), b, e,
                                   out.ctypes.data_as(_c_u64p))
    assert rc == 0
    return out[:passes]


_REF_KEY_CODE = {np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2, np.dtype(np.int64): 3,
                 np.dtype(np.uint8): 4, np.dtype(np.uint16): 5, np.dtype(np.uint32): 6, np.dtype(np.uint64): 7,
                 np.dtype(np.float32): 9, np.dtype(np.float64): 10}

REF_MODE_DEFAULT, REF_MODE_FUNCTOR, REF_MODE_WINDOW = 0, 1, 2


def ref_sort_inplace(keys: np.ndarray, values: np.ndarray | None = None, descending=False, begin_bit=None,
                     end_bit=None, mode=REF_MODE_DEFAULT, threads=0) -> float:
    """Reference Thrust OMP stable sort of host arrays in place; returns seconds."""
    assert keys.flags.c_contiguous and (values is None or values.flags.c_contiguous)
    code = _REF_KEY_CODE[keys.dtype]
    b, e = _bits(keys.dtype, begin_bit, end_bit)
    vb = 0 if values is None else values.dtype.itemsize
    t = _load_ref().ref_stable_sort(code, _ptr(keys), _ptr(values), keys.size, vb, int(descending), b, e,
                                    mode, threads)
    if t < 0:
        raise ValueError("reference sort rejected the arguments")
    return t


def ref_max_threads() -> int:
    return _load_ref().ref_max_threads()


def topk_order_bits(keys: np.ndarray, select_max: bool) -> np.ndarray:
    """Order image of each key for cub::DeviceTopK, as uint64
    (cub/cub/device/dispatch/dispatch_topk.cuh:57-82): Traits<T>::TwiddleIn on the key
    bits -- unsigned identity, signed sign flip, floats all bits flipped when negative
    and the sign bit set otherwise, with no -0.0 / +0.0 merging -- then NOT for Max.
    DeviceTopK returns the k keys of smallest image."""
    keys = np.ascontiguousarray(keys)
    kb, kd = key_kind(keys.dtype)
    u = keys.view(f"u{kb}").astype(np.uint64)
    bits = 8 * kb
    mask = np.uint64((1 << bits) - 1)
    high = np.uint64(1 << (bits - 1))
    if kd == KIND_FLOAT:
        u = np.where((u & high) != 0, ~u & mask, u | high)
    elif kd == KIND_SIGNED:
        u = u ^ high
    if select_max:
        u = ~u & mask
    return u


def topk(keys: np.ndarray, k: int, select_max: bool) -> np.ndarray:
    """Sorted order images of the k selected keys (k capped to the key count): every
    valid top-k output's keys have exactly these images, whichever tied items it took."""
    o = topk_order_bits(keys, select_max)
    k = min(int(k), o.size)
    return np.sort(np.partition(o, k - 1)[:k]) if k > 0 else o[:0]
