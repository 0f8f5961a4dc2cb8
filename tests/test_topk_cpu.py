"""cub::DeviceTopK without a GPU: the oracle's order restatement pinned three ways, and
the entry point's temp-storage query and argument checks.

Pins of oracle.topk_order_bits / oracle.topk:
* against the oracle's radix sort (itself pinned against the reference's Thrust): for
  keys without -0.0 the k largest are the first k of the descending stable sort;
* brute force on small arrays with an order written from the values themselves
  (Python comparison of numbers, -0.0 below +0.0), independent of the bit transform;
* the special cases cub/test/catch2_test_device_topk_keys.cu exercises: k = num_items,
  k > num_items (capped), all-equal keys.
"""

import ctypes
import math

import numpy as np
import pytest

import oracle
from cccl_amd import _bindings as B
from sortutil import float_keys_with_specials, random_keys

DTYPES = [np.uint8, np.int8, np.uint16, np.int16, np.float16, np.uint32, np.int32, np.float32, np.uint64, np.int64,
          np.float64]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("select_max", [True, False])
def test_topk_matches_the_sort_oracle(dtype, select_max):
    rng = np.random.default_rng(np.dtype(dtype).itemsize * 7 + select_max)
    n = 5000
    if np.dtype(dtype).kind == "f":
        keys = rng.standard_normal(n).astype(dtype)
        keys[keys == 0] = 1  # no -0.0: the sort merges the zeros, top-k does not
    else:
        keys = random_keys(rng, dtype, n, "and2")
    for k in (1, 17, 2500, n - 1, n, n + 5):
        ek, _ = oracle.sort(keys, descending=select_max)
        expect = np.sort(oracle.topk_order_bits(ek[: min(k, n)], select_max))
        assert np.array_equal(oracle.topk(keys, k, select_max), expect)


def _value_order_key(x, select_max):
    """Order by numeric value, -0.0 below +0.0; Max selects from the top."""
    x = float(x) if isinstance(x, (float, np.floating)) else int(x)
    sign = math.copysign(1.0, x) if isinstance(x, float) else 0
    key = (x, sign)
    return tuple(-v for v in key) if select_max else key


@pytest.mark.parametrize("dtype", [np.int8, np.uint16, np.float32, np.float64, np.int64])
def test_topk_brute_force_small(dtype):
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(1, 12))
        if np.dtype(dtype).kind == "f":
            keys = np.array(rng.choice([-2.5, -1.0, -0.0, 0.0, 0.5, 3.0, np.inf, -np.inf], n), dtype)
        else:
            keys = random_keys(rng, dtype, n, "uniform")
        for select_max in (True, False):
            for k in range(0, n + 2):
                brute = sorted(keys.tolist(), key=lambda v: _value_order_key(v, select_max))[: min(k, n)]
                expect = np.sort(oracle.topk_order_bits(np.array(brute, dtype), select_max))
                assert np.array_equal(oracle.topk(keys, k, select_max), expect), (keys, k, select_max)


def test_topk_special_cases():
    keys = np.array([0.0, -0.0, -0.0, 0.0, 1.0], np.float32)
    # Max-2: 1.0 and one +0.0 (+0.0 ranks above -0.0); Min-2: both -0.0
    assert np.array_equal(oracle.topk(keys, 2, True), np.sort(oracle.topk_order_bits(np.array([1.0, 0.0], np.float32), True)))
    assert np.array_equal(oracle.topk(keys, 2, False),
                          oracle.topk_order_bits(np.array([-0.0, -0.0], np.float32), False))
    same = np.full(100, 7, np.int32)
    assert np.array_equal(oracle.topk(same, 30, True), np.full(30, oracle.topk_order_bits(same[:1], True)[0]))
    nan = np.array([np.nan, -np.nan, 1.0, -1.0], np.float32)
    # positive NaN ranks above +inf and every number; negative NaN below everything
    assert oracle.topk(nan, 1, True)[0] == oracle.topk_order_bits(nan[:1], True)[0]
    assert oracle.topk(nan, 1, False)[0] == oracle.topk_order_bits(nan[1:2], False)[0]


def test_topk_pin_catches_a_dropped_negative_flip():
    """A transform that only flips the sign bit of floats (a plausible mistake) orders
    negative floats backwards: its Min-1 pick differs from the value order's, which the
    oracle follows."""
    keys = np.array([-1.0, -2.0, 3.0], np.float32)
    wrong_pick = keys[np.argmin(keys.view(np.uint32) ^ np.uint32(1 << 31))]
    brute_pick = min(keys.tolist())
    assert wrong_pick != brute_pick
    assert oracle.topk(keys, 1, False)[0] == oracle.topk_order_bits(np.array([brute_pick], np.float32), False)[0]


def _query(key_type, value_bytes, n, k):
    nbytes = ctypes.c_size_t(0)
    rc = B.lib().cccl_amd_topk(None, ctypes.byref(nbytes), None, None, None, None, key_type, value_bytes, n, k, 1,
                               None)
    return rc, nbytes.value


def test_entry_point_query_and_rejections():
    for kt in (B.CCCL_UINT8, B.CCCL_INT16, B.CCCL_FLOAT32, B.CCCL_UINT64, B.CCCL_BOOLEAN):
        for vb in (0, 1, 4, 16):
            rc, nbytes = _query(kt, vb, 1 << 20, 100)
            assert rc == 0 and nbytes > 0
    assert _query(B.CCCL_UINT32, 0, 0, 0)[0] == 0 and _query(B.CCCL_UINT32, 0, 0, 0)[1] > 0
    # the candidate buffer of the filter pass: keys + 4-byte indices for a sixteenth of the input
    n = 1 << 24
    assert _query(B.CCCL_UINT32, 4, n, 3)[1] >= (n // 16) * (4 + 4)
    assert _query(B.CCCL_UINT8, 4, n, 3)[1] < 4096  # one digit pass: no filter, no buffer
    assert _query(B.CCCL_UINT32, 3, 100, 3)[0] == 801  # value width
    assert _query(B.CCCL_STORAGE, 0, 100, 3)[0] == 801  # key type
    assert _query(B.CCCL_UINT32, 0, 1 << 32, 3)[0] == 1  # 32-bit counters
    assert B.lib().cccl_amd_topk(None, None, None, None, None, None, B.CCCL_UINT32, 0, 10, 3, 1, None) == 1


def test_python_api_exports():
    import cccl_amd

    for name in ("max_keys", "min_keys", "max_pairs", "min_pairs", "temp_storage_bytes"):
        assert callable(getattr(cccl_amd.topk, name))
    assert "cccl_amd_topk" in B.EXPORTED_SYMBOLS


def test_float_specials_generator_has_both_zeros():
    k = float_keys_with_specials(np.random.default_rng(0), np.float32, 10000, zero_frac=0.05)
    bits = k.view(np.uint32)
    assert (bits == 0x80000000).any() and (bits == 0).any()
