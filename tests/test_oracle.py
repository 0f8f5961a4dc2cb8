"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it.

Independent pins:
  * the reference's own Thrust OpenMP stable_sort[_by_key] compiled from
    /root/reference (oracle/_ref) on the same inputs;
  * numpy's stable argsort on an independently computed bit-ordered key;
  * brute force over every permutation of tiny inputs;
  * the special cases of cub/test/catch2_test_radix_operations.cu:582-603 and the key
    orderings CUB documents (cub/cub/device/device_radix_sort.cuh:95-143);
  * histogram invariants and numpy bincount.
"""

import itertools

import numpy as np
import pytest

import oracle
from sortutil import float_keys_with_specials, random_keys

INT_TYPES = [np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int16, np.int32, np.int64]
FLOAT_TYPES = [np.float16, np.float32, np.float64]


def np_bit_ordered(keys: np.ndarray) -> np.ndarray:
    """Order-preserving unsigned image of keys, written with numpy independently of
    the oracle: float -> (negative ? ~bits : bits | sign) after -0 -> +0."""
    dt = keys.dtype
    u = np.dtype(f"u{dt.itemsize}")
    bits = keys.view(u).astype(np.uint64)
    width = dt.itemsize * 8
    high = np.uint64(1) << np.uint64(width - 1)
    full = np.uint64((1 << width) - 1) if width < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    if dt.kind == "f":
        bits = np.where(bits == high, np.uint64(0), bits)
        neg = (bits & high) != 0
        return np.where(neg, (~bits) & full, bits | high)
    if dt.kind == "i":
        return bits ^ high
    return bits


def np_window(o: np.ndarray, width: int, b: int, e: int) -> np.ndarray:
    if b == 0 and e == width:
        return o
    m = ((1 << (e - b)) - 1) << b
    return o & np.uint64(m)


def np_stable_perm(keys, descending, b=None, e=None):
    width = keys.dtype.itemsize * 8
    b = 0 if b is None else b
    e = width if e is None else e
    o = np_window(np_bit_ordered(keys), width, b, e)
    if descending:
        full = np.uint64((1 << width) - 1) if width < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        o = (~o) & full
    return np.argsort(o, kind="stable")


@pytest.mark.parametrize("dtype", INT_TYPES + FLOAT_TYPES)
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("entropy", ["uniform", "and3", "equal"])
def test_oracle_vs_numpy_stable_argsort(dtype, descending, entropy):
    rng = np.random.default_rng(7)
    n = 20011
    if np.dtype(dtype).kind == "f" and entropy == "uniform":
        keys = float_keys_with_specials(rng, dtype, n, zero_frac=0.05, inf_frac=0.01)
    else:
        keys = random_keys(rng, dtype, n, entropy)
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = oracle.sort(keys, vals, descending=descending)
    perm = np_stable_perm(keys, descending)
    assert np.array_equal(vo, perm.astype(np.uint32))
    assert np.array_equal(ko.view(f"u{keys.itemsize}"), keys[perm].view(f"u{keys.itemsize}"))


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32, np.uint16])
@pytest.mark.parametrize("window", [(0, 1), (3, 11), (5, 16), (8, 8)])
def test_oracle_bit_windows_vs_numpy(dtype, window):
    rng = np.random.default_rng(11)
    keys = random_keys(rng, dtype, 5000, "uniform")
    b, e = window
    for desc in (False, True):
        perm = oracle.permutation(keys, descending=desc, begin_bit=b, end_bit=e)
        assert np.array_equal(perm, np_stable_perm(keys, desc, b, e).astype(np.uint64))


def test_oracle_brute_force_tiny():
    """Unique stable order: the one permutation that is sorted and keeps ties in input order."""
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(1, 7))
        keys = rng.integers(-3, 3, n).astype(np.int8)
        for desc in (False, True):
            s = keys.astype(np.int64) * (-1 if desc else 1)
            ok = [p for p in itertools.permutations(range(n))
                  if all(s[p[i]] < s[p[i + 1]] or (s[p[i]] == s[p[i + 1]] and p[i] < p[i + 1])
                         for i in range(n - 1))]
            assert len(ok) == 1
            assert list(oracle.permutation(keys, descending=desc)) == list(ok[0])


def test_oracle_float_special_values_order():
    """-inf < -1 < -0 == +0 < 1 < inf; +-0 keep input order (device_radix_sort.cuh:95-143)."""
    keys = np.array([1.0, 0.0, -0.0, np.inf, -1.0, -np.inf, -0.0, 0.0], dtype=np.float32)
    ko, vo = oracle.sort(keys, np.arange(8, dtype=np.uint32))
    assert list(vo) == [5, 4, 1, 2, 6, 7, 0, 3]
    ko, vo = oracle.sort(keys, np.arange(8, dtype=np.uint32), descending=True)
    assert list(vo) == [3, 0, 1, 2, 6, 7, 4, 5]
    # stored bits are untouched: the -0.0 items come out as -0.0
    assert np.signbit(ko[3]) and np.signbit(ko[4])


def test_oracle_pm_zero_share_every_digit():
    """catch2_test_radix_operations.cu:582-603: -0.0 and +0.0 give equal digits."""
    for dt in FLOAT_TYPES:
        z = np.array([0.0, -0.0], dtype=dt)
        s = oracle.striped(z)
        assert s[0] == s[1]
        h = oracle.histograms(z)
        assert all(int(row.max()) == 2 for row in h)


def test_oracle_signed_order():
    keys = np.array([0, -1, 127, -128, 1, -2], dtype=np.int8)
    ko, _ = oracle.sort(keys)
    assert list(ko) == [-128, -2, -1, 0, 1, 127]


def test_oracle_nan_by_bit_pattern():
    """NaNs sort by bit pattern: +NaN above +inf, -NaN below -inf (device_radix_sort.cuh:100-101)."""
    keys = np.array([np.nan, 1.0, -np.nan, np.inf, -np.inf], dtype=np.float32)
    ko, vo = oracle.sort(keys, np.arange(5, dtype=np.uint32))
    assert list(vo) == [2, 4, 1, 3, 0]


@pytest.mark.parametrize("dtype", [np.uint32, np.float32, np.int16, np.uint64])
@pytest.mark.parametrize("descending", [False, True])
def test_oracle_histograms(dtype, descending):
    rng = np.random.default_rng(5)
    keys = random_keys(rng, dtype, 30000, "and2")
    b, e = 0, keys.itemsize * 8
    if keys.itemsize == 8:
        e = 40
    h = oracle.histograms(keys, descending=descending, begin_bit=b, end_bit=e)
    o = np_bit_ordered(keys)
    width = keys.itemsize * 8
    if descending:
        full = np.uint64((1 << width) - 1) if width < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        o = (~o) & full
    passes = (e - b + 7) // 8
    assert h.shape == (passes, 256)
    for p in range(passes):
        sh = b + 8 * p
        bits = min(8, e - sh)
        d = ((o >> np.uint64(sh)) & np.uint64((1 << bits) - 1)).astype(np.int64)
        assert np.array_equal(h[p], np.bincount(d, minlength=256).astype(np.uint64))
        assert int(h[p].sum()) == keys.size


def test_oracle_segmented_untouched_outside():
    rng = np.random.default_rng(9)
    n = 1000
    keys = random_keys(rng, np.uint32, n)
    vals = np.arange(n, dtype=np.uint32)
    sb = np.array([10, 100, 500, 500, 900])
    se = np.array([50, 400, 500, 700, 1000])
    init_k = np.full(n, 0xDEADBEEF, np.uint32)
    init_v = np.full(n, 7, np.uint32)
    ko, vo = oracle.segmented_sort(keys, vals, sb, se, init_k, init_v)
    covered = np.zeros(n, bool)
    for b, e in zip(sb, se):
        covered[b:e] = True
        seg = keys[b:e]
        perm = np.argsort(seg, kind="stable")
        assert np.array_equal(ko[b:e], seg[perm])
        assert np.array_equal(vo[b:e], vals[b:e][perm])
    assert (ko[~covered] == 0xDEADBEEF).all() and (vo[~covered] == 7).all()


# ---- against the reference's own Thrust OpenMP sort -----------------------------

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no /root/reference)")


@needs_ref
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int32, np.int64,
                                   np.float32, np.float64])
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("entropy", ["uniform", "and4", "equal"])
def test_oracle_vs_reference_thrust_omp(dtype, descending, entropy):
    rng = np.random.default_rng(123)
    n = 200003
    keys = random_keys(rng, dtype, n, entropy)
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = oracle.sort(keys, vals, descending=descending)
    k2, v2 = keys.copy(), vals.copy()
    # floats use the functor path, which keeps +-0 in input order like CUB; integer keys
    # use Thrust's default comparator (its per-thread radix path)
    mode = oracle.REF_MODE_FUNCTOR if np.dtype(dtype).kind == "f" else oracle.REF_MODE_DEFAULT
    oracle.ref_sort_inplace(k2, v2, descending=descending, mode=mode, threads=4)
    assert np.array_equal(ko.view(f"u{keys.itemsize}"), k2.view(f"u{keys.itemsize}"))
    assert np.array_equal(vo, v2)


@needs_ref
def test_oracle_vs_reference_float_specials_descending():
    """The descending f32 workload with +-0 and +-inf, against Thrust's comparator form."""
    rng = np.random.default_rng(42)
    n = 300007
    keys = float_keys_with_specials(rng, np.float32, n)
    vals = np.arange(n, dtype=np.uint32)
    ko, vo = oracle.sort(keys, vals, descending=True)
    k2, v2 = keys.copy(), vals.copy()
    oracle.ref_sort_inplace(k2, v2, descending=True, mode=oracle.REF_MODE_FUNCTOR, threads=8)
    assert np.array_equal(vo, v2)
    assert np.array_equal(ko.view(np.uint32), k2.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("window", [(0, 40), (7, 29), (0, 1)])
def test_oracle_vs_reference_bit_window(window):
    rng = np.random.default_rng(17)
    n = 100000
    keys = random_keys(rng, np.uint64, n)
    vals = np.arange(n, dtype=np.uint32)
    b, e = window
    for desc in (False, True):
        ko, vo = oracle.sort(keys, vals, descending=desc, begin_bit=b, end_bit=e)
        k2, v2 = keys.copy(), vals.copy()
        oracle.ref_sort_inplace(k2, v2, descending=desc, begin_bit=b, end_bit=e, mode=oracle.REF_MODE_WINDOW,
                                threads=4)
        assert np.array_equal(ko, k2) and np.array_equal(vo, v2)


@needs_ref
def test_reference_default_comparator_pm_zero_hazard_is_thread_dependent():
    """Documents the known hazard: Thrust's default-comparator float path orders -0.0
    before +0.0 inside each thread's tile (thrust/thrust/system/detail/sequential/
    stable_radix_sort.h:105-119) but merges them as equal; the oracle follows CUB."""
    keys = np.array([0.0, -0.0] * 64, dtype=np.float32)
    vals = np.arange(keys.size, dtype=np.uint32)
    ko, vo = oracle.sort(keys, vals)
    assert np.array_equal(vo, vals)  # CUB: all zeros equal, input order kept
    k1, v1 = keys.copy(), vals.copy()
    oracle.ref_sort_inplace(k1, v1, mode=oracle.REF_MODE_DEFAULT, threads=1)
    assert not np.array_equal(v1, vals)  # one thread: -0.0 first
