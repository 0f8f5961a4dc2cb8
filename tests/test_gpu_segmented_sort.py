"""GPU parity of cub::DeviceSegmentedSort (the cccl_device_segmented_sort_* C-ABI, through
the cuda.compute-style segmented_sort / make_segmented_sort) against the CPU oracle.

Mirrors c/parallel/test/test_segmented_sort.cpp: keys-only with equal-size segments
(0 / 13 / ~1,500 segments of 1 / 12 / ~2,500 items, :198-326), key-value pairs checked
against a stable sort (:328-466), the variable segment-size pattern (:593-726), both
orders, pointer and DoubleBuffer (is_overwrite_okay) forms, plus large segments that
take the batched digit passes, int64 offsets, ±0.0 order and sorter reuse.  With
DoubleBuffer, the selector and the buffer holding the result follow CUB's
DeviceSegmentedSort rule (dispatch_segmented_sort.cuh:1293-1299): (0 + P) & 1, P =
ceil(8K / 6) passes for K-byte keys, 2 for one-byte keys.
"""

import numpy as np
import pytest

import oracle
from sortutil import float_keys_with_specials, lognormal_segments, random_keys, to_dev, to_host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FILL_K, FILL_V = 0x5A, 0x5B


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def cub_final_selector(key_bytes: int, overwrite: bool) -> int:
    if not overwrite:
        return 1
    passes = -(-8 * key_bytes // (6 if key_bytes > 1 else 4))
    return passes & 1


def gen_keys(rng, dtype, n):
    dt = np.dtype(dtype)
    if dt.kind == "f":
        return float_keys_with_specials(rng, dt, n, zero_frac=0.02, inf_frac=0.005)
    if dt.kind == "b":
        return rng.integers(0, 2, n).astype(np.bool_)
    return random_keys(rng, dt, n, "and2")


def run(keys, vals, begins, ends, descending=False, overwrite=False, offset_dtype=np.int32, sorter=None):
    """Sorts through the Python mirror; returns both key / value buffers and the selector."""
    import cccl_amd as ca

    n = keys.size
    kd = to_dev(keys)
    ko = to_dev(np.full(keys.nbytes, FILL_K, np.uint8)).view(kd.dtype)
    vd = vo = None
    if vals is not None:
        vd = to_dev(vals)
        vo = to_dev(np.full(vals.nbytes, FILL_V, np.uint8)).view(vd.dtype)
    sb = to_dev(np.asarray(begins, offset_dtype))
    se = to_dev(np.asarray(ends, offset_dtype))
    order = ca.SortOrder.DESCENDING if descending else ca.SortOrder.ASCENDING
    if overwrite:
        dk = ca.DoubleBuffer(kd, ko)
        dv = ca.DoubleBuffer(vd, vo) if vals is not None else None
        args = dict(d_in_keys=dk, d_out_keys=None, d_in_values=dv, d_out_values=None)
    else:
        args = dict(d_in_keys=kd, d_out_keys=ko, d_in_values=vd, d_out_values=vo)
    if sorter is None:
        ca.segmented_sort(**args, num_items=n, num_segments=len(begins), start_offsets_in=sb, end_offsets_in=se,
                          order=order)
    else:
        kw = dict(args, num_items=n, num_segments=len(begins), start_offsets_in=sb, end_offsets_in=se)
        nbytes = sorter(temp_storage=None, **kw)
        sorter(temp_storage=torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda"), **kw)
    torch.cuda.synchronize()
    sel = args["d_in_keys"].selector if overwrite else 1
    if overwrite and vals is not None:
        assert args["d_in_values"].selector == sel
    bk = [to_host(kd, keys.dtype), to_host(ko, keys.dtype)]
    bv = [to_host(vd, vals.dtype), to_host(vo, vals.dtype)] if vals is not None else [None, None]
    return bk, bv, sel


def check(keys, vals, begins, ends, descending=False, overwrite=False, offset_dtype=np.int32, sorter=None):
    bk, bv, sel = run(keys, vals, begins, ends, descending, overwrite, offset_dtype, sorter)
    n = keys.size
    if n == 0 or len(begins) == 0:
        assert sel == (0 if overwrite else 1)
        return
    assert sel == cub_final_selector(keys.itemsize, overwrite)
    if not overwrite:
        assert np.array_equal(bk[0].view(np.uint8), keys.view(np.uint8)), "input keys written"
        if vals is not None:
            assert np.array_equal(bv[0].view(np.uint8), vals.view(np.uint8)), "input values written"
    # items outside every segment keep what the result buffer held before the call
    init_k = keys if sel == 0 else np.full(keys.nbytes, FILL_K, np.uint8).view(keys.dtype)
    init_v = None
    if vals is not None:
        init_v = vals if sel == 0 else np.full(vals.nbytes, FILL_V, np.uint8).view(vals.dtype)
    ek, ev = oracle.segmented_sort(keys.view(np.uint8) if keys.dtype == np.bool_ else keys, vals,
                                   np.asarray(begins, np.int64), np.asarray(ends, np.int64),
                                   init_k.view(np.uint8) if keys.dtype == np.bool_ else init_k, init_v,
                                   descending=descending)
    u = f"u{keys.itemsize}"
    assert np.array_equal(bk[sel].view(u), ek.view(u)), "keys differ"
    if vals is not None:
        assert np.array_equal(bv[sel].view(np.uint8), ev.view(np.uint8)), "values differ"


ALL_KEYS = [np.uint8, np.int8, np.bool_, np.uint16, np.int16, np.float16, np.uint32, np.int32, np.float32,
            np.uint64, np.int64, np.float64]


@pytest.mark.parametrize("dtype", ALL_KEYS)
@pytest.mark.parametrize("descending,overwrite", [(False, False), (True, False), (False, True), (True, True)])
def test_keys_only_equal_segments(dtype, descending, overwrite):
    """test_segmented_sort.cpp:198-326."""
    rng = np.random.default_rng(100 + np.dtype(dtype).itemsize)
    for nseg, size in ((0, 12), (13, 1), (13, 12), (1500, 12), (13, 2500), (300, 1500)):
        n = nseg * size
        keys = gen_keys(rng, dtype, n)
        begins = np.arange(nseg, dtype=np.int64) * size
        check(keys, None, begins, begins + size, descending, overwrite)


@pytest.mark.parametrize("dtype", [np.int8, np.uint16, np.int32, np.float32, np.uint64, np.float64])
@pytest.mark.parametrize("descending,overwrite", [(False, False), (True, False), (False, True), (True, True)])
def test_pairs_are_stable(dtype, descending, overwrite):
    """test_segmented_sort.cpp:328-466: pairs against std::stable_sort per segment."""
    rng = np.random.default_rng(7)
    nseg, size = 400, 700
    keys = gen_keys(rng, dtype, nseg * size)
    if np.dtype(dtype).kind in "iu":
        keys = (keys % 16).astype(dtype)  # many equal keys
    vals = np.arange(keys.size, dtype=np.uint32)
    begins = np.arange(nseg, dtype=np.int64) * size
    check(keys, vals, begins, begins + size, descending, overwrite)


@pytest.mark.parametrize("nseg", [20, 600])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.int64])
@pytest.mark.parametrize("descending,overwrite", [(False, False), (True, True)])
def test_variable_segment_sizes(nseg, dtype, descending, overwrite):
    """test_segmented_sort.cpp:593-726 (the same size pattern), float values."""
    pattern = [1, 5, 10, 20, 30, 50, 100, 3, 25, 600, 7, 18, 300, 4, 35, 9, 14, 700, 28, 11]
    sizes = np.array((pattern * (nseg // len(pattern) + 1))[:nseg])
    rng = np.random.default_rng(nseg)
    keys = gen_keys(rng, dtype, int(sizes.sum()))
    vals = rng.standard_normal(keys.size).astype(np.float32)
    ends = np.cumsum(sizes)
    check(keys, vals, ends - sizes, ends, descending, overwrite)


@pytest.mark.parametrize("dtype,vdtype", [(np.uint32, np.uint32), (np.uint8, np.uint32), (np.float64, None),
                                          (np.int16, np.uint64)])
@pytest.mark.parametrize("overwrite", [False, True])
def test_large_segments_take_the_batched_passes(dtype, vdtype, overwrite):
    """Lognormal sizes up to 2^18 items: one-wave, one-workgroup and batched classes mixed."""
    rng = np.random.default_rng(3)
    sizes = lognormal_segments(rng, 300, 512, 2.0, 1, 1 << 18)
    keys = gen_keys(rng, dtype, int(sizes.sum()))
    vals = None if vdtype is None else np.arange(keys.size, dtype=vdtype)
    ends = np.cumsum(sizes)
    check(keys, vals, ends - sizes, ends, descending=bool(overwrite), overwrite=overwrite)


@pytest.mark.parametrize("overwrite", [False, True])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.uint64])
def test_gaps_and_empty_segments_int64_offsets(overwrite, dtype):
    """Segments that leave gaps, overlap nothing, and include empty ones; items outside
    every segment stay untouched in the result buffer (for one-byte keys under
    is_overwrite_okay that is the input buffer)."""
    rng = np.random.default_rng(21)
    n = 50_000
    keys = gen_keys(rng, dtype, n)
    vals = np.arange(n, dtype=np.uint32)
    begins = np.array([0, 100, 100, 5000, 20000, 49999], np.int64)
    ends = np.array([50, 100, 3000, 19000, 20000, 50000], np.int64)
    check(keys, vals, begins, ends, False, overwrite, offset_dtype=np.int64)


def test_signed_zero_keeps_input_order():
    """-0.0 and +0.0 compare equal (`<`): they keep input order, bits preserved."""
    keys = np.array([0.0, -0.0, 1.0, -0.0, 0.0, -1.0, -0.0] * 300, np.float32)
    vals = np.arange(keys.size, dtype=np.uint32)
    for desc in (False, True):
        check(keys, vals, [0, 700], [700, keys.size], descending=desc)


def test_sorter_reuse_and_empty_input():
    import cccl_amd as ca

    rng = np.random.default_rng(5)
    keys = gen_keys(rng, np.int32, 20_000)
    vals = np.arange(keys.size, dtype=np.uint32)
    kd, vd = to_dev(keys), to_dev(vals)
    sb = to_dev(np.array([0, 7000], np.int32))
    se = to_dev(np.array([7000, 20000], np.int32))
    sorter = ca.make_segmented_sort(d_in_keys=kd, d_out_keys=torch.empty_like(kd), d_in_values=vd,
                                    d_out_values=torch.empty_like(vd), start_offsets_in=sb, end_offsets_in=se,
                                    order=ca.SortOrder.ASCENDING)
    for seed in (1, 2):
        k = gen_keys(np.random.default_rng(seed), np.int32, 20_000)
        check(k, vals, [0, 7000], [7000, 20000], sorter=sorter)
    check(np.zeros(0, np.int32), np.zeros(0, np.uint32), [], [], sorter=sorter)
    check(np.zeros(0, np.int32), None, [], [], overwrite=True)
