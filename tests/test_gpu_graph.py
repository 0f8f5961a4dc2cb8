"""hipGraph capture/replay of sort calls: every replay must equal the eager result, with
the program's own eager work and synchronisations between replays."""

import numpy as np
import pytest

import oracle
from sortutil import lognormal_segments, random_keys, to_dev, to_host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_segmented_graph_replays_match_eager():
    import cccl_amd as ca
    from cccl_amd.graph import SortGraph

    rng = np.random.default_rng(8)
    sizes = lognormal_segments(rng, 2000, 512, 2.0, 1, 200000)
    ends = np.cumsum(sizes)
    begins = ends - sizes
    n = int(ends[-1])
    keys = random_keys(rng, np.uint32, n)
    vals = np.arange(n, dtype=np.uint32)
    kd, vd = to_dev(keys), to_dev(vals)
    ko, vo = torch.empty_like(kd), torch.empty_like(vd)
    sb, se = to_dev(begins.astype(np.int32)), to_dev(ends.astype(np.int32))
    s = ca.SegmentedSorter(torch.uint32, 4, n, sizes.size)
    ek, ev = oracle.segmented_sort(keys, vals, begins, ends, np.zeros(n, np.uint32), np.zeros(n, np.uint32))

    stream = torch.cuda.Stream()
    g = SortGraph()
    with torch.cuda.stream(stream):
        with g.capture(stream):
            s.sort(kd, ko, sb, se, vd, vo, stream=stream)
    torch.cuda.synchronize()
    for replay in range(4):
        ko.view(torch.int32).fill_(-1)  # eager work between replays
        vo.view(torch.int32).fill_(-1)
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(to_host(ko, np.uint32), ek), f"replay {replay} keys"
        assert np.array_equal(to_host(vo, np.uint32), ev), f"replay {replay} values"


def test_plain_sort_graph_replays_match_eager():
    import cccl_amd as ca
    from cccl_amd.graph import SortGraph

    rng = np.random.default_rng(9)
    n = 1 << 20
    keys = random_keys(rng, np.float32, n)
    kd = to_dev(keys)
    ko = torch.empty_like(kd)
    s = ca.Sorter(torch.float32, 0, n, descending=True)
    ek, _ = oracle.sort(keys, None, descending=True)
    stream = torch.cuda.Stream()
    g = SortGraph()
    with torch.cuda.stream(stream):
        with g.capture(stream):
            s.sort(kd, ko, stream=stream)
    for replay in range(3):
        ko.zero_()
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(to_host(ko, np.float32).view(np.uint32), ek.view(np.uint32)), f"replay {replay}"


def test_topk_graph_replays_match_eager():
    """DeviceTopK's launch sequence (state cleared by a kernel, LDS-staged output) captured
    once and replayed: every replay returns a valid top-k set (slot order may differ)."""
    import cccl_amd as ca
    from cccl_amd.graph import SortGraph

    rng = np.random.default_rng(21)
    n, k = 1 << 21, 5000
    keys = random_keys(rng, np.uint32, n, "and2")
    vals = np.arange(n, dtype=np.uint32)
    kd, vd = to_dev(keys), to_dev(vals)
    ko = torch.empty(k, dtype=kd.dtype, device="cuda")
    vo = torch.empty(k, dtype=vd.dtype, device="cuda")
    temp = torch.empty(ca.topk.temp_storage_bytes(kd, n, k, 4), dtype=torch.uint8, device="cuda")
    expect = oracle.topk(keys, k, True)
    stream = torch.cuda.Stream()
    g = SortGraph()
    with torch.cuda.stream(stream):
        with g.capture(stream):
            ca.topk.max_pairs(kd, ko, vd, vo, n, k, temp_storage=temp, stream=stream)
    torch.cuda.synchronize()
    for replay in range(4):
        ko.fill_(0)  # eager work between replays
        vo.fill_(0)
        temp.fill_(0xFF)
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        out_k, out_v = to_host(ko, np.uint32), to_host(vo, np.uint32)
        assert np.array_equal(np.sort(oracle.topk_order_bits(out_k, True)), expect), f"replay {replay} keys"
        assert np.unique(out_v).size == k and np.array_equal(keys[out_v], out_k), f"replay {replay} values"


def test_segmented_sort_c_abi_graph_replays_match_eager():
    """cccl_device_segmented_sort with one-byte keys under is_overwrite_okay (the path that
    stages the input with a device copy) captured and replayed on restored inputs."""
    import cccl_amd as ca
    from cccl_amd.graph import SortGraph

    rng = np.random.default_rng(22)
    sizes = lognormal_segments(rng, 500, 256, 2.0, 1, 50000)
    ends = np.cumsum(sizes)
    begins = ends - sizes
    n = int(ends[-1])
    keys = random_keys(rng, np.uint8, n)
    vals = np.arange(n, dtype=np.uint32)
    k0, k1 = to_dev(keys), torch.empty(n, dtype=torch.uint8, device="cuda")
    v0, v1 = to_dev(vals), torch.empty(n, dtype=torch.uint32, device="cuda")
    pristine_k, pristine_v = k0.clone(), v0.clone()
    sb, se = to_dev(begins.astype(np.int32)), to_dev(ends.astype(np.int32))
    dk, dv = ca.DoubleBuffer(k0, k1), ca.DoubleBuffer(v0, v1)
    sorter = ca.make_segmented_sort(d_in_keys=dk, d_in_values=dv, start_offsets_in=sb, end_offsets_in=se,
                                    order=ca.SortOrder.ASCENDING)
    kw = dict(d_in_keys=dk, d_out_keys=None, d_in_values=dv, d_out_values=None, num_items=n,
              num_segments=sizes.size, start_offsets_in=sb, end_offsets_in=se)
    temp = torch.empty(max(sorter(temp_storage=None, **kw), 1), dtype=torch.uint8, device="cuda")
    ek, ev = oracle.segmented_sort(keys, vals, begins, ends, keys, vals)
    stream = torch.cuda.Stream()
    g = SortGraph()
    with torch.cuda.stream(stream):
        with g.capture(stream):
            sorter(temp_storage=temp, stream=stream, **kw)
    torch.cuda.synchronize()
    assert dk.selector == 0  # one-byte keys: CUB's two 4-bit passes end in the current buffer
    for replay in range(4):
        k0.copy_(pristine_k)  # eager work between replays: restore the input
        v0.copy_(pristine_v)
        k1.fill_(0)
        torch.cuda.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert np.array_equal(to_host(k0, np.uint8), ek), f"replay {replay} keys"
        assert np.array_equal(to_host(v0, np.uint32), ev), f"replay {replay} values"
