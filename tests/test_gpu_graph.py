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
