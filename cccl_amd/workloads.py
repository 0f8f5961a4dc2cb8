"""Synthetic workloads of the benchmark configurations (device-side generation).

* keys of a given entropy: AND of k uniform words (k = 1..5) or one repeated value,
  the entropy-reduction scheme of nvbench_helper/nvbench_helper/nvbench_helper.cu:344-402;
* N(0,1) floats with a share of -0.0/+0.0 and -inf/+inf;
* lognormal segment sizes (median 512, sigma 2.0, clipped to [1, 2^21]) rescaled to a
  fixed total, for the batched segmented workload.
Random streams are torch's Philox generator (device) and numpy's PCG64 (segment sizes),
seeded explicitly.
"""

from __future__ import annotations

import numpy as np
import torch

ENTROPY_WORDS = {"uniform": 1, "and2": 2, "and3": 3, "and4": 4, "and5": 5}


def int_keys(n: int, itemsize: int, entropy: str, seed: int, device) -> torch.Tensor:
    """Raw key bytes (uint8 tensor of n * itemsize) at the requested entropy."""
    g = torch.Generator(device=device).manual_seed(seed)

    def words():
        return torch.randint(-(2**31), 2**31, ((n * itemsize + 3) // 4,), dtype=torch.int32, device=device,
                             generator=g).view(torch.uint8)[: n * itemsize]

    if entropy == "equal":
        one = words()[:itemsize].clone()
        return one.view(1, itemsize).expand(n, itemsize).contiguous().view(-1)
    out = words()
    for _ in range(ENTROPY_WORDS[entropy] - 1):
        out &= words()
    return out


def u32_keys(n: int, entropy: str, seed: int, device) -> torch.Tensor:
    return int_keys(n, 4, entropy, seed, device).view(torch.uint32)


def float_keys_with_specials(n: int, seed: int, device, zero_frac=0.01, inf_frac=0.001) -> torch.Tensor:
    g = torch.Generator(device=device).manual_seed(seed)
    k = torch.randn(n, device=device, generator=g, dtype=torch.float32)
    nz, ni = int(n * zero_frac), int(n * inf_frac)
    idx = torch.randperm(n, device=device, generator=g)[: nz + ni]
    k[idx[: nz // 2]] = -0.0
    k[idx[nz // 2: nz]] = 0.0
    k[idx[nz: nz + ni // 2]] = float("-inf")
    k[idx[nz + ni // 2: nz + ni]] = float("inf")
    return k


def lognormal_sizes(num_segments: int, total: int, seed: int, median=512.0, sigma=2.0, lo=1,
                    hi=1 << 21) -> np.ndarray:
    """Sizes drawn lognormal, clipped to [lo, hi], rescaled to sum to `total` (then
    re-clipped; the remainder is spread one item at a time over random segments)."""
    rng = np.random.default_rng(seed)
    s = np.exp(np.log(median) + sigma * rng.standard_normal(num_segments))
    s = np.clip(np.rint(s), lo, hi)
    s = np.clip(np.rint(s * (total / s.sum())), lo, hi).astype(np.int64)
    diff = total - int(s.sum())
    order = rng.permutation(num_segments)
    i = 0
    while diff != 0:
        j = order[i % num_segments]
        room = (hi - s[j]) if diff > 0 else (s[j] - lo)
        take = int(min(abs(diff), room, max(1, abs(diff) // num_segments + 1)))
        if take > 0:
            s[j] += take if diff > 0 else -take
            diff += -take if diff > 0 else take
        i += 1
    return s


def offsets_from_sizes(sizes: np.ndarray, device):
    ends = np.cumsum(sizes)
    begins = ends - sizes
    return (torch.from_numpy(begins.astype(np.int32)).to(device),
            torch.from_numpy(ends.astype(np.int32)).to(device))
