"""Host logic of the device clustering: the library's head draws equal the reference's torch.rand(num) < 0.5 from a
manual_seed generator round after round (cluster.py:65), and the merge cap is the reference's float32 expression
(cluster.py:61) as the oracle restates it (oracle/cluster.py, whose cluster ids the reference's pin)."""
import numpy as np
import pytest
import torch


@pytest.mark.parametrize("seed", [0, 1, 5, 4242, 2 ** 32 - 1])
def test_head_draws_are_torch_rand_below_half(seed):
    """The library's head draws (host code, ap_head_draws): the stream of torch.manual_seed(seed), consumed round
    after round in pieces of shrinking size, gives torch.rand(num) < 0.5 of every round -- across MT19937's 624-word
    blocks and at the ends of pieces."""
    import ctypes
    from autoplace_amd import _lib
    gen = torch.Generator().manual_seed(seed)
    pieces = (200000, 126671, 87234, 87234, 624, 623, 50000, 3, 1, 0)     # shrinking, as the rounds go
    stream = np.empty(sum(pieces), np.uint8)
    _lib.check(_lib.lib().ap_head_draws(seed, len(stream), stream.ctypes.data_as(ctypes.c_void_p)), "head_draws")
    at = 0
    for num in pieces:
        ref = (torch.rand(num, generator=gen) < 0.5).numpy()
        assert np.array_equal(stream[at:at + num].astype(bool), ref), num
        at += num


def test_merge_cap_is_the_reference_float32_expression():
    from autoplace_amd.cluster import merge_cap
    rng = np.random.default_rng(3)
    for _ in range(20):
        area = rng.integers(1000, 60000, 5000).astype(np.int64)
        area[:3] = 10 ** 7                            # macros above `big`
        big = 20 * int(area.sum()) // len(area)
        small_area = area[area < big]
        ratio = float(rng.choice([2.0, 5.0, 7.5]))
        ref = np.floor_divide(np.float32(3 * ratio) * np.float32(small_area.sum()), np.float32(2 * len(small_area)),
                              dtype=np.float32)
        assert merge_cap(torch.tensor(area), big, ratio) == float(ref)
