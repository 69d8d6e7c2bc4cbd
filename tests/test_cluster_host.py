"""Host logic of the device clustering: the head draws equal the reference's torch.rand(num) < 0.5 from a
manual_seed generator round after round (cluster.py:65), and the merge cap is the reference's float32 expression
(cluster.py:61) as the oracle restates it (oracle/cluster.py, whose cluster ids the reference's pin)."""
import numpy as np
import pytest
import torch


@pytest.mark.parametrize("seed", [0, 1, 5, 4242])
def test_head_draws_are_torch_rand_below_half(seed):
    from autoplace_amd.cluster import head_draws
    gen = torch.Generator().manual_seed(seed)
    draws = head_draws(seed)
    next(draws)
    for num in (200000, 126671, 87234, 87234, 50000, 3, 1):      # shrinking, as the rounds go
        ref = (torch.rand(num, generator=gen) < 0.5).numpy()
        mine = draws.send(num)
        assert mine.dtype == bool and np.array_equal(mine, ref), num


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
