"""Prepared once before the optimization: the overflow grid, the fillers, the start positions.

Follows src/initializer.py:8-47 of the reference."""
import math

import numpy as np
import torch


def choose_num_bin(data, num_bin, max_bin=512):
    """0 = automatic: the largest power of 2 not above the row count, at most max_bin (initializer.py:8-13)."""
    if num_bin > 0:
        return num_bin
    rows = max(int(data.height / data.row_height), 2)
    return min(max_bin, 2 ** int(math.log2(rows)))


def compute_fillers(data, fixed_fraction, bin_area, args, log=None):
    """Fillers take the free area the cells do not need (initializer.py:16-37). May raise args.target_density to
    the utilization. -> filler sizes [n, 2] numpy."""
    placeable = float((1.0 - fixed_fraction).sum() * bin_area)
    util = data.mov_area / placeable
    if util > 1:
        raise ValueError("the movable cells need %.3f x the free area of the region: they do not fit" % util)
    if util > args.target_density:
        if log:
            log("utilization %.3f > target density %.3f: raised" % (util, args.target_density))
        args.target_density = util
    widths = np.sort(data.size[:, 0])
    n = len(widths)
    w = float(widths[int(n * 0.05):max(int(n * 0.95), int(n * 0.05) + 1)].mean())
    h = data.row_height
    count = int(round(max(args.target_density * placeable - data.mov_area, 0.0) / (w * h)))
    return np.tile(np.array([[w, h]]), (count, 1))


def init_positions(data, node_size, seed):
    """Movable cells: a Gaussian cloud (sigma 0.001 x region) at the centre; fillers uniform over the region.
    Drawn from torch's CPU generator seeded with `seed` (initializer.py:40-47). -> float32 [nodes, 2] (CPU)."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    region = torch.tensor([data.width, data.height], dtype=torch.float32)
    num_fill = len(node_size) - data.num_mov
    return torch.cat([torch.randn(data.num_mov, 2, generator=gen) * region * 0.001 + region * 0.5,
                      torch.rand(num_fill, 2, generator=gen) * region], 0)
