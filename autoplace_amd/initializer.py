"""Once-per-run preparation: the size of the overflow grid, the filler cells, the start placement.

Behaviour of src/initializer.py:8-47 (the reference). The results are pinned bit for bit by the golden fixtures
(filler count and size, start positions), so the arithmetic below is the reference's: the same float64 sums and
rounding, and the start drawn from torch's seeded CPU generator in the same order (cells' normals first, then the
fillers' uniforms).
"""
import numpy as np
import torch


def choose_num_bin(data, num_bin, max_bin=512):
    """The overflow grid's side: `num_bin` when given (> 0), else the largest power of two at or below the number
    of rows (at least 2), capped at max_bin -- so a bin is never shorter than a row."""
    if num_bin > 0:
        return num_bin
    rows = max(int(data.height / data.row_height), 2)
    return min(max_bin, 1 << (rows.bit_length() - 1))


def _filler_width(widths):
    """Mean cell width over the middle 90 % of the sorted widths (at least one width)."""
    w = np.sort(widths)
    lo = int(len(w) * 0.05)
    hi = max(int(len(w) * 0.95), lo + 1)
    return float(w[lo:hi].mean())


def compute_fillers(data, fixed_fraction, bin_area, args, log=None):
    """Filler cells that bring the free area of every bin up to the target density.

    The free area is sum(1 - fixed_fraction) x bin_area. Cells that fill more of it than the target raise
    args.target_density to their utilization; cells that do not fit raise ValueError. Fillers are one row tall and
    as wide as a typical cell; their count covers target x free area - cell area. -> sizes [count, 2] (numpy)."""
    free_area = float((1.0 - fixed_fraction).sum() * bin_area)
    util = data.mov_area / free_area
    if util > 1:
        raise ValueError("cell area is %.3f times the area the fixed objects leave free: no placement exists"
                         % util)
    if util > args.target_density:
        if log:
            log("target density %.3f below the utilization %.3f: raised to it" % (args.target_density, util))
        args.target_density = util
    width, height = _filler_width(data.size[:, 0]), data.row_height
    room = max(args.target_density * free_area - data.mov_area, 0.0)
    count = int(round(room / (width * height)))
    return np.repeat(np.array([[width, height]]), count, axis=0)


def init_positions(data, node_size, seed):
    """The start placement, float32 [nodes, 2] on the CPU: the cells in a normal cloud of 0.1 % of the region
    around its centre, the fillers spread uniformly over the region; drawn from a torch CPU generator seeded with
    `seed`, the cells' normals before the fillers' uniforms."""
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    extent = torch.tensor([data.width, data.height], dtype=torch.float32)
    cells, fillers = data.num_mov, len(node_size) - data.num_mov
    cloud = torch.randn(cells, 2, generator=gen)
    spread = torch.rand(fillers, 2, generator=gen)
    return torch.cat([cloud * extent * 0.001 + extent * 0.5, spread * extent])
