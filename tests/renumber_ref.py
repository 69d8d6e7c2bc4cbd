"""The spatial renumbering of a level in torch operators: the statement the library's ap_level_renumber is checked
against (bit for bit, tests/test_gpu_renumber.py), and a Morton variant for comparison.

Cells (and, separately, fillers) are ordered along a space-filling curve of their bins; nets by their first cell in
the new order; the pins of every cell follow. Every sort is stable and of integer keys, so the permutation is one.
"""
import torch


def _spread_bits(v):
    v = v.long() & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def morton_order(pos, nbx, nby, bx, by, num_mov):
    """Permutation new -> old of the nodes: cells, then fillers, each sorted by the Morton code of their bin."""
    ix, iy = _bins(pos, nbx, nby, bx, by)
    key = _spread_bits(ix) | (_spread_bits(iy) << 1)
    cells = torch.argsort(key[:num_mov], stable=True)
    fill = torch.argsort(key[num_mov:], stable=True) + num_mov
    return torch.cat([cells, fill])


def hilbert_key(ix, iy, order):
    """Index of the cells (ix, iy) along the Hilbert curve of a 2^order x 2^order grid (the classic xy -> d walk,
    vectorised). Consecutive indices are edge neighbours, so any run of indices covers a compact region."""
    x, y = ix.long().clone(), iy.long().clone()
    d = torch.zeros_like(x)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((x & s) > 0).long()
        ry = ((y & s) > 0).long()
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant
        flip = ry == 0
        swap_x = torch.where(flip & (rx == 1), s - 1 - x, x)
        swap_y = torch.where(flip & (rx == 1), s - 1 - y, y)
        x = torch.where(flip, swap_y, swap_x)
        y = torch.where(flip, swap_x, swap_y)
        s >>= 1
    return d


def _bins(pos, nbx, nby, bx, by):
    """Bin of every node: floor(x * (1 / bx)) in single precision (the reciprocal rounded to float32 first), clamped
    to the grid."""
    one = torch.tensor(1.0, dtype=torch.float32, device=pos.device)
    ibx = one / torch.tensor(bx, dtype=torch.float32, device=pos.device)
    iby = one / torch.tensor(by, dtype=torch.float32, device=pos.device)
    return (pos[:, 0] * ibx).floor().clamp_(0, nbx - 1), (pos[:, 1] * iby).floor().clamp_(0, nby - 1)


def hilbert_order(pos, nbx, nby, bx, by, num_mov):
    """Permutation new -> old of the nodes: cells, then fillers, each sorted along the Hilbert curve of their bin."""
    ix, iy = _bins(pos, nbx, nby, bx, by)
    order = max(1, int(max(nbx, nby) - 1).bit_length())
    key = hilbert_key(ix, iy, order)
    cells = torch.argsort(key[:num_mov], stable=True)
    fill = torch.argsort(key[num_mov:], stable=True) + num_mov
    return torch.cat([cells, fill])


def renumber_netlist(cell32, off, net_start, num_mov, perm):
    """Pins grouped by net (optimized-net CSR) -> the same netlist with cells renumbered by `perm` (new -> old) and
    nets ordered by their smallest new cell index. -> dict(cell32, off, net_start, cell_start, cell_pin)."""
    dev = cell32.device
    inv = torch.empty(num_mov + 1, dtype=torch.long, device=dev)
    inv[perm[:num_mov]] = torch.arange(num_mov, device=dev)
    inv[num_mov] = num_mov                                     # a fixed pin stays a fixed pin
    new_cell = inv[cell32.long()]
    ns = net_start.long()
    deg = ns[1:] - ns[:-1]
    nets = len(deg)
    pins = len(cell32)
    net_of = torch.repeat_interleave(torch.arange(nets, device=dev), deg)
    key = torch.full((nets,), num_mov + 1, dtype=torch.long, device=dev)
    key.scatter_reduce_(0, net_of, new_cell, "amin", include_self=True)
    order = torch.argsort(key, stable=True)
    deg_new = deg[order]
    start_new = torch.zeros(nets + 1, dtype=torch.long, device=dev)
    start_new[1:] = torch.cumsum(deg_new, 0)
    src = torch.repeat_interleave(ns[:-1][order], deg_new) + \
        (torch.arange(pins, device=dev) - torch.repeat_interleave(start_new[:-1], deg_new))
    c = new_cell[src]
    o = off[src].contiguous()
    by_cell = torch.argsort(c, stable=True)
    cell_start = torch.searchsorted(c[by_cell], torch.arange(num_mov + 1, device=dev))
    return {"cell32": c.int().contiguous(), "off": o, "net_start": start_new.int().contiguous(),
            "cell_start": cell_start.int().contiguous(),
            "cell_pin": by_cell[:int(cell_start[-1])].int().contiguous(), "num_nets": nets}
