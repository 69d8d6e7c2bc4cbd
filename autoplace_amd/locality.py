"""Spatial renumbering of a level's nodes and nets for the device engine (ap_level_renumber in the library).

The engine's scatter, WA and gather kernels read node data by index. In the problem's own order, the nodes of a
workgroup (and the cells of a net) are spread over the whole die, so every access is a separate cache line and every
density contribution a separate global atomic. Renumbering cells (and, separately, fillers) along the Hilbert curve of
their bins at the start of a level makes index neighbours spatial neighbours; nets are renumbered by their first
cell in the new order so that the pins of neighbouring nets, and the pins a cell gathers, are close too.

The renumbering is a permutation: the integer density sums do not change, per-net and per-cell sums keep their pin
order within the net / cell, so results change only by the order of FP64 norm sums. It is deterministic (stable
radix sorts of integer keys, on the device).
"""
import ctypes

import torch

from . import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def renumber_level(pos, nbx, nby, bx, by, num_mov, cell32, off, net_start):
    """pos f32 [nodes, 2]; the optimized-net CSR cell32 i32 [P] (num_mov = fixed pin), off f32 [P, 2], net_start
    i32 [nets + 1], all on the device. -> (perm int64 [nodes], new -> old: cells then fillers, each along the Hilbert
    curve of their bin on the nbx x nby grid of bx x by bins; dict(cell32, off, net_start, cell_start, cell_pin,
    num_nets) of the renumbered netlist, cell_pin holding every pin's entry, the fixed pins' after
    cell_start[num_mov])."""
    dev = pos.device
    for t, dt in ((pos, torch.float32), (cell32, torch.int32), (off, torch.float32), (net_start, torch.int32)):
        if t.dtype != dt or not t.is_contiguous() or t.device != dev:
            raise ValueError("renumber_level: %s contiguous on %s expected" % (dt, dev))
    n_nodes, n_pins, n_nets = pos.shape[0], cell32.numel(), net_start.numel() - 1
    perm = torch.empty(n_nodes, dtype=torch.int32, device=dev)
    cell = torch.empty(n_pins, dtype=torch.int32, device=dev)
    o = torch.empty((n_pins, 2), dtype=torch.float32, device=dev)
    start = torch.empty(n_nets + 1, dtype=torch.int32, device=dev)
    cell_start = torch.empty(num_mov + 1, dtype=torch.int32, device=dev)
    cell_pin = torch.empty(max(n_pins, 1), dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(_lib.lib().ap_level_renumber(_p(pos), n_nodes, num_mov, nbx, nby, float(bx), float(by), _p(cell32),
                                             _p(off), _p(net_start), n_nets, n_pins, _p(perm), _p(cell), _p(o),
                                             _p(start), _p(cell_start), _p(cell_pin), stream), "level_renumber")
    return perm.long(), {"cell32": cell, "off": o, "net_start": start, "cell_start": cell_start,
                         "cell_pin": cell_pin, "num_nets": n_nets}
