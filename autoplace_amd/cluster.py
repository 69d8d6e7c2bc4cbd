"""Two-level placement: the cells merged into clusters, the clusters placed first, every cell then started inside its
cluster (the behaviour of src/cluster.py:1-144 of the reference).

Clustering runs in the library's kernels (ap_cluster_*, include/autoplace_gp.h part 5): the weighted cell graph of
the small nets, then rounds of star matching. The reference draws each round's heads from torch's CPU generator; the
library makes the same draws (MT19937 seeded as torch.manual_seed, on the host, the next round's while the device
runs a round) and hands them to the device. Weights and areas are integers and a score is one IEEE division, so the
clusters are the reference's bit for bit. The coarse problem -- cluster sizes, the clusters' pins
and netlist CSR, the free-area fraction of its grid and its density layer -- is built from device tensors, inside the
GP clock as the reference builds it, without a round trip through the host.
"""
import ctypes
import threading
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from .density import ElectronicDensityLayer
from .resolution import force_num_bin
from .wirelength import Wirelength

AREA_SCALE = 1024          # areas as integers: 1 / 1024 of a site area, so sums are exact (cluster.py:94)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def merge_cap(area, big, ratio):
    """The merged-area limit of two clusters: 3 ratio / 2 times the mean area of the cells below `big`, in float32
    arithmetic as the reference's tensor expression 3 * ratio * area[area < big].sum() // count evaluates
    (cluster.py:61). -> a float exactly representable in float32."""
    below = area < big
    total = area[below].sum().cpu().float()
    return float((torch.tensor(3 * ratio, dtype=torch.float32) * total).div(float(2 * int(below.sum())),
                                                                            rounding_mode="floor"))


def cluster_cells(pin_cell, net_start, area, ratio, seed, max_degree=8, per_head=3, max_rounds=12):
    """pin_cell int32 [P] grouped by net (index len(area) = fixed pin) and net_start int32 [nets + 1] on the device;
    area int64 [n] on the device. -> (cluster of every cell int64 [n], cluster areas int64 [K]), on the device.
    Rounds stop once the clusters are `ratio` times fewer than the cells, or when a round merges nothing."""
    n, device = area.numel(), area.device
    lib = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(lib.ap_cluster_create(n, _p(pin_cell), _p(net_start), net_start.numel() - 1, _p(area), max_degree,
                                     _stream(device), ctypes.byref(h)), "cluster_create")
    try:
        big = int(20 * area.sum() // n)                # cells above 20 x the mean area stay alone
        cap = merge_cap(area, big, ratio)
        _lib.check(lib.ap_cluster_seed_heads(h, seed), "cluster_seed_heads")
        num = n
        for _ in range(max_rounds):
            if num * ratio <= n:
                break
            after = ctypes.c_int()
            _lib.check(lib.ap_cluster_round_seeded(h, cap, big, per_head, ctypes.byref(after), _stream(device)),
                       "cluster_round_seeded")
            if after.value == num:
                break
            num = after.value
        cid = torch.empty(n, dtype=torch.int64, device=device)
        carea = torch.empty(num, dtype=torch.int64, device=device)
        _lib.check(lib.ap_cluster_result(h, _p(cid), _p(carea), _stream(device)), "cluster_result")
    finally:
        lib.ap_cluster_destroy(h)
    return cid, carea


def cluster_sizes(cid, carea, cell_size):
    """A merged cluster is the square of its area; a cluster of one cell keeps the cell's size (cluster.py:97-102).
    cell_size float32 [n, 2] on the device. -> (sizes float32 [K, 2], single-cell flags [K])."""
    num = carea.numel()
    side = (carea.double() / AREA_SCALE).sqrt().float()
    size = torch.stack([side, side], 1)
    single = torch.bincount(cid, minlength=num) == 1
    member = torch.zeros(num, dtype=torch.int64, device=cid.device)
    member.scatter_(0, cid, torch.arange(cid.numel(), device=cid.device))   # a single cluster's only cell
    size[single] = cell_size[member[single]]
    return size, single


def cluster_pins(cid, single, pins):
    """The clusters' netlist (cluster.py:103-116): every pin goes to its cell's cluster; a merged cluster keeps one
    pin per net, the first in pin order, at its centre; single-cell clusters and fixed pins keep theirs.
    pins: dict(cell, off, net) of every net (fixed pin = cell n). -> (owner [P'], offset float32 [P', 2], net [P'])."""
    num = single.numel()
    owner = torch.cat([cid, cid.new_tensor([num])])[pins["cell"]]
    keeps_offset = torch.cat([single, single.new_ones(1)])[owner]
    # the first pin of each (net, owner): the smallest pin index among the pins with that key
    key = pins["net"] * (num + 1) + owner
    uniq, which = torch.unique(key, return_inverse=True)
    at = torch.arange(key.numel(), device=key.device)
    first = torch.full((uniq.numel(),), key.numel(), dtype=torch.int64, device=key.device)
    first.scatter_reduce_(0, which, at, "amin")
    keep = keeps_offset | (first[which] == at)
    off = torch.where(keeps_offset[keep].unsqueeze(1), pins["off"][keep], torch.zeros((), device=key.device))
    return owner[keep], off, pins["net"][keep]


def coarse_problem(data, node_size, wirelength, density, fixed_fraction, args, poisson="auto"):
    """The clusters' problem (cluster.py:88-134) from the cell level's device data: cid, size, wirelength, density
    and the scheduler's totals. The fillers and the blocked area are the cell level's; the grid follows the cluster
    count as the cell grid follows the cells (resolution.force_num_bin), between the overflow grid and the cells'.
    node_size / fixed_fraction: the cell level's host arrays (the device copies in `density` are used)."""
    n, device = data.num_mov, density.device
    cell_size = density.real_size[:n].float()
    area = (cell_size.prod(1).double() * AREA_SCALE).round().long()
    all_pins = wirelength.all
    cid, carea = cluster_cells(all_pins["cell"], all_pins["net_start"], area, args.cluster_ratio, args.seed)
    num = carea.numel()
    csize, single = cluster_sizes(cid, carea, cell_size)
    owner, off, net = cluster_pins(cid, single, wirelength.pins())
    cwl = Wirelength(owner, off, net, all_pins["num_nets"], num, args.ignore_net_degree, device)
    cpins = torch.bincount(owner, minlength=num + 1)[:num].double()
    cnode = torch.cat([csize.double(), density.real_size[n:]])
    judge_bin = density.judge.nbx if density.judge is not None else density.nbx
    nb = min(density.nbx, force_num_bin(num, 1.0 - density.fixed_mean, judge_bin))
    fraction = density.fraction_on_device() if nb == density.nbx else data.fixed_fraction_on(device, nb, nb)
    node_pins = torch.cat([cpins, cpins.new_zeros(cnode.shape[0] - num)])
    cden = ElectronicDensityLayer(data.width, data.height, nb, nb, fraction, density.target, cnode, num, device,
                                  node_pins, poisson=poisson)
    return SimpleNamespace(cid=cid, num=num, size=csize, cell_size=cell_size, wirelength=cwl, density=cden,
                           total_pins=float(cpins.sum()), total_area=float(cnode.prod(1).sum()),
                           num_nodes=cnode.shape[0])


class ExpansionDraws(object):
    """The expansion's uniform draws, torch.rand(n, 2) of a generator seeded with `seed` (as the reference draws them,
    cluster.py:137-144), made on a host thread into pinned memory while the device places the clusters: they depend on
    nothing else. get() waits for them."""

    def __init__(self, n, seed):
        self._out = torch.empty((n, 2), dtype=torch.float32, pin_memory=torch.cuda.is_available())
        self._seed = seed
        self._thread = threading.Thread(target=self._draw, daemon=True)
        self._thread.start()

    def _draw(self):
        torch.rand(self._out.shape, generator=torch.Generator().manual_seed(self._seed), out=self._out)

    def get(self):
        self._thread.join()
        return self._out


def expand(cpos, coarse, node_size, seed, draws=None):
    """Clusters + fillers -> cells + fillers (cluster.py:137-144): every cell at a uniform point (torch's CPU
    generator, as the reference) of the room its cluster's square leaves around it; fillers unchanged.
    draws: an ExpansionDraws of the same n and seed made beforehand (else drawn here)."""
    n, device = coarse.cid.numel(), cpos.device
    cell_size = coarse.cell_size if getattr(coarse, "cell_size", None) is not None else \
        torch.tensor(node_size[:n], dtype=torch.float32, device=device)
    u = draws.get() if draws is not None else torch.rand(n, 2, generator=torch.Generator().manual_seed(seed))
    room = (coarse.size[coarse.cid] - cell_size).clamp_(min=0)
    cells = cpos[:coarse.num][coarse.cid] + (u.to(device, non_blocking=True) - 0.5) * room
    return torch.cat([cells, cpos[coarse.num:]])
