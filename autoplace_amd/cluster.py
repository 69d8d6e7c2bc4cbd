"""Two-level placement: cells merged into clusters, the clusters placed first, then every cell starts inside its
cluster (src/cluster.py:1-144 of the reference).

Clustering on the device, in rounds of star matching: half of the clusters, drawn from torch's CPU generator, are
heads; every other cluster joins the head it shares the most net weight with per unit of merged area (a net of d pins
gives 420 / (d - 1) to each pair of its cells, an integer for d <= 8), a head taking at most `per_head`. Weights and
areas are integers, so the same input and seed give the same clusters on any device.
"""
import math
from types import SimpleNamespace

import numpy as np
import torch

from .density import ElectronicDensityLayer
from .resolution import force_num_bin
from .wirelength import Wirelength


def _merge_edges(parts, n):
    """[(i, j, w)] -> one edge per (i, j), sorted by (i, j), weights summed; int32 out."""
    key = torch.cat([a.long() * n + b.long() for a, b, _ in parts])
    w = torch.cat([c for _, _, c in parts]).int()
    key, inv = torch.unique(key, return_inverse=True)
    w = torch.zeros(len(key), dtype=torch.int32, device=w.device).index_add_(0, inv, w)
    return (key // n).int(), (key % n).int(), w


def _top_per_group(group, score, limit):
    """Indices of the `limit` highest scores of every group (ties: lower index first)."""
    order = torch.argsort(score, descending=True, stable=True)
    order = order[torch.argsort(group[order], stable=True)]
    g = group[order]
    at = torch.arange(len(order), device=group.device)
    start = torch.ones_like(g, dtype=torch.bool)
    start[1:] = g[1:] != g[:-1]
    run_start = torch.cummax(torch.where(start, at, torch.zeros_like(at)), 0)[0]
    return order[at - run_start < limit]


def cluster_cells(pin_cell, pin_net, area, ratio, seed, max_degree=8, per_head=3, max_rounds=12):
    """pin_cell / pin_net [P] grouped by net (cell len(area) = fixed pin); area [n] long.
    -> (cluster of every cell [n] long, cluster areas [K] long)."""
    n, device = len(area), area.device
    pin_cell, pin_net = pin_cell.long(), pin_net.long()
    degree = torch.bincount(pin_net)[pin_net]
    small = degree <= max_degree
    pin_cell, pin_net, degree = pin_cell[small], pin_net[small], degree[small]
    i = j = w = None
    for k in range(1, max_degree):
        a, b, d = pin_cell[:-k], pin_cell[k:], degree[:-k]
        m = (pin_net[:-k] == pin_net[k:]) & (a < n) & (b < n) & (a != b)
        a, b, d = a[m], b[m], 420 // (d[m] - 1)
        parts = [(a, b, d), (b, a, d)] + ([(i, j, w)] if i is not None else [])
        i, j, w = _merge_edges(parts, n)
    del pin_cell, pin_net, degree
    cid, num = torch.arange(n, device=device), n
    big = 20 * area.sum() // n
    small_area = area[area < big]
    # float32, as the tensor expression 3 * ratio * long_sum // count evaluates
    cap = (torch.tensor(3 * ratio, dtype=torch.float32) * small_area.sum().cpu().float()) \
        .div(float(2 * len(small_area)), rounding_mode="floor").to(device)
    gen = torch.Generator().manual_seed(seed)
    for _ in range(max_rounds):
        if num * ratio <= n:
            break
        head = (torch.rand(num, generator=gen) < 0.5).to(device)
        total = area[i] + area[j]
        m = ~head[i] & head[j] & (total.float() <= cap) & (area[i] < big) & (area[j] < big)
        tail, to, score = i[m], j[m], w[m].double() / total[m].double()
        best = _top_per_group(tail, score, 1)
        tail, to, score = tail[best], to[best], score[best]
        taken = _top_per_group(to, score, per_head)
        parent = torch.arange(num, device=device)
        parent[tail[taken].long()] = to[taken].long()
        kept, new = torch.unique(parent, return_inverse=True)
        if len(kept) == num:
            break
        cid, num = new[cid], len(kept)
        area = torch.zeros(num, dtype=torch.long, device=device).index_add_(0, new, area)
        i, j = new[i.long()].int(), new[j.long()].int()
        keep = i != j
        i, j, w = _merge_edges([(i[keep], j[keep], w[keep])], num)
    return cid, area


def coarse_problem(data, node_size, wirelength, density, fixed_fraction, args, poisson="auto"):
    """The problem of the clusters (cluster.py:88-134): cid, size, wirelength, density and the scheduler's totals.
    A merged cluster is a square of its area with one pin per net at its centre; a single cell keeps its shape and
    pins. Same fillers and blockages; grid from the cluster count (force_num_bin), between the overflow grid and the
    cells' grid."""
    n, device = data.num_mov, density.device
    size = torch.tensor(node_size[:n], dtype=torch.float32, device=device)
    scale = 1024
    cid, area = cluster_cells(wirelength.all["cell"], wirelength.all["net"],
                              (size.prod(1).double() * scale).round().long(), args.cluster_ratio, args.seed)
    pins = wirelength.pins()
    num = len(area)
    side = (area.double() / scale).sqrt().float()
    csize = torch.stack([side, side], 1)
    single = torch.bincount(cid, minlength=num) == 1
    member = torch.zeros(num, dtype=torch.long, device=device)
    member[cid] = torch.arange(n, device=device)
    csize[single] = size[member[single]]
    owner = torch.cat([cid, cid.new_tensor([num])])[pins["cell"]]
    keep_off = torch.cat([single, single.new_tensor([True])])[owner]
    key = pins["net"] * (num + 1) + owner
    order = torch.argsort(key, stable=True)
    sorted_key = key[order]
    repeat = torch.zeros_like(keep_off)
    repeat[order[1:]] = (sorted_key[1:] == sorted_key[:-1]) & ~keep_off[order[1:]]
    sel = ~repeat
    off = torch.where(keep_off[sel].unsqueeze(1), pins["off"][sel], torch.zeros(1, 2, device=device))
    c_owner, c_net = owner[sel], pins["net"][sel]
    cwl = Wirelength(c_owner, off, c_net, pins["num_nets"], num, args.ignore_net_degree, device)   # CSR on the device
    cpins = torch.bincount(c_owner, minlength=num + 1)[:num].double().cpu().numpy()
    cnode = np.concatenate([csize.double().cpu().numpy(), node_size[n:]])
    judge_bin = density.judge.nbx if density.judge is not None else density.nbx
    nb = min(density.nbx, force_num_bin(num, 1.0 - float(fixed_fraction.mean()), judge_bin))
    if nb != density.nbx:
        fixed_fraction = data.fixed_fraction(nb, nb)
    cden = ElectronicDensityLayer(data.width, data.height, nb, nb, fixed_fraction, density.target, cnode, num, device,
                                  np.concatenate([cpins, np.zeros(len(cnode) - num)]), poisson=poisson)
    return SimpleNamespace(cid=cid, num=num, size=csize, wirelength=cwl, density=cden, total_pins=float(cpins.sum()),
                           total_area=float(cnode.prod(1).sum()), num_nodes=len(cnode))


def expand(cpos, coarse, node_size, seed):
    """Cells at uniform points of their cluster's square (cluster.py:137-144); fillers unchanged."""
    n, device = len(coarse.cid), cpos.device
    gen = torch.Generator().manual_seed(seed)
    room = (coarse.size[coarse.cid] - torch.tensor(node_size[:n], dtype=torch.float32, device=device)).clamp_(min=0)
    cells = cpos[:coarse.num][coarse.cid] + (torch.rand(n, 2, generator=gen) - 0.5).to(device) * room
    return torch.cat([cells, cpos[coarse.num:]])
