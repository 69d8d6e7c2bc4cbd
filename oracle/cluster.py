"""Oracle: star-matching clustering and the cluster -> cell expansion (test infrastructure only).

Restates src/cluster.py:18-85 (cluster_cells), :88-134 (the coarse netlist) and :137-144 (expand) in numpy. The
random draws (heads, expansion offsets) come from torch's CPU generator, as in the reference, so the same seed gives
the same clusters; everything else is integer numpy.
"""
import numpy as np
import torch


def _merge_edges(parts, n):
    """[(i, j, w)] -> one edge per (i, j), sorted by (i, j), weights summed (cluster.py:18-25)."""
    key = np.concatenate([a.astype(np.int64) * n + b for a, b, _ in parts])
    w = np.concatenate([c for _, _, c in parts]).astype(np.int32)
    key, inv = np.unique(key, return_inverse=True)
    ws = np.zeros(len(key), dtype=np.int64)
    np.add.at(ws, inv.reshape(-1), w)
    return (key // n).astype(np.int32), (key % n).astype(np.int32), ws.astype(np.int32)


def _first_of_group(group, score, limit):
    """Indices of the `limit` highest scores of every group, ties in index order (cluster.py:28-35)."""
    order = np.argsort(-score, kind="stable")
    order = order[np.argsort(group[order], kind="stable")]
    g = group[order]
    at = np.arange(len(order))
    start = np.ones(len(g), dtype=bool)
    start[1:] = g[1:] != g[:-1]
    run = np.maximum.accumulate(np.where(start, at, 0)) if len(at) else at
    return order[at - run < limit]


def cluster_cells(pin_cell, pin_net, area, ratio, seed, max_degree=8, per_head=3, max_rounds=12):
    """pin_cell / pin_net int [P] grouped by net (cell len(area) = fixed pin), area int64 [n].
    -> (cluster id of every cell int64 [n], cluster areas int64 [K])."""
    pin_cell = np.asarray(pin_cell, dtype=np.int64)
    pin_net = np.asarray(pin_net, dtype=np.int64)
    area = np.asarray(area, dtype=np.int64)
    n = len(area)
    degree = np.bincount(pin_net)[pin_net]
    small = degree <= max_degree
    pin_cell, pin_net, degree = pin_cell[small], pin_net[small], degree[small]
    i = j = w = None
    for k in range(1, max_degree):
        a, b, d = pin_cell[:-k], pin_cell[k:], degree[:-k]
        m = (pin_net[:-k] == pin_net[k:]) & (a < n) & (b < n) & (a != b)
        a, b, d = a[m], b[m], 420 // (d[m] - 1)
        parts = [(a, b, d), (b, a, d)] + ([(i, j, w)] if i is not None else [])
        i, j, w = _merge_edges(parts, n)
    cid, num = np.arange(n), n
    big = 20 * int(area.sum()) // n
    small_area = area[area < big]
    # float32 arithmetic, as the reference's tensor expression 3 * ratio * long_sum // int
    cap = np.floor_divide(np.float32(3 * ratio) * np.float32(small_area.sum()), np.float32(2 * len(small_area)),
                          dtype=np.float32)
    gen = torch.Generator().manual_seed(seed)
    for _ in range(max_rounds):
        if num * ratio <= n:
            break
        head = (torch.rand(num, generator=gen) < 0.5).numpy()
        total = area[i] + area[j]
        m = ~head[i] & head[j] & (total.astype(np.float32) <= cap) & (area[i] < big) & (area[j] < big)
        tail, to = i[m], j[m]
        score = w[m].astype(np.float64) / total[m].astype(np.float64)
        best = _first_of_group(tail, score, 1)
        tail, to, score = tail[best], to[best], score[best]
        taken = _first_of_group(to, score, per_head)
        parent = np.arange(num)
        parent[tail[taken]] = to[taken]
        kept, new = np.unique(parent, return_inverse=True)
        new = new.reshape(-1)
        if len(kept) == num:
            break
        cid, num = new[cid], len(kept)
        area = _sum_int(new, area, num)
        i, j = new[i].astype(np.int32), new[j].astype(np.int32)
        keep = i != j
        i, j, w = _merge_edges([(i[keep], j[keep], w[keep])], num)
    return cid.astype(np.int64), area


def _sum_int(idx, vals, n):
    out = np.zeros(n, dtype=np.int64)
    np.add.at(out, idx, vals)
    return out


def coarse_pins(cid, num, pin_cell, pin_off, pin_net):
    """The clusters' pins (cluster.py:100-122): a merged cluster has one pin per net at its centre; a cluster of
    one cell keeps that cell's pins. -> (owner int64, off float32 [P', 2], net int64), pins grouped by net."""
    cid = np.asarray(cid, dtype=np.int64)
    pin_cell = np.asarray(pin_cell, dtype=np.int64)
    pin_net = np.asarray(pin_net, dtype=np.int64)
    single = np.bincount(cid, minlength=num) == 1
    owner = np.concatenate([cid, [num]])[pin_cell]
    keep_off = np.concatenate([single, [True]])[owner]
    key = pin_net * (num + 1) + owner
    order = np.argsort(key, kind="stable")
    sk = key[order]
    repeat = np.zeros(len(key), dtype=bool)
    repeat[order[1:]] = (sk[1:] == sk[:-1]) & ~keep_off[order[1:]]
    sel = ~repeat
    off = np.where(keep_off[sel][:, None], np.asarray(pin_off, dtype=np.float32)[sel], np.float32(0))
    return owner[sel], off.astype(np.float32), pin_net[sel]


def cluster_sizes(cid, area, num, cell_size):
    """Square of the summed area for merged clusters, the cell's own size for single ones (cluster.py:101-107)."""
    side = np.sqrt(np.asarray(area, dtype=np.float64) / 1024).astype(np.float32)
    csize = np.stack([side, side], 1)
    single = np.bincount(cid, minlength=num) == 1
    member = np.zeros(num, dtype=np.int64)
    member[cid] = np.arange(len(cid))
    csize[single] = np.asarray(cell_size, dtype=np.float32)[member[single]]
    return csize


def expand(cpos, cid, csize, num, cell_size, seed):
    """Every cell at a uniform point of its cluster's square (cluster.py:137-144). -> float32 [n + fillers, 2]."""
    n = len(cid)
    gen = torch.Generator().manual_seed(seed)
    room = np.maximum(csize[cid] - np.asarray(cell_size, dtype=np.float32), np.float32(0))
    r = (torch.rand(n, 2, generator=gen) - 0.5).numpy()
    cells = cpos[:num][cid] + r * room
    return np.concatenate([cells, cpos[num:]]).astype(np.float32)
