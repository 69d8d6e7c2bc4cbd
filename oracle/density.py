"""Oracle: electrostatic density map, overflow and the field gather (test infrastructure only).

Restates the arithmetic of cpp_to_py/gp_cuda_kernel.cu:67-139 in numpy:
    span / overlap       :73-82   a node's extent in bin units, its first / last bin, the overlap with bin b
    density_map_kernel   :84-95   fixed point: u64((double)(ratio ovx ovy) * 2^32) added per bin, cells -> map 0,
                                  fillers -> map 1 (integer sums: order-free)
    density_post_kernel  :130-139 mov = map0 / 2^32 + fixed, total = mov + map1 / 2^32, over = max(mov - target, 0)
    density_grad_kernel  :102-125 grad = (wl_grad + weight * -(sum_b w_b field_b)) / max(1, pins + precond area)
and the layer set-up of src/core/electronic_density_layer.py:34-77 (smoothing to sqrt(2) bins, density ratio,
fixed map = fixed_fraction x target, position bounds) and the overflow of :96-120.
`exact=True` gives the same maps in float64 without fixed-point rounding (the FP64 error reference).
"""
import math

import numpy as np

SCALE = 4294967296.0


def smooth_nodes(node_size, bin_x, bin_y):
    """electronic_density_layer.py:48-55 -> (smoothed size float32 [n, 2], ratio float32 [n], pos_lo, pos_hi)."""
    real = np.asarray(node_size, dtype=np.float64)
    smooth = np.maximum(real, math.sqrt(2) * np.array([bin_x, bin_y]))
    ratio = real.prod(1) / smooth.prod(1)
    return smooth.astype(np.float32), ratio.astype(np.float32)


def _spans(c, size, inv_b, nb, dtype):
    c = c.astype(dtype)
    size = size.astype(dtype)
    lo = (c - dtype(0.5) * size) * dtype(inv_b)
    hi = (c + dtype(0.5) * size) * dtype(inv_b)
    b0 = np.maximum(np.floor(lo).astype(np.int64), 0)
    b1 = np.minimum(np.floor(hi).astype(np.int64), nb - 1)
    return lo, hi, b0, b1


def _overlap(lo, hi, b, dtype):
    bf = b.astype(dtype)
    return np.maximum(np.minimum(hi, bf + dtype(1)) - np.maximum(lo, bf), dtype(0))


def node_bin_weights(pos, size, ratio, bx, by, nbx, nby, dtype=np.float32):
    """Every (node, bin, weight) with weight = ratio ovx ovy > 0 as the kernel computes it (float32 products).

    -> (node int64 [Q], flat bin bx * nby + by int64 [Q], weight dtype [Q]). Vectorized over the bins of each
    node's span: nodes are grouped by span size."""
    dt = np.dtype(dtype).type
    pos = np.asarray(pos)
    inv_bx = dt(1) / dt(bx)
    inv_by = dt(1) / dt(by)
    lx, hx, x0, x1 = _spans(pos[:, 0], size[:, 0], inv_bx, nbx, dt)
    ly, hy, y0, y1 = _spans(pos[:, 1], size[:, 1], inv_by, nby, dt)
    nx = np.maximum(x1 - x0 + 1, 0)
    ny = np.maximum(y1 - y0 + 1, 0)
    ratio = np.asarray(ratio).astype(dt)
    nodes, bins, ws = [], [], []
    for kx in range(int(nx.max()) if len(nx) else 0):
        mx = nx > kx
        if not mx.any():
            continue
        ix = np.nonzero(mx)[0]
        bxs = x0[ix] + kx
        wx = ratio[ix] * _overlap(lx[ix], hx[ix], bxs, dt)
        for ky in range(int(ny[ix].max())):
            m = ny[ix] > ky
            if not m.any():
                continue
            jx = ix[m]
            bys = y0[jx] + ky
            w = wx[m] * _overlap(ly[jx], hy[jx], bys, dt)
            nodes.append(jx)
            bins.append(bxs[m] * nby + bys)
            ws.append(w)
    if not nodes:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, dt)
    return np.concatenate(nodes), np.concatenate(bins), np.concatenate(ws)


def density_maps(pos, size, ratio, num_mov, bx, by, nbx, nby, exact=False):
    """-> (map0, map1) [nbx * nby]: uint64 fixed point (kernel semantics) or float64 sums (exact=True)."""
    dtype = np.float64 if exact else np.float32
    node, b, w = node_bin_weights(pos, size, ratio, bx, by, nbx, nby, dtype)
    nb = nbx * nby
    which = (node >= num_mov).astype(np.int64)
    if exact:
        m = np.bincount(which * nb + b, weights=w.astype(np.float64), minlength=2 * nb)
        return m[:nb], m[nb:]
    q = (w.astype(np.float64) * SCALE).astype(np.uint64)         # truncation toward zero, as the (u64) cast
    m = np.zeros(2 * nb, dtype=np.uint64)
    np.add.at(m, which * nb + b, q)
    return m[:nb], m[nb:]


def density_post(map0, map1, fixed, target, nbx, nby, exact=False):
    """-> (total density [nbx, nby], over [nbx, nby]) (gp_cuda_kernel.cu:130-139)."""
    fixed = np.asarray(fixed).reshape(-1)
    if exact:
        mov = map0 + fixed.astype(np.float64)
        r = mov + map1
        over = np.maximum(mov - target, 0.0)
    else:
        mov = (map0.astype(np.float64) / SCALE).astype(np.float32) + fixed.astype(np.float32)
        r = mov + (map1.astype(np.float64) / SCALE).astype(np.float32)
        over = np.maximum(mov - np.float32(target), np.float32(0))
    return r.reshape(nbx, nby), over.reshape(nbx, nby)


def mirror(rho):
    """[nbx, nby] -> [2 nbx, 2 nby] mirrored in both axes (the reference's FFT input)."""
    r = np.concatenate([rho, rho[::-1]], 0)
    return np.concatenate([r, r[:, ::-1]], 1)


def overflow(over, bin_area, mov_area):
    """sum(over) x bin_area / movable area (electronic_density_layer.py:110), summed in float64."""
    return float(np.asarray(over, dtype=np.float64).sum() * (bin_area / mov_area))


def density_grad(pos, size, ratio, field_x, field_y, wl_grad, pins, area, weight, precond, bx, by, nbx, nby,
                 exact=False):
    """Preconditioned gradient of every node (gp_cuda_kernel.cu:102-125). field_x / field_y: [nbx, nby]."""
    dtype = np.float64 if exact else np.float32
    dt = np.dtype(dtype).type
    n = len(pos)
    node, b, w = node_bin_weights(pos, size, ratio, bx, by, nbx, nby, dtype)
    fx = np.asarray(field_x, dtype=dtype).reshape(-1)
    fy = np.asarray(field_y, dtype=dtype).reshape(-1)
    gx = -np.bincount(node, weights=(w * fx[b]).astype(np.float64), minlength=n)
    gy = -np.bincount(node, weights=(w * fy[b]).astype(np.float64), minlength=n)
    g = np.stack([gx, gy], 1).astype(dtype) * dt(weight)
    nw = len(wl_grad)
    g[:nw] += np.asarray(wl_grad, dtype=dtype)
    if precond >= 0:
        scale = dt(1) / np.maximum(np.asarray(pins, dtype=dtype) + dt(precond) * np.asarray(area, dtype=dtype), dt(1))
        g = g * scale[:, None]
    return g.astype(dtype)


def brute_overlap_map(pos, size, ratio, bx, by, nbx, nby):
    """Per-bin charge by direct rectangle intersection in float64 (tiny grids only)."""
    out = np.zeros((nbx, nby))
    for (cx, cy), (w, h), r in zip(np.asarray(pos, np.float64), np.asarray(size, np.float64), np.asarray(ratio, np.float64)):
        x0, x1, y0, y1 = cx - w / 2, cx + w / 2, cy - h / 2, cy + h / 2
        for i in range(nbx):
            ox = max(0.0, min(x1, (i + 1) * bx) - max(x0, i * bx))
            if ox <= 0:
                continue
            for j in range(nby):
                oy = max(0.0, min(y1, (j + 1) * by) - max(y0, j * by))
                out[i, j] += r * ox * oy / (bx * by)
    return out
