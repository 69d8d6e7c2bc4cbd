"""Oracle: weighted-average (WA) wirelength, its gradient and the HPWL (test infrastructure only).

Restates src/core/wa_wirelength_hpwl.py:17-46 (optimized-net filter and the net / cell CSR) and the arithmetic of
cpp_to_py/gp_cuda_kernel.cu:12-57 (per net and axis: max / min shift, e^{(x-hi)/gamma}, e^{(lo-x)/gamma}, the
per-pin gradient, the per-cell sum) in numpy, in float32 (the kernel's precision) or float64 (the error reference).
"""
import numpy as np


def build_csr(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree):
    """Optimized nets (2 <= degree <= ignore_net_degree, renumbered) as the kernel's CSR view
    (wa_wirelength_hpwl.py:26-32,38-44). Pins must be grouped by net (ascending net id).

    -> dict(cell int32 [P], off float32 [P, 2], net_start int32 [nets + 1], cell_start int32 [num_mov + 1],
            cell_pin int32 [P'], num_nets, num_mov)"""
    pin_cell = np.asarray(pin_cell, dtype=np.int64)
    pin_net = np.asarray(pin_net, dtype=np.int64)
    degree = np.bincount(pin_net, minlength=num_nets)
    use = (degree >= 2) & (degree <= ignore_net_degree)
    new_id = np.cumsum(use.astype(np.int64)) - 1
    sel = use[pin_net]
    cell, off, net = pin_cell[sel], np.asarray(pin_off, dtype=np.float32)[sel], new_id[pin_net[sel]]
    n_opt = int(use.sum())
    net_start = np.searchsorted(net, np.arange(n_opt + 1)).astype(np.int32)
    order = np.argsort(cell, kind="stable")
    cell_start = np.searchsorted(cell[order], np.arange(num_mov + 1)).astype(np.int32)
    return {"cell": cell.astype(np.int32), "off": np.ascontiguousarray(off), "net_start": net_start,
            "cell_start": cell_start, "cell_pin": order.astype(np.int32)[:int(cell_start[-1])],
            "num_nets": n_opt, "num_mov": num_mov}


def pin_coords(pos, csr, dtype):
    """x of every pin: pos[cell] + off for a movable pin, off for a fixed one (cell == num_mov)."""
    n = csr["num_mov"]
    ext = np.concatenate([np.asarray(pos[:n], dtype=dtype), np.zeros((1, 2), dtype=dtype)])
    return ext[csr["cell"]] + csr["off"].astype(dtype)


def wa_forward(pos, csr, gamma, dtype=np.float32):
    """-> (WA total, grad [num_mov, 2], HPWL total, out [nets, 2] = per-net (WA, HPWL), pin_grad [P, 2]).

    Totals are summed in float64 whatever `dtype` (the device reduces in FP64)."""
    dtype = np.dtype(dtype)
    ns = csr["net_start"].astype(np.int64)
    x = pin_coords(pos, csr, dtype)
    deg = np.diff(ns)
    net = np.repeat(np.arange(len(deg)), deg)
    inv_g = dtype.type(1.0) / dtype.type(gamma)
    hi = np.maximum.reduceat(x, ns[:-1], axis=0)
    lo = np.minimum.reduceat(x, ns[:-1], axis=0)
    ep = np.exp((x - hi[net]) * inv_g)
    en = np.exp((lo[net] - x) * inv_g)
    sp = np.add.reduceat(ep, ns[:-1], axis=0)
    sn = np.add.reduceat(en, ns[:-1], axis=0)
    xp = np.add.reduceat(x * ep, ns[:-1], axis=0) / sp
    xn = np.add.reduceat(x * en, ns[:-1], axis=0) / sn
    pin_grad = (ep * (1 + (x - xp[net]) * inv_g) / sp[net] - en * (1 - (x - xn[net]) * inv_g) / sn[net]).astype(dtype)
    out = np.stack([(xp - xn).sum(1), (hi - lo).sum(1)], 1).astype(dtype)
    grad = cell_sum(pin_grad, csr, dtype)
    return float(out[:, 0].astype(np.float64).sum()), grad, float(out[:, 1].astype(np.float64).sum()), out, pin_grad


def cell_sum(pin_grad, csr, dtype=np.float32):
    """grad[c] = sum of the pin gradients of cell c in cell_pin order (gp_cuda_kernel.cu:46-57)."""
    n = csr["num_mov"]
    cs = csr["cell_start"].astype(np.int64)
    g = pin_grad[csr["cell_pin"]].astype(np.float64)
    out = np.zeros((n, 2))
    has = np.diff(cs) > 0
    if len(g):
        sums = np.add.reduceat(g, cs[:-1][has], axis=0)
        out[has] = sums
    return out.astype(dtype)


def hpwl_exact(pos, pin_cell, pin_off, pin_net, num_nets, num_mov):
    """Exact HPWL over every net in float64 (wa_wirelength_hpwl.py:58-70, evaluator.py:4-8)."""
    pin_cell = np.asarray(pin_cell, dtype=np.int64)
    ext = np.concatenate([np.asarray(pos[:num_mov], dtype=np.float64), np.zeros((1, 2))])
    x = ext[pin_cell] + np.asarray(pin_off, dtype=np.float64)
    hi = np.full((num_nets, 2), -np.inf)
    lo = np.full((num_nets, 2), np.inf)
    np.maximum.at(hi, np.asarray(pin_net), x)
    np.minimum.at(lo, np.asarray(pin_net), x)
    ok = np.isfinite(hi[:, 0])
    return float((hi[ok] - lo[ok]).sum())


def wa_brute(xs, gamma):
    """WA of one net along one axis straight from the definition, float64 (no max / min shift). -> (wa, d wa / d x)."""
    xs = np.asarray(xs, dtype=np.float64)
    a = np.exp(xs / gamma)
    b = np.exp(-xs / gamma)
    sp, sn = a.sum(), b.sum()
    xp, xn = (xs * a).sum() / sp, (xs * b).sum() / sn
    gp = a / sp * (1 + (xs - xp) / gamma)
    gn = b / sn * (1 - (xs - xn) / gamma)
    return xp - xn, gp - gn
