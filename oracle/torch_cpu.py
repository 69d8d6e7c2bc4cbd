"""Oracle: the reference's CPU path of one GP pass in PyTorch on the host cores (test and baseline infrastructure
only: bench.py times it as the CPU baseline; nothing in the product imports it).

One pass is what src/gpu_loop.py:58-64 does per iteration, on the CPU branches the reference takes when its CUDA
module is absent (--gpu -1):
    Nesterov step            src/nesterov_optimizer.py:41-42 (torch.add, torch.lerp, clamp)
    WA value / gradient      src/core/wa_wirelength_hpwl.py:83-94 (scatter_reduce amax / amin, exp, index_add)
    density map, overflow    src/core/electronic_density_layer.py:80-94,96-110 (nodes cut into pieces of at most
                             2 x 2 bins, each piece spread over a 3 x 3 bin window, one index_add; the map mirrored)
    overflow grid            :112-120,127-128 (smoothed cells on the coarser grid)
    Poisson field            :62-77,133 (rfft2 of the mirrored map times the spectral multiplier, irfft2)
    field gather             :142-148 (charge x field over the same windows, index_add, preconditioner)
    BB norms                 src/gpu_loop.py:62 (torch.dist)
and the control decisions on the host in FP64 (oracle.control.control_pass), as the reference's host loop makes
them. Everything is float32 as in the reference.
"""
import math

import numpy as np
import torch


def _pieces(size, bin_xy):
    """Cut every node into n_x x n_y equal pieces of at most 2 x 2 bins. -> (node of each piece, offset of its
    centre from the node's centre [Q, 2], its half size [Q, 2]) as numpy arrays."""
    size = np.asarray(size, np.float64)
    cuts = np.maximum(np.ceil(size / (2.0 * bin_xy) - 1e-9), 1).astype(np.int64)
    per = cuts[:, 0] * cuts[:, 1]
    node = np.repeat(np.arange(len(size)), per)
    first = np.repeat(np.cumsum(per) - per, per)
    k = np.arange(int(per.sum())) - first
    piece = size[node] / cuts[node]
    col = np.stack([k // cuts[node, 1], k % cuts[node, 1]], 1)
    return node, (col + 0.5) * piece - size[node] / 2, piece / 2


class Grid(object):
    """One density grid: the smoothed nodes cut into pieces, the fixed map and (for the force grid) the spectral
    multiplier of the field."""

    def __init__(self, width, height, nb, fixed_fraction, target, node_size, num_mov, node_pins=None, field=True):
        self.nb, self.num_mov, self.target = nb, num_mov, target
        self.bin_np = np.array([width / nb, height / nb])
        self.bin = torch.tensor(self.bin_np, dtype=torch.float32)
        self.bin_area = float(self.bin_np.prod())
        real = np.asarray(node_size, np.float64)
        self.mov_area = float(real[:num_mov].prod(1).sum())
        smooth = np.maximum(real, math.sqrt(2) * self.bin_np)
        ratio = real.prod(1) / smooth.prod(1)
        node, off, half = _pieces(smooth, self.bin_np)
        self.node = torch.tensor(node)
        self.off = torch.tensor(off, dtype=torch.float32)
        self.half = torch.tensor(half, dtype=torch.float32)
        self.weight = torch.tensor(ratio[node], dtype=torch.float32).view(-1, 1, 1)
        self.filler = (self.node >= num_mov).long().view(-1, 1, 1) * (nb * nb)
        self.fixed = torch.tensor(np.asarray(fixed_fraction) * target, dtype=torch.float32).reshape(-1)
        self.lo = torch.tensor(smooth / 2, dtype=torch.float32)
        self.hi = torch.tensor([width, height], dtype=torch.float32) - self.lo
        self.steps = torch.arange(3)
        idx = np.minimum(np.arange(2 * nb), 2 * nb - 1 - np.arange(2 * nb))
        self.mirror = torch.tensor((idx[:, None] * nb + idx[None, :]).reshape(-1))
        self.pins = torch.tensor(np.zeros(len(real)) if node_pins is None else node_pins, dtype=torch.float32)
        self.area = torch.tensor(real.prod(1), dtype=torch.float32)
        if field:
            bx, by = float(self.bin_np[0]), float(self.bin_np[1])
            kx = (2 * math.pi * torch.fft.fftfreq(2 * nb, d=bx)).unsqueeze(1)
            ky = (2 * math.pi * torch.fft.rfftfreq(2 * nb, d=by)).unsqueeze(0)
            inv = 1.0 / (kx * kx + ky * ky).clamp(min=1e-30)
            inv[0, 0] = 0.0
            inv = inv / bx
            kx[nb, 0] = 0.0
            ky[0, nb] = 0.0
            self.spectral = torch.stack([-1j * kx * inv, -1j * ky * inv]).to(torch.complex64)

    def windows(self, pos):
        """-> (flat bin of each of a piece's 3 x 3 bins [Q, 3, 3], the charge it adds there [Q, 3, 3])."""
        centre = pos[self.node] + self.off
        lo = (centre - self.half) / self.bin
        hi = (centre + self.half) / self.bin
        first = lo.floor().clamp(min=0).unsqueeze(2) + self.steps
        cover = (torch.minimum(hi.unsqueeze(2), first + 1) - torch.maximum(lo.unsqueeze(2), first)).clamp(min=0)
        bx = first[:, 0].long().clamp(max=self.nb - 1)
        by = first[:, 1].long().clamp(max=self.nb - 1)
        return bx.unsqueeze(2) * self.nb + by.unsqueeze(1), self.weight * cover[:, 0, :, None] * cover[:, 1, None, :]

    def maps(self, pos):
        """-> (mirrored total map [2 nb, 2 nb] or None, overflow) of the nodes at pos."""
        nb2 = self.nb * self.nb
        at, q = self.windows(pos)
        both = torch.zeros(2 * nb2).index_add_(0, (at + self.filler).reshape(-1), q.reshape(-1))
        mov = both[:nb2] + self.fixed
        overflow = (mov - self.target).clamp(min=0).sum() * (self.bin_area / self.mov_area)
        return (mov + both[nb2:])[self.mirror].view(2 * self.nb, 2 * self.nb), overflow

    def gradient(self, pos, mirrored, wl_grad, weight, precond):
        """Preconditioned gradient from the field of the mirrored map."""
        nb = self.nb
        field = torch.fft.irfft2(torch.fft.rfft2(mirrored) * self.spectral, s=(2 * nb, 2 * nb))
        at, q = self.windows(pos)
        per_bin = field[:, :nb, :nb].reshape(2, -1).t()
        force = (q.unsqueeze(3) * per_bin[at]).sum((1, 2))
        grad = torch.zeros_like(pos).index_add_(0, self.node, force).mul_(-weight)
        grad[:len(wl_grad)] += wl_grad
        return grad / (self.pins + precond * self.area).clamp(min=1.0).unsqueeze(1)


class Level(object):
    """One level's problem on the CPU: the optimized nets, the force grid and the overflow grid."""

    def __init__(self, prob, ignore_net_degree=100):
        n = prob.num_mov
        cell = np.asarray(prob.pin_cell, np.int64)
        net = np.asarray(prob.pin_net, np.int64)
        degree = np.bincount(net, minlength=prob.num_nets)
        keep = (degree >= 2) & (degree <= ignore_net_degree)
        renum = np.cumsum(keep) - 1
        sel = keep[net]
        self.num_mov, self.num_nets = n, int(keep.sum())
        self.cell = torch.tensor(cell[sel])
        self.net = torch.tensor(renum[net[sel]])
        self.net2 = self.net.unsqueeze(1).expand(-1, 2).contiguous()
        self.pin_off = torch.tensor(np.asarray(prob.pin_off, np.float32)[sel])
        self.force = Grid(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size,
                          n, prob.node_pins)
        self.judge = None
        if prob.num_bin != prob.force_bin:
            self.judge = Grid(prob.width, prob.height, prob.num_bin, prob.judge_fraction, prob.target,
                              prob.node_size[:n], n, field=False)

    def wirelength(self, pos, gamma):
        """-> (WA, gradient [num_mov, 2], HPWL of the optimized nets)."""
        ext = torch.cat([pos[:self.num_mov], pos.new_zeros(1, 2)])
        x = ext[self.cell] + self.pin_off
        shape = (self.num_nets, 2)
        hi = x.new_zeros(shape).scatter_reduce_(0, self.net2, x, "amax", include_self=False)
        lo = x.new_zeros(shape).scatter_reduce_(0, self.net2, x, "amin", include_self=False)
        ep = torch.exp((x - hi[self.net]) / gamma)
        en = torch.exp((lo[self.net] - x) / gamma)
        sp = x.new_zeros(shape).index_add_(0, self.net, ep)
        sn = x.new_zeros(shape).index_add_(0, self.net, en)
        xp = x.new_zeros(shape).index_add_(0, self.net, x * ep) / sp
        xn = x.new_zeros(shape).index_add_(0, self.net, x * en) / sn
        pin = ep * (1 + (x - xp[self.net]) / gamma) / sp[self.net] - en * (1 - (x - xn[self.net]) / gamma) / sn[self.net]
        grad = x.new_zeros(self.num_mov + 1, 2).index_add_(0, self.cell, pin)[:self.num_mov]
        return (xp - xn).sum(), grad, (hi - lo).sum()

    def objective(self, pos, kp):
        """Objective.on_gpu's quantities on the CPU: -> (gradient, WA, HPWL, overflow)."""
        wa, wl_grad, hpwl = self.wirelength(pos, float(kp[0]))
        mirrored, overflow = self.force.maps(pos)
        if self.judge is not None:
            overflow = self.judge.maps(pos[:self.num_mov])[1]
        grad = self.force.gradient(pos, mirrored, wl_grad, float(kp[1]), float(kp[2]))
        return grad, wa, hpwl, overflow


def nesterov_step(u, v, g, lo, hi, alpha, coef):
    """nesterov_optimizer.py:41-42 -> (u', v')."""
    u_next = torch.add(v, g, alpha=-alpha)
    return u_next, torch.lerp(u, u_next, 1 + coef).clamp_(min=lo, max=hi)
