"""Electrostatic density (ePlace) on the device (src/core/electronic_density_layer.py:34-149 of the reference, on the
HIP kernels): cells are charges, the density map is the charge density, the force on a cell is the field.

Smoothing: a side narrower than sqrt(2) bins is widened to sqrt(2) bins and the density scaled to keep the charge.
Field: Poisson's equation with zero normal field at the region boundary, solved by the library's own cosine-series
solver (GEMM or FFT form) instead of torch.fft on the mirrored map. Two grids: the overflow may be measured on a
coarser grid (`judge`) than the force.
"""
import math

import numpy as np
import torch

from . import gp_cuda


class ElectronicDensityLayer(object):
    def __init__(self, width, height, nbx, nby, fixed_fraction, target_density, node_size, num_mov, device,
                 node_pins=None, field=True, poisson="auto"):
        """node_size [nodes, 2] (real sizes), fixed_fraction [nbx, nby], node_pins [nodes] (for the preconditioner):
        numpy arrays, or float64 tensors on the device (a layer built inside the GP clock: no host round trip; the
        float32 arrays are the same, elementwise IEEE operations either way). field=False: a grid that only measures
        overflows."""
        device = torch.device(device)
        if device.type != "cuda":
            raise RuntimeError("the density kernels run on a HIP device (got %s)" % device)
        self.nbx, self.nby, self.num_mov, self.num_nodes = nbx, nby, num_mov, len(node_size)
        self.target, self.device = target_density, device
        self.width, self.height = width, height
        self.bin_np = np.array([width / nbx, height / nby])
        self.bin_area = float(self.bin_np.prod())
        t = lambda a, dt=torch.float32: torch.as_tensor(a, dtype=dt, device=device)   # noqa: E731
        self.bin = t(self.bin_np)
        on_device = torch.is_tensor(node_size)
        real = node_size.to(device, torch.float64) if on_device else np.asarray(node_size, dtype=np.float64)
        frac = fixed_fraction.to(device, torch.float64) if torch.is_tensor(fixed_fraction) else \
            np.asarray(fixed_fraction, dtype=np.float64)
        self.mov_area = float(real[:num_mov].prod(1).sum())
        self.fixed_map = t(frac * target_density)
        self._fraction = frac
        self.fixed_mean = float(frac.mean())           # the blocked share of the region (cluster.py:126)
        self.real_size = t(real, torch.float64)        # float64 real sizes, kept for a layer built from this one
        floor = math.sqrt(2) * self.bin_np
        smooth = torch.maximum(real, t(floor, torch.float64)) if on_device else np.maximum(real, floor)
        self.smooth = {"size": t(smooth), "ratio": t(real.prod(1) / smooth.prod(1))}
        self.real = {"size": t(real[:num_mov]), "ratio": torch.ones(num_mov, device=device)}
        self.pos_lo = self.smooth["size"] / 2
        self.pos_hi = t([width, height]) - self.pos_lo
        if node_pins is None:
            self.pins = torch.zeros(len(real), device=device)
        else:
            self.pins = t(node_pins)
        self.area = t(real.prod(1))
        self.no_grad = torch.zeros(0, 2, device=device)
        self._scalars = torch.empty(2, device=device)
        self.judge = None
        self.plan = gp_cuda.PoissonPlan(nbx, nby, float(self.bin_np[0]), float(self.bin_np[1]), poisson) \
            if field else None

    def fraction_on_device(self):
        """The grid's blocked fraction as a float64 tensor [nbx, nby] on the device."""
        if not torch.is_tensor(self._fraction):
            self._fraction = torch.as_tensor(self._fraction, dtype=torch.float64, device=self.device)
        return self._fraction

    def _density(self, pos, n, num_mov):
        """-> (total density mirrored [2 nbx, 2 nby], overflow as a 0-d float64 tensor)."""
        rho, over = gp_cuda.density_map(pos, n["size"], n["ratio"], num_mov, float(self.bin_np[0]),
                                        float(self.bin_np[1]), self.nbx, self.nby, self.fixed_map, self.target)
        return rho, over.double().sum() * (self.bin_area / self.mov_area)

    def exact_overflow(self, pos):
        """Overflow of the movable cells with their real sizes (no smoothing, no fillers) on the overflow grid."""
        if self.judge is not None:
            return self.judge.exact_overflow(pos)
        return self._density(pos[:self.num_mov].contiguous(), self.real, self.num_mov)[1]

    def smoothed_overflow(self, pos):
        """Overflow of the smoothed movable cells (a layer built without fillers)."""
        return self._density(pos[:self.num_mov].contiguous(), self.smooth, self.num_mov)[1]

    def field(self, mirrored):
        return gp_cuda.poisson_field(mirrored, self.plan)

    def forward(self, pos, wl_grad=None, weight=1.0, precond=-1.0):
        """pos [nodes, 2] -> (gradient [nodes, 2], overflow). gradient = (wl_grad + weight x density gradient) /
        max(1, pins + precond x area); precond < 0: no division. weight / precond: floats or 1-element tensors."""
        n = self.smooth
        mirrored, overflow = self._density(pos, n, self.num_mov)
        if self.judge is not None:
            overflow = self.judge.smoothed_overflow(pos)
        field = self.field(mirrored)
        wl_grad = self.no_grad if wl_grad is None else wl_grad
        if not torch.is_tensor(weight):
            weight = self._scalars[0:1].fill_(weight)
        if not torch.is_tensor(precond):
            precond = self._scalars[1:2].fill_(precond)
        grad = gp_cuda.density_grad(pos, n["size"], n["ratio"], field, wl_grad, self.pins, self.area, weight, precond,
                                    float(self.bin_np[0]), float(self.bin_np[1]), self.nbx, self.nby)
        return grad, overflow
