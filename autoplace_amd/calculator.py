"""Objective of the host-driven part of a level (the wirelength-only start and the first density-weighted step).

The optimizer needs only the gradient. With w the density weight it is, per node,

    (dWA/dpos + w dE/dpos) / max(1, pins + w area)        (w > 0: the density operator composes all of it)
    (dWA/dpos) / max(1, pins)                             (w = 0: no density map, no field; fillers get 0)

and the pass leaves the WA value, the HPWL of the optimized nets and the overflow of the evaluated point in
`wa`, `hpwl`, `overflow` (overflow is NaN while w = 0: it is not measured). Behaviour of src/calculator.py:12-35;
the device loop evaluates the same gradient in its own gather (gpu_loop.Engine).
"""
import torch


class Objective(object):
    """Gradient evaluations of one level: `wirelength` (wirelength.Wirelength), `density`
    (density.ElectronicDensityLayer), `ps` (the level's schedule: wa_coeff, density_weight). `evals` counts them."""

    def __init__(self, wirelength, density, ps):
        self.wirelength = wirelength
        self.density = density
        self.ps = ps
        self.evals = 0
        self.wa = None
        self.hpwl = None
        self.overflow = None
        cells = density.num_mov
        # the w = 0 preconditioner of the cells, [cells, 1] for the broadcast over (x, y)
        self._cell_divisor = density.pins[:cells].clamp(min=1.0).unsqueeze(1)
        self._nan = torch.full((), float("nan"), dtype=torch.float64, device=density.device)

    def __call__(self, pos):
        """pos [nodes, 2] inside the region (density.pos_lo .. pos_hi) -> gradient [nodes, 2]."""
        gamma, weight = self.ps.wa_coeff, self.ps.density_weight
        self.wa, wl_grad, self.hpwl = self.wirelength.forward(pos, gamma)
        if weight == 0:
            grad = self._wirelength_only(pos, wl_grad)
            self.overflow = self._nan
        else:
            grad, self.overflow = self.density.forward(pos, wl_grad, weight, weight)
        self.evals += 1
        return grad

    def _wirelength_only(self, pos, wl_grad):
        grad = pos.new_zeros(pos.shape)
        grad[:wl_grad.size(0)] = wl_grad / self._cell_divisor
        return grad

    def tension_unit(self, pos):
        """The density weight, in which the run's first weight is given, at which the density gradient of the cells
        pulls as hard as all their pins together: sum of pins / L1 norm of the cells' unweighted, unpreconditioned
        density gradient (sums in FP64)."""
        cells = self.density.num_mov
        dens = self.density.forward(pos)[0][:cells]
        return (self.density.pins.double().sum() / dens.double().abs().sum()).item()
