"""The two energies of global placement as torch.autograd.Function objects whose backward passes are the library's
hand-written gradient kernels.

    WAEnergy.apply(pos, wirelength, gamma)  -> WA wirelength of the optimized nets (0-d float32)
        backward: the WA kernel's per-cell gradient (the same pass computes value and gradient; fillers get 0)
    DensityEnergy.apply(pos, density)       -> electrostatic energy 1/2 sum rho phi of the smoothed map (0-d float32)
        backward: the field gather (density_grad with weight 1, no preconditioner), i.e. ePlace's field
        approximation -(sum_b w_ib E_b) of the energy's gradient, which the reference uses as the density gradient

The reference (src/calculator.py:4-5) never evaluates either energy; these wrappers let the objective be composed
and differentiated with ordinary torch code (e.g. W(pos) + lam * D(pos)).
"""
import torch

from . import gp_cuda


class WAEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, wirelength, gamma):
        wa, grad, _ = wirelength.forward(pos.detach().contiguous(), float(gamma))
        ctx.save_for_backward(grad)
        ctx.n = pos.shape[0]
        return wa.float()

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        full = grad.new_zeros((ctx.n, 2))
        full[:grad.shape[0]] = grad * g
        return full, None, None


class DensityEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, density):
        p = pos.detach().contiguous()
        rho, _ = density._density(p, density.smooth, density.num_mov)          # mirrored [2 nbx, 2 nby]
        plan = getattr(density, "_energy_plan", None)
        if plan is None:
            plan = density._energy_plan = gp_cuda.PoissonPlan(density.nbx, density.nby, float(density.bin_np[0]),
                                                              float(density.bin_np[1]), "gemm")
        e = plan.energy(rho, 2 * density.nby)                                   # the top-left quadrant
        ctx.save_for_backward(p)
        ctx.density = density
        return e.float()

    @staticmethod
    def backward(ctx, g):
        (p,) = ctx.saved_tensors
        grad, _ = ctx.density.forward(p)                                        # weight 1, no preconditioner
        return grad * g, None
