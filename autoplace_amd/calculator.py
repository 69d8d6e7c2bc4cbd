"""The gradient of the objective: WA wirelength + density_weight x density energy, divided per node by
max(1, #pins + density_weight x area) (src/calculator.py:12-43 of the reference). The objective's value is never
needed by the optimizer; the HPWL and overflow of the evaluated point come from the same pass.
While the density weight is 0 (the wirelength-only start) no density map or field is computed."""
import torch


class Objective(object):
    def __init__(self, wirelength, density, ps):
        self.wirelength, self.density, self.ps = wirelength, density, ps
        self.wa = self.hpwl = self.overflow = None
        self.evals = 0
        self.not_measured = torch.tensor(float("nan"), dtype=torch.float64, device=density.device)
        self.pin_precond = density.pins[:density.num_mov].clamp(min=1.0).unsqueeze(1)

    def tension_unit(self, pos):
        """#pins / L1 norm of the cells' plain density gradient: the unit of the first density weight."""
        n = self.density.num_mov
        g = self.density.forward(pos)[0][:n]
        return (self.density.pins.double().sum() / g.double().abs().sum()).item()

    def __call__(self, pos):
        self.wa, wl_grad, self.hpwl = self.wirelength.forward(pos, self.ps.wa_coeff)
        if self.ps.density_weight == 0:
            grad, self.overflow = torch.zeros_like(pos), self.not_measured
            grad[:len(wl_grad)] = wl_grad / self.pin_precond
        else:
            grad, self.overflow = self.density.forward(pos, wl_grad, self.ps.density_weight, self.ps.density_weight)
        self.evals += 1
        return grad

    def on_gpu(self, pos, kp):
        """The same gradient with the parameters in device memory: kp = (gamma, weight, precond, ...)."""
        self.wa, wl_grad, self.hpwl = self.wirelength.forward(pos, kp[0:1])
        grad, self.overflow = self.density.forward(pos, wl_grad, kp[1:2], kp[2:3])
        self.evals += 1
        return grad
