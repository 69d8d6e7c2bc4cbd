"""Nesterov's method with a Barzilai-Borwein step (src/nesterov_optimizer.py:14-61 of the reference).

u is the main solution, v the look-ahead point where the gradient is taken. The step length alpha = |v' - v| /
|g' - g| measured on a step is the length of the next one; every step is taken.
"""
import math

import torch

from . import gp_cuda


def _dist(a, b):
    """L2 norm of a - b accumulated in float64."""
    return torch.linalg.vector_norm((a.double() - b.double()))


class NesterovOptimizer(object):
    def __init__(self, pos, grad_fn, lr, bounds, stats=None, start=None):
        """pos: updated in place (holds v). grad_fn(x) -> gradient. bounds = (lo, hi). stats: callable -> 0-d tensors
        of the last evaluated point, fetched with the step estimate in one transfer (left in `self.fetched`).
        start: the (point, gradient) estimate_initial_learning_rate(..., keep=True) left, used instead of a second
        evaluation of grad_fn when that point is pos itself (the reference evaluates it again, nesterov_optimizer.py:
        22-23; the objective is the same, so is the gradient)."""
        self.grad_fn = grad_fn
        self.stats, self.fetched = stats or (lambda: ()), []
        self.lo, self.hi = bounds
        self.kernel = True
        self._scalars = pos.new_empty(2)
        self.v = pos.clamp_(min=self.lo, max=self.hi)
        if start is not None and start[0] is pos:
            self.g = start[1]
        else:
            self.g = grad_fn(self.v)
        self.u = self.v.clone()
        probe = torch.add(self.v, self.g, alpha=-lr).clamp_(min=self.lo, max=self.hi)
        self.alpha = (_dist(self.v, probe) / _dist(self.g, grad_fn(probe))).item()
        self.a = 1.0
        self.moved = 0.0
        self.diverged = False

    def step(self):
        a_next = (1 + math.sqrt(4 * self.a * self.a + 1)) / 2
        coef = (self.a - 1) / a_next
        u_next, v_next = gp_cuda.nesterov_step(self.u, self.v, self.g, self.lo, self.hi,
                                               self._scalars[0:1].fill_(self.alpha), self._scalars[1:2].fill_(coef))
        g_next = self.grad_fn(v_next)
        vals = [_dist(v_next, self.v), _dist(g_next, self.g)] + [s.double() for s in self.stats()]
        moved, turned, *self.fetched = torch.stack(vals).tolist()
        if not all(map(math.isfinite, [moved, turned] + self.fetched[:1])):
            self.diverged = True
            return
        self.moved = moved
        self.alpha = moved / turned if turned > 0 else self.alpha
        self.u = u_next
        self.v.copy_(v_next)
        self.g = g_next
        self.a = a_next


def estimate_initial_learning_rate(pos, grad_fn, lr, bounds, keep=False):
    """Step length from one probe step of length lr (nesterov_optimizer.py:56-61). keep: -> (step length,
    (pos, its gradient)), the start NesterovOptimizer can take over."""
    pos.clamp_(min=bounds[0], max=bounds[1])
    g = grad_fn(pos)
    probe = (pos - lr * g).clamp_(min=bounds[0], max=bounds[1])
    step = (_dist(pos, probe) / _dist(g, grad_fn(probe))).item()
    return (step, (pos, g)) if keep else step
