"""The parameters that change during global placement, and when to stop: the host side of the run's control.

Follows src/param_scheduler.py:24-116 of the reference. The device control kernel (csrc/gp_control.hip) makes the
same decisions in the same order in double precision; this class runs the wirelength-only start, carries the
state into and out of the device loop, and hands a level over (`resume`).

    density_weight  0 during the wirelength-only start; then `density_weight` x the tension unit, multiplied each
                    iteration by resolution.growth
    wa_coeff        the WA smoothing gamma, from 10 x base_gamma; resolution.smoothing (--wa_rule accuracy) or
                    Xplace's overflow formula (--wa_rule overflow)
    handover_gamma  clusters only: the level ends ("handover") once gamma is down to it
"""
import math

from . import resolution


class ParamScheduler(object):
    def __init__(self, data, args):
        """data: bin (bin size x, y), refine, total_pins, total_area, pitch, num_nodes."""
        self.args = args
        self.iter = 0
        self.wa, self.hpwl, self.overflow = [], [], []
        self.base_gamma = args.wa_coeff * (data.bin[0] + data.bin[1])
        self.refine = data.refine
        self.wa_coeff = 10 * self.base_gamma
        self.density_weight = 0.0
        self.handover_gamma = None
        self.total_pins, self.total_area = data.total_pins, data.total_area
        self.pitch = data.pitch
        self.motion_unit = data.pitch * math.sqrt(data.num_nodes)
        self.motion = 0.0
        self.weighted_weight = 0.0
        self.mu = 1.0
        self.settled, self.best_hpwl, self.best_overflow = 0, float("inf"), float("inf")
        self.is_best, self.improved_iter = False, 0

    def start_step(self, hpwl, wa, window=5):
        """One wirelength-only iteration -> True once WA fell by less than start_tol per iteration over `window`."""
        self.hpwl.append(hpwl)
        self.wa.append(wa)
        self.iter += 1
        return not (len(self.wa) <= window or self.wa[-1 - window] - wa > self.args.start_tol * window * wa)

    def gamma(self, overflow):
        return self.base_gamma * 10 ** ((overflow - 0.1) * 20 / 9 - 1)

    def resume(self, coarse, density_weight, hpwl, overflow=None):
        """Continue where the cluster level (`coarse`) stopped, with no wirelength-only start."""
        self.density_weight, self.iter = density_weight, coarse.iter
        if self.args.wa_rule == "accuracy":
            self.wa_coeff = coarse.wa_coeff / coarse.base_gamma * self.base_gamma
        else:
            self.wa_coeff = self.gamma(overflow)
        self.hpwl.append(hpwl)

    def step(self, hpwl, overflow, moved, wa):
        self.motion = moved / self.motion_unit
        part = self.density_weight * self.total_area
        self.weighted_weight = part / (self.total_pins + part)
        self.hpwl.append(hpwl)
        self.overflow.append(overflow)
        self.mu = resolution.growth(self.motion, self.pitch, self.wa_coeff, self.hpwl[-1] < self.hpwl[-2])
        self.density_weight *= self.mu
        if self.args.wa_rule != "accuracy":
            self.wa_coeff = self.gamma(overflow)
        else:
            self.wa_coeff = resolution.smoothing(self.wa_coeff, hpwl, wa, overflow, self.refine, self.base_gamma,
                                                 self.mu)
        self.iter += 1

    def need_to_stop(self):
        """-> reason or None; sets is_best when the current placement is to be kept."""
        hpwl, overflow = self.hpwl[-1], self.overflow[-1]
        self.is_best = False
        if not (0 <= hpwl < 1e30 and 0 <= overflow < 1e30):
            return "diverged"
        if self.iter >= self.args.min_iter and overflow < self.args.stop_overflow:
            self.settled += 1
            self.is_best = hpwl < self.best_hpwl
            self.best_hpwl = min(hpwl, self.best_hpwl)
            if self.settled > self.args.settle_iter:
                return "converged"
        elif self.settled == 0:
            self.is_best = overflow < self.best_overflow
            if overflow < self.best_overflow * 0.999:
                self.improved_iter = self.iter
            self.best_overflow = min(overflow, self.best_overflow)
        if (self.is_best and self.settled) or self.weighted_weight < 0.95:
            self.improved_iter = self.iter
        if self.iter - self.improved_iter > self.args.stall_iter:
            return "stalled"
        if self.handover_gamma and self.density_weight > 0 and resolution.handover(self.wa_coeff, self.handover_gamma):
            return "handover"
        return None
