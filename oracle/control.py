"""Oracle: one pass of the run's control in double precision (test infrastructure only).

Restates cpp_to_py/gp_cuda_kernel.cu:172-249 (control_kernel, select_kernel), which is the device form of
src/nesterov_optimizer.py:34-53 (the Barzilai-Borwein step), src/param_scheduler.py:63-116 (step, need_to_stop)
and src/resolution.py:33-56, in the kernel's order. State layout: src/gpu_loop.py:20-23.
"""
import math

import numpy as np

FIELDS = ("lam gam alpha a iter hpwl_prev motion ww mu settled best_hpwl best_ovf is_best improved stop accept "
          "loop_it evals res_hpwl res_ovf has_best").split()
K = {name: i for i, name in enumerate(FIELDS)}
CONSTS = ("base_gamma motion_unit total_pins total_area pitch min_iter stop_overflow settle_iter stall_iter "
          "inner_iter refine accuracy handover_gamma").split()
C = {name: i for i, name in enumerate(CONSTS)}
REASON = {1: "converged", 2: "stalled", 3: "diverged", 4: "max_iter", 5: "handover"}


def next_coef(a):
    a_next = (1 + math.sqrt(4 * a * a + 1)) / 2
    return (a - 1) / a_next


def control_pass(S, kp, c, moved, turned, hpwl, ovf, wa, lam_factor=1.0, gam_factor=1.0):
    """Mutates S (list of 21 floats) and kp (float32 [5]) as control_kernel does. lam_factor / gam_factor: the learned
    controller's factors on the density-weight growth and on the new smoothing (1: the plain rule)."""
    S[K["accept"]] = 0.0
    S[K["is_best"]] = 0.0
    if S[K["stop"]] != 0:
        return
    moved, turned, hpwl, ovf, wa = float(moved), float(turned), float(hpwl), float(ovf), float(wa)
    if not all(math.isfinite(v) for v in (moved, turned, hpwl, ovf)):
        S[K["stop"]] = 3.0
        return
    alpha_next = moved / turned if turned > 0 else S[K["alpha"]]
    S[K["alpha"]] = alpha_next
    S[K["evals"]] += 1
    S[K["accept"]] = 1.0
    S[K["a"]] = (1 + math.sqrt(4 * S[K["a"]] * S[K["a"]] + 1)) / 2
    it = S[K["iter"]]
    S[K["motion"]] = moved / c[C["motion_unit"]]
    part = S[K["lam"]] * c[C["total_area"]]
    S[K["ww"]] = part / (c[C["total_pins"]] + part)
    if hpwl < S[K["hpwl_prev"]]:
        mu = 1.05 * 0.98
    else:
        mu = 1.05 * max(math.pow(1.05, -S[K["motion"]] * c[C["pitch"]] / S[K["gam"]]), 0.95)
    mu *= lam_factor
    S[K["hpwl_prev"]] = hpwl
    S[K["mu"]] = mu
    S[K["lam"]] *= mu
    if c[C["accuracy"]] != 0:
        err = max(1e-6, 1 - wa / max(hpwl, 1e-12))
        m = max(mu, 1 / mu)
        g = S[K["gam"]] * min(m, max(1 / m, math.sqrt(max(ovf, 1e-3) / c[C["refine"]] / err))) * gam_factor
        S[K["gam"]] = min(100 * c[C["base_gamma"]], max(0.01 * c[C["base_gamma"]], g))
    else:
        S[K["gam"]] = c[C["base_gamma"]] * math.pow(10.0, (ovf - 0.1) * 20 / 9 - 1) * gam_factor
    it = it + 1
    S[K["iter"]] = it
    stop, is_best = 0.0, 0.0
    if not (hpwl < 1e30 and ovf < 1e30):
        stop = 3.0
    else:
        if it >= c[C["min_iter"]] and ovf < c[C["stop_overflow"]]:
            S[K["settled"]] += 1
            is_best = 1.0 if hpwl < S[K["best_hpwl"]] else 0.0
            S[K["best_hpwl"]] = min(hpwl, S[K["best_hpwl"]])
            if S[K["settled"]] > c[C["settle_iter"]]:
                stop = 1.0
        elif S[K["settled"]] == 0:
            is_best = 1.0 if ovf < S[K["best_ovf"]] else 0.0
            if ovf < S[K["best_ovf"]] * 0.999:
                S[K["improved"]] = it
            S[K["best_ovf"]] = min(ovf, S[K["best_ovf"]])
        if stop == 0:
            if (is_best and S[K["settled"]] > 0) or S[K["ww"]] < 0.95:
                S[K["improved"]] = it
            if it - S[K["improved"]] > c[C["stall_iter"]]:
                stop = 2.0
        if stop == 0 and c[C["handover_gamma"]] > 0 and S[K["lam"]] > 0 and S[K["gam"]] <= c[C["handover_gamma"]]:
            stop = 5.0
    S[K["is_best"]] = is_best
    if is_best:
        S[K["res_hpwl"]], S[K["res_ovf"]], S[K["has_best"]] = hpwl, ovf, 1.0
    S[K["loop_it"]] += 1
    if stop == 0 and S[K["loop_it"]] >= c[C["inner_iter"]]:
        stop = 4.0
    S[K["stop"]] = stop
    if not S[K["has_best"]]:
        S[K["res_hpwl"]], S[K["res_ovf"]] = hpwl, ovf
    a_next = (1 + math.sqrt(4 * S[K["a"]] * S[K["a"]] + 1)) / 2
    kp[0] = np.float32(S[K["gam"]])
    kp[1] = np.float32(S[K["lam"]])
    kp[2] = np.float32(S[K["lam"]])
    kp[3] = np.float32(alpha_next)
    kp[4] = np.float32((S[K["a"]] - 1) / a_next)


def nesterov_step(u, v, g, lo, hi, alpha, coef):
    """u' = v - alpha g; v' = clamp(u' + coef (u' - u), lo, hi) in float32 (gp_cuda_kernel.cu:157-165)."""
    alpha, coef = np.float32(alpha), np.float32(coef)
    un = (v - alpha * g).astype(np.float32)
    vn = np.minimum(np.maximum(un + coef * (un - u), lo), hi).astype(np.float32)
    return un, vn


def dist(a, b):
    """L2 norm of a - b accumulated in float64."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(math.sqrt((d * d).sum()))
