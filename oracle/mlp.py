"""Oracle: the learned schedule controller in FP64 (test infrastructure only).

The controller has no counterpart in the reference; this restates the build's own definition (the library's
gp_mlp.h / gp_mlp.hip, include/autoplace_gp.h part 4) to check the HIP kernels:
    forward / loss_grad   the 12 -> 64 -> 64 -> 2 ReLU MLP and the gradient of sum_s w_s |o_s - t_s|^2, by hand
                          (pinned by central finite differences in tests/test_oracle_mlp.py)
    features              the 12 per-pass features (pinned only against this definition: parity unpinned)
    normal_draw           the counter-based exploration draws (splitmix64 + Box-Muller)
    controller_pass       features -> MLP -> noise -> clip -> oracle.control.control_pass with the factors, and the
                          fallback bookkeeping
"""
import math

import numpy as np

from . import control as octl

IN, H, OUT = 12, 64, 2
W1, B1 = 0, H * IN
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + OUT * H
PARAMS = B3 + OUT
CLIP = 0.2
X_STARTED, X_PREV_OVF, X_HPWL0, X_ALPHA0, X_BEST_OVF, X_LAST_IMPROVE, X_FALLBACK, X_COUNT = range(8)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(H, IN), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(OUT, H),
            p[B3:PARAMS])


def forward(p, x):
    """x [n, 12] -> o [n, 2] (unclipped), FP64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    h1 = np.maximum(np.asarray(x, np.float64) @ w1.T + b1, 0)
    h2 = np.maximum(h1 @ w2.T + b2, 0)
    return h2 @ w3.T + b3


def loss_grad(p, x, t, w):
    """(sum_s w_s |o_s - t_s|^2, its gradient in the parameter layout), FP64, by backpropagation."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    a1 = x @ w1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w2.T + b2
    h2 = np.maximum(a2, 0)
    o = h2 @ w3.T + b3
    e = o - np.asarray(t, np.float64)
    wt = np.asarray(w, np.float64)[:, None]
    loss = float((wt * e * e).sum())
    do = 2 * wt * e
    d2 = (do @ w3) * (a2 > 0)
    d1 = (d2 @ w2) * (a1 > 0)
    g = np.zeros(PARAMS)
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(0)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    g[W3:B3] = (do.T @ h2).ravel()
    g[B3:PARAMS] = do.sum(0)
    return loss, g


def features(S, c, X, moved, turned, hpwl, ovf, wa, wl2, dn2, phase):
    K, C = octl.K, octl.C
    alpha = moved / turned if turned > 0 else S[K["alpha"]]
    dens = S[K["lam"]] * c[C["total_area"]]
    hp = S[K["hpwl_prev"]]
    f = [ovf,
         100.0 * (ovf - X[X_PREV_OVF]),
         math.log(max(hpwl, 1e-30) / max(X[X_HPWL0], 1e-30)),
         10.0 * (hpwl - hp) / max(hp, 1e-30) if hp < 1e29 else 0.0,
         1.0 - wa / max(hpwl, 1e-30),
         math.log(max(S[K["gam"]], 1e-30) / max(c[C["base_gamma"]], 1e-30)),
         dens / (c[C["total_pins"]] + dens),
         math.log(max(alpha, 1e-30) / max(X[X_ALPHA0], 1e-30)),
         0.2 * math.log(math.sqrt(max(wl2, 1e-60) / max(dn2, 1e-60))),
         S[K["iter"]] / 1000.0,
         math.log1p(moved / c[C["motion_unit"]]),
         phase]
    return np.array([min(10.0, max(-10.0, v if math.isfinite(v) else 0.0)) for v in f])


_M64 = (1 << 64) - 1


def _mix(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def normal_draw(seed, n, k):
    a = _mix(seed ^ _mix((n * 4 + k) & _M64))
    b = _mix(a)
    u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0)
    u2 = (b >> 11) * (1.0 / 9007199254740992.0)
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)


def controller_pass(S, kp, c, X, p, moved, turned, hpwl, ovf, wa, wl2, dn2, phase=0.0, sigma=0.0, seed=0, window=50,
                    hold=1):
    """One pass with the learned controller (X: list of 8 floats, mutated). -> (features or None, outputs)."""
    K = octl.K
    x, o = None, (0.0, 0.0)
    lam_f = gam_f = 1.0
    if S[K["stop"]] == 0:
        if X[X_STARTED] == 0:
            X[X_STARTED] = 1
            X[X_PREV_OVF], X[X_HPWL0] = ovf, hpwl
            X[X_ALPHA0] = moved / turned if turned > 0 else S[K["alpha"]]
            X[X_BEST_OVF], X[X_LAST_IMPROVE] = ovf, S[K["iter"]]
        if X[X_FALLBACK] == 0 and all(math.isfinite(v) for v in (moved, turned, hpwl, ovf)):
            x = features(S, c, X, moved, turned, hpwl, ovf, wa, wl2, dn2, phase)
            o0, o1 = forward(p, x[None, :].astype(np.float32).astype(np.float64))[0]
            n = int(X[X_COUNT])
            if sigma > 0:
                o0 += sigma * normal_draw(seed, n // hold, 0)
                o1 += sigma * normal_draw(seed, n // hold, 1)
            o = (min(CLIP, max(-CLIP, o0)), min(CLIP, max(-CLIP, o1)))
            lam_f, gam_f = math.exp(o[0]), math.exp(o[1])
            X[X_COUNT] = n + 1
    octl.control_pass(S, kp, c, moved, turned, hpwl, ovf, wa, lam_f, gam_f)
    if S[K["accept"]] != 0:
        X[X_PREV_OVF] = ovf
        better = ovf < X[X_BEST_OVF] * 0.999
        if better:
            X[X_BEST_OVF] = ovf
        if better or S[K["ww"]] < 0.95:
            X[X_LAST_IMPROVE] = S[K["iter"]]
        elif S[K["iter"]] - X[X_LAST_IMPROVE] >= window:
            X[X_FALLBACK] = 1
    return x, o
