"""The controller oracle's MLP gradient against central finite differences (FP64), and its pass with zero outputs
against the plain rule."""
import math

import numpy as np

from oracle import control as octl
from oracle import mlp as omlp


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, omlp.PARAMS)
    x = rng.normal(0, 1, (37, omlp.IN))
    t = rng.normal(0, 0.1, (37, 2))
    w = rng.random(37)
    loss, g = omlp.loss_grad(p, x, t, w)
    assert np.isclose(loss, (w[:, None] * (omlp.forward(p, x) - t) ** 2).sum())
    h = 1e-6
    for k in list(rng.choice(omlp.PARAMS, 200, replace=False)) + [omlp.B3, omlp.B3 + 1, omlp.W1, omlp.B1]:
        q1, q2 = p.copy(), p.copy()
        q1[k] += h
        q2[k] -= h
        fd = (omlp.loss_grad(q1, x, t, w)[0] - omlp.loss_grad(q2, x, t, w)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_zero_network_is_the_plain_rule():
    """Zero parameters give zero outputs, factors exp(0) = 1: the pass equals control_pass bit for bit."""
    S0 = [0.0] * len(octl.FIELDS)
    K = octl.K
    S0[K["lam"]], S0[K["gam"]], S0[K["alpha"]], S0[K["a"]], S0[K["iter"]] = 1e-3, 20.0, 0.5, 3.0, 50
    S0[K["hpwl_prev"]], S0[K["best_hpwl"]], S0[K["best_ovf"]] = 1e30, 1e30, 1e30
    c = [20.0, 2.0, 4e4, 1e5, 1.0, 100, 0.07, 30, 200, 10000, 2.0, 1.0, 0.0]
    S1, kp1 = list(S0), np.zeros(5, np.float32)
    S2, kp2 = list(S0), np.zeros(5, np.float32)
    X = [0.0] * 8
    p = np.zeros(omlp.PARAMS)
    for i, (hp, ov) in enumerate([(1e6, 0.6), (9.8e5, 0.55), (9.9e5, 0.5), (9.7e5, 0.52)]):
        octl.control_pass(S1, kp1, c, 3.0 + i, 2.0, hp, ov, hp * 0.9)
        x, o = omlp.controller_pass(S2, kp2, c, X, p, 3.0 + i, 2.0, hp, ov, hp * 0.9, 4.0, 1.0)
        assert o == (0.0, 0.0) and x is not None and len(x) == 12
        assert S1 == S2 and np.array_equal(kp1, kp2)
    assert X[omlp.X_COUNT] == 4


def test_fallback_after_a_stalled_overflow():
    """With the overflow stuck the controller hands over to the rule after `window` passes -- counted once the
    density share is 0.95 or more, as the rule's stall test."""
    K = octl.K
    S = [0.0] * len(octl.FIELDS)
    S[K["lam"]], S[K["gam"]], S[K["alpha"]], S[K["a"]], S[K["iter"]] = 1e-3, 20.0, 0.5, 3.0, 10
    S[K["hpwl_prev"]], S[K["best_hpwl"]], S[K["best_ovf"]] = 1e30, 1e30, 1e30
    c = [20.0, 2.0, 4e4, 1e5, 1.0, 100, 0.07, 30, 1000, 10000, 2.0, 1.0, 0.0]
    kp = np.zeros(5, np.float32)
    X = [0.0] * 8
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.05, omlp.PARAMS)
    learned = 0
    for i in range(80):                            # density share 0.0025: no stall is counted
        x, o = omlp.controller_pass(S, kp, c, X, p, 3.0, 2.0, 1e6, 0.5, 9e5, 4.0, 1.0, window=50)
        learned += x is not None
    assert X[omlp.X_FALLBACK] == 0 and learned == 80
    S[K["lam"]] = 10.0                              # share 0.96: the clock runs
    for i in range(80):
        x, o = omlp.controller_pass(S, kp, c, X, p, 3.0, 2.0, 1e6, 0.5, 9e5, 4.0, 1.0, window=50)
        learned += x is not None
    assert X[omlp.X_FALLBACK] == 1 and learned == 80 + 50 and X[omlp.X_COUNT] == 130


def test_normal_draws_are_standard_normal():
    z = np.array([omlp.normal_draw(7, n, k) for n in range(4000) for k in (0, 1)])
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05
    assert omlp.normal_draw(7, 3, 1) == omlp.normal_draw(7, 3, 1) and math.isfinite(omlp.normal_draw(0, 0, 0))
