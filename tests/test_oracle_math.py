"""The oracle against what the mathematics fixes: closed forms, invariants, brute force on tiny inputs and central
finite differences (each check chosen so that a dropped term, a sign or an index error fails it)."""
import math

import numpy as np
import pytest

from conftest import norm_err, rel_err
from oracle import control as octl
from oracle import density as oden
from oracle import poisson as opo
from oracle import wa as owa


def _csr_one_net(xs, ys):
    n = len(xs)
    return owa.build_csr(np.arange(n), np.zeros((n, 2)), np.zeros(n, np.int64), 1, n, 100), \
        np.stack([xs, ys], 1).astype(np.float64)


@pytest.mark.parametrize("d,gamma", [(3.0, 1.0), (0.2, 5.0), (40.0, 2.0)])
def test_two_pin_closed_form(d, gamma):
    """A 2-pin net at distance d: WA = d tanh(d / 2 gamma), dWA/dx_hi = tanh(u) + u sech^2(u) with u = d / 2 gamma."""
    csr, pos = _csr_one_net([1.0, 1.0 + d], [0.0, 0.0])
    wa, grad, hpwl, _, _ = owa.wa_forward(pos, csr, gamma, np.float64)
    u = d / (2 * gamma)
    assert abs(wa - d * math.tanh(u)) < 1e-12 * d and abs(hpwl - d) < 1e-12
    g_hi = math.tanh(u) + u / math.cosh(u) ** 2
    assert abs(grad[1, 0] - g_hi) < 1e-12 and abs(grad[0, 0] + g_hi) < 1e-12
    assert grad[0, 1] == 0 and grad[1, 1] == 0


def test_wa_invariants_and_limits():
    rng = np.random.default_rng(0)
    xs, ys = rng.uniform(0, 50, 9), rng.uniform(0, 50, 9)
    csr, pos = _csr_one_net(xs, ys)
    hp = np.ptp(xs) + np.ptp(ys)
    prev = 0.0
    for gamma in (20.0, 5.0, 1.0, 0.1, 0.01):
        wa, grad, hpwl, _, _ = owa.wa_forward(pos, csr, gamma, np.float64)
        assert wa <= hpwl + 1e-12 and abs(hpwl - hp) < 1e-12
        assert np.abs(grad.sum(0)).max() < 1e-10           # translation invariance: the gradient sums to zero
        assert wa >= prev - 1e-12                           # WA grows to the HPWL as gamma -> 0
        prev = wa
    assert abs(wa - hp) < 1e-6 * hp
    # shift-free brute force of the definition agrees with the max / min-shifted form
    w1, g1 = owa.wa_brute(xs, 5.0)
    wa, grad, _, _, _ = owa.wa_forward(pos, csr, 5.0, np.float64)
    w2, g2 = owa.wa_brute(ys, 5.0)
    assert abs(wa - (w1 + w2)) < 1e-10 and np.abs(grad[:, 0] - g1).max() < 1e-12 and np.abs(grad[:, 1] - g2).max() < 1e-12


def test_wa_finite_differences(golden_small):
    """Central differences (step 1e-3 site, FP64) on 200 random cells of the 1.5k-cell design."""
    g = golden_small
    n = int(g["num_mov"])
    csr = owa.build_csr(g["pin_cell"], g["pin_off"], g["pin_net"], int(g["num_nets"]), n, 100)
    pos = g["pos_rand"].astype(np.float64)
    _, grad, _, _, _ = owa.wa_forward(pos, csr, 6.0, np.float64)
    rng = np.random.default_rng(1)
    h = 1e-3
    for c in rng.choice(n, 200, replace=False):
        for a in range(2):
            p1, p2 = pos.copy(), pos.copy()
            p1[c, a] += h
            p2[c, a] -= h
            fd = (owa.wa_forward(p1, csr, 6.0, np.float64)[0] - owa.wa_forward(p2, csr, 6.0, np.float64)[0]) / (2 * h)
            assert abs(fd - grad[c, a]) < 1e-6 + 1e-5 * abs(grad[c, a])


def test_density_total_charge_and_brute_force():
    rng = np.random.default_rng(2)
    nbx, nby, bx, by = 7, 5, 3.0, 4.0
    pos = np.stack([rng.uniform(3, 18, 40), rng.uniform(3, 17, 40)], 1).astype(np.float32)
    size = rng.uniform(0.5, 6, (40, 2)).astype(np.float32)
    ratio = rng.uniform(0.3, 1, 40).astype(np.float32)
    m0, m1 = oden.density_maps(pos, size, ratio, 25, bx, by, nbx, nby, exact=True)
    brute = oden.brute_overlap_map(pos[:25], size[:25], ratio[:25], bx, by, nbx, nby)
    assert np.abs(m0.reshape(nbx, nby) - brute).max() < 1e-12
    # charge inside the region, in bin areas
    pos, size, ratio = pos.astype(np.float64), size.astype(np.float64), ratio.astype(np.float64)
    x0, x1 = np.maximum(pos[:, 0] - size[:, 0] / 2, 0), np.minimum(pos[:, 0] + size[:, 0] / 2, nbx * bx)
    y0, y1 = np.maximum(pos[:, 1] - size[:, 1] / 2, 0), np.minimum(pos[:, 1] + size[:, 1] / 2, nby * by)
    charge = ratio * np.maximum(x1 - x0, 0) * np.maximum(y1 - y0, 0) / (bx * by)
    assert abs(m0.sum() + m1.sum() - charge.sum()) < 1e-9
    # fixed point: within 2^-32 per (node, bin) of the exact map
    q0, q1 = oden.density_maps(pos.astype(np.float32), size.astype(np.float32), ratio.astype(np.float32), 25, bx, by,
                               nbx, nby)
    assert np.abs(q0 / oden.SCALE - m0).max() < 40 * 12 * 2 ** -32 + 1e-6


def test_aligned_node_pattern():
    """A node covering exactly bins [1, 3) x [2, 3) puts ratio in those bins and nothing elsewhere."""
    m0, _ = oden.density_maps(np.array([[4.0, 2.5]], np.float32), np.array([[4.0, 1.0]], np.float32),
                              np.array([0.5], np.float32), 1, 2.0, 1.0, 4, 4)
    m = m0.reshape(4, 4).astype(np.float64) / oden.SCALE
    expect = np.zeros((4, 4))
    expect[1:3, 2] = 0.5
    assert np.array_equal(m, expect)


@pytest.mark.parametrize("N,M,bx,by", [(8, 8, 1.0, 1.0), (16, 32, 3.7, 2.1), (64, 64, 12.5, 12.5)])
def test_dct_form_equals_mirrored_fft(N, M, bx, by):
    rho = np.random.default_rng(N).random((N, M))
    f = opo.field_mirrored_fft(oden.mirror(rho), opo.to_field(N, M, bx, by))[:, :N, :M]
    fx, fy = opo.field_dct(rho, bx, by)
    assert np.abs(f[0] - fx).max() < 1e-12 and np.abs(f[1] - fy).max() < 1e-12


def test_field_of_a_cosine_mode_and_of_a_uniform_map():
    N, M, bx, by, p = 64, 32, 2.0, 3.0, 3
    x = np.arange(N)
    rho = np.repeat(np.cos(np.pi * p * (x + 0.5) / N)[:, None], M, 1) + 1.0
    f = opo.field_mirrored_fft(oden.mirror(rho), opo.to_field(N, M, bx, by))[:, :N, :M]
    w = np.pi * p / (N * bx)
    expect = np.repeat((np.sin(np.pi * p * (x + 0.5) / N) / (w * bx))[:, None], M, 1)
    assert np.abs(f[0] - expect).max() < 1e-12 and np.abs(f[1]).max() < 1e-12
    f = opo.field_mirrored_fft(oden.mirror(np.full((N, M), 0.3)), opo.to_field(N, M, bx, by))
    assert np.abs(f).max() < 1e-14


def test_gather_linearity_and_brute_force():
    rng = np.random.default_rng(4)
    nb, bx = 6, 2.0
    pos = rng.uniform(2, 10, (12, 2)).astype(np.float32)
    size = rng.uniform(0.5, 3, (12, 2)).astype(np.float32)
    ratio = np.ones(12, np.float32)
    f1, f2 = rng.normal(size=(2, 2, nb, nb)).astype(np.float64)
    z = np.zeros((0, 2))
    args = (np.zeros(12), np.ones(12), 1.0, -1.0, bx, bx, nb, nb)
    g1 = oden.density_grad(pos, size, ratio, f1[0], f1[1], z, *args, exact=True)
    g2 = oden.density_grad(pos, size, ratio, f2[0], f2[1], z, *args, exact=True)
    g12 = oden.density_grad(pos, size, ratio, f1[0] + 2 * f2[0], f1[1] + 2 * f2[1], z, *args, exact=True)
    assert np.abs(g12 - (g1 + 2 * g2)).max() < 1e-12
    for i in range(12):
        w = oden.brute_overlap_map(pos[i:i + 1], size[i:i + 1], ratio[i:i + 1], bx, bx, nb, nb)
        assert abs(g1[i, 0] + (w * f1[0]).sum()) < 1e-10 and abs(g1[i, 1] + (w * f1[1]).sum()) < 1e-10


def test_control_hand_cases():
    """Hand-derived single passes: BB step, growth when the HPWL falls, diverged on NaN, frozen once stopped."""
    K = octl.K
    S = [0.0] * 21
    S[K["lam"]], S[K["gam"]], S[K["a"]], S[K["hpwl_prev"]] = 1e-3, 10.0, 1.0, 200.0
    S[K["best_hpwl"]] = S[K["best_ovf"]] = math.inf
    c = [4.0, 100.0, 1000.0, 5000.0, 2.0, 0, 0.07, 3, 50, 1000, 1.0, 1.0, 0.0]
    kp = np.zeros(5, np.float32)
    octl.control_pass(S, kp, c, 6.0, 3.0, 150.0, 0.5, 140.0)
    assert S[K["alpha"]] == 2.0 and S[K["accept"]] == 1 and S[K["evals"]] == 1
    assert S[K["mu"]] == 1.05 * 0.98 and S[K["lam"]] == 1e-3 * (1.05 * 0.98)
    assert S[K["a"]] == (1 + math.sqrt(5)) / 2 and S[K["iter"]] == 1 and S[K["is_best"]] == 1
    assert kp[3] == np.float32(2.0)
    octl.control_pass(S, kp, c, float("nan"), 1.0, 150.0, 0.5, 140.0)
    assert S[K["stop"]] == 3 and S[K["accept"]] == 0
    before = list(S)
    octl.control_pass(S, kp, c, 1.0, 1.0, 100.0, 0.01, 99.0)
    assert S == before
