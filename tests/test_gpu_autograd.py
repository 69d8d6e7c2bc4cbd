"""The autograd energies on the GPU: backward against the FP64 oracle gradient and central finite differences; the
density energy's value against the oracle's closed form, and the gap between its exact gradient and the field
gather (the ePlace approximation the reference uses) reported."""
import numpy as np
import pytest
import torch

from conftest import norm_err, rel_err
from oracle import density as oden
from oracle import poisson as opo
from oracle import wa as owa

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def setup(small_design):
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators
    _, _, data = small_design
    args = default_args()
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV, "gemm")
    rng = np.random.default_rng(8)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    return data, prob, wl, den, pos


def test_wa_energy_backward_vs_fp64_and_finite_differences(setup):
    from autoplace_amd.autograd import WAEnergy
    data, prob, wl, den, pos = setup
    p = torch.tensor(pos, device=DEV, requires_grad=True)
    e = WAEnergy.apply(p, wl, 6.0)
    e.backward()
    g = p.grad.cpu().numpy()
    csr = owa.build_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov, 100)
    wa64, g64, _, _, _ = owa.wa_forward(pos.astype(np.float64), csr, 6.0, np.float64)
    n = data.num_mov
    assert rel_err(float(e), wa64) < 1e-5
    assert norm_err(g[:n], g64) < 1e-5 and np.all(g[n:] == 0)
    rng = np.random.default_rng(2)
    h, fd, an = 1e-3, [], []
    p64 = pos.astype(np.float64)
    for c in rng.choice(n, 1000, replace=False):
        a = int(rng.integers(2))
        p1, p2 = p64.copy(), p64.copy()
        p1[c, a] += h
        p2[c, a] -= h
        fd.append((owa.wa_forward(p1, csr, 6.0, np.float64)[0] - owa.wa_forward(p2, csr, 6.0, np.float64)[0]) / (2 * h))
        an.append(g[c, a])
    e_fd = norm_err(np.array(an), np.array(fd))
    print("WA energy: backward vs FP64 FD on 1000 cells: %.2e" % e_fd)
    assert e_fd < 1e-4


def test_density_energy_value_backward_and_fd_gap(setup):
    from autoplace_amd.autograd import DensityEnergy
    data, prob, wl, den, pos = setup
    nb = den.nbx
    bx, by = float(den.bin_np[0]), float(den.bin_np[1])
    p = torch.tensor(pos, device=DEV, requires_grad=True)
    e = DensityEnergy.apply(p, den)
    e.backward()
    g = p.grad.cpu().numpy()
    smooth, ratio = den.smooth["size"].cpu().numpy(), den.smooth["ratio"].cpu().numpy()
    fixed = den.fixed_map.cpu().numpy()
    m0, m1 = oden.density_maps(pos, smooth, ratio, data.num_mov, bx, by, nb, nb)
    rho, _ = oden.density_post(m0, m1, fixed, den.target, nb, nb)
    e64 = opo.density_energy(rho.astype(np.float64), bx, by)
    print("density energy: GPU %.8e, FP64 oracle %.8e (rel %.2e)" % (float(e), e64, rel_err(float(e), e64)))
    assert rel_err(float(e), e64) < 1e-5
    fx, fy = opo.field_dct(rho.astype(np.float64), bx, by)
    g_ref = oden.density_grad(pos, smooth, ratio, fx, fy, np.zeros((0, 2)), np.zeros(len(pos)),
                              np.ones(len(pos)), 1.0, -1.0, bx, by, nb, nb, exact=True)
    assert norm_err(g, g_ref) < 1e-5
    # exact discrete energy (no fixed-point rounding) by central differences on 150 nodes: the field gather is
    # ePlace's approximation of its gradient, so they agree in direction and scale, not to rounding
    rng = np.random.default_rng(3)
    p64 = pos.astype(np.float64)
    sel = rng.choice(len(pos), 150, replace=False)

    def energy(q):
        a0, a1 = oden.density_maps(q, smooth, ratio, data.num_mov, bx, by, nb, nb, exact=True)
        r, _ = oden.density_post(a0, a1, fixed.astype(np.float64), den.target, nb, nb, exact=True)
        return opo.density_energy(r, bx, by)

    h = 1e-4
    fd = np.zeros((len(sel), 2))
    for k, i in enumerate(sel):
        for a in range(2):
            p1, p2 = p64.copy(), p64.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd[k, a] = (energy(p1) - energy(p2)) / (2 * h)
    gs = g[sel].astype(np.float64)
    corr = np.corrcoef(fd.ravel(), gs.ravel())[0, 1]
    scale = float((fd * gs).sum() / (gs * gs).sum())
    print("density energy: exact FD vs field gather on 150 nodes: rel gap %.3f, correlation %.4f, scale %.3f"
          % (norm_err(gs, fd), corr, scale))
    assert corr > 0.9 and 0.8 < scale < 1.2


def test_composed_objective_gradient(setup):
    """W + lam D through autograd equals the kernels' WA gradient plus lam times the plain density gradient."""
    from autoplace_amd.autograd import DensityEnergy, WAEnergy
    data, prob, wl, den, pos = setup
    lam = 3e-3
    p = torch.tensor(pos, device=DEV, requires_grad=True)
    (WAEnergy.apply(p, wl, 6.0) + lam * DensityEnergy.apply(p, den)).backward()
    with torch.no_grad():
        q = torch.tensor(pos, device=DEV)
        _, wlg, _ = wl.forward(q, 6.0)
        dg, _ = den.forward(q)
        ref = lam * dg
        ref[:data.num_mov] += wlg
    assert norm_err(p.grad.cpu().numpy(), ref.cpu().numpy()) < 1e-6
