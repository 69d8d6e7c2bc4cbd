"""Parity of the HIP kernels (through the C ABI) with the oracle, on the GPU.

Tolerances: integer / index work and the fixed-point density map are bit-exact; FP32 results are within 1e-5
relative (max-norm or 2-norm as stated) of the FP32 oracle, and their error against the FP64 oracle is printed."""
import math

import numpy as np
import pytest
import torch

from conftest import norm_err, rel_err
from oracle import control as octl
from oracle import density as oden
from oracle import poisson as opo
from oracle import wa as owa

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(a, dt=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)


@pytest.fixture(scope="module")
def small(small_design):
    from autoplace_amd.problem import build_problem, default_args
    _, _, data = small_design
    p = build_problem(data, default_args())
    nb = p.force_bin
    bx, by = data.width / nb, data.height / nb
    smooth, ratio = oden.smooth_nodes(p.node_size, bx, by)
    rng = np.random.default_rng(3)
    lo, hi = smooth / 2, np.array([data.width, data.height], np.float32) - smooth / 2
    pos = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    return data, p, pos, smooth, ratio, bx, by


@pytest.mark.parametrize("planned", [True, False])
@pytest.mark.parametrize("gamma", [0.5, 4.0, 60.0])
def test_wa_parity(small, gamma, planned):
    """gp_cuda.wa with the netlist's plan (lane groups per net) and without it (one lane per net)."""
    from autoplace_amd import gp_cuda
    data, p, pos, *_ = small
    csr = owa.build_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov, 100)
    wa, grad, hpwl, out, pin_grad = owa.wa_forward(pos, csr, gamma, np.float32)
    wa64, grad64, hpwl64, out64, _ = owa.wa_forward(pos, csr, gamma, np.float64)
    g_dev, out_dev = gp_cuda.wa(_t(pos), _t(csr["cell"], torch.int32), _t(csr["off"]), _t(csr["net_start"], torch.int32),
                                _t(csr["cell_start"], torch.int32), _t(csr["cell_pin"], torch.int32), data.num_mov,
                                _t([gamma]), planned=planned)
    g_dev, out_dev = g_dev.cpu().numpy(), out_dev.cpu().numpy()
    e_out, e_grad = rel_err(out_dev, out), norm_err(g_dev, grad)
    print("WA gamma %g (%s): out err vs fp32 %.2e (vs fp64 %.2e); grad err vs fp32 %.2e (vs fp64 %.2e)"
          % (gamma, "planned" if planned else "plan-free", e_out, rel_err(out_dev, out64), e_grad,
             norm_err(g_dev, grad64)))
    assert e_out < 1e-5 and e_grad < 1e-5
    assert abs(out_dev[:, 1].astype(np.float64).sum() - hpwl) / hpwl < 1e-6


@pytest.mark.parametrize("planned", [True, False])
def test_wa_big_net_and_fixed_pins(planned):
    """A net larger than the LDS stage and the 64-lane groups (the block-wide path), fixed pins, nets of one pin,
    degree-2 nets around them."""
    from autoplace_amd import gp_cuda
    rng = np.random.default_rng(9)
    n = 3000
    deg = [2, 3, 700, 2, 1, 1500, 5, 64, 65, 2, 1]
    cells = np.concatenate([rng.choice(n + 1, d, replace=False) for d in deg])     # n = fixed pin
    net = np.repeat(np.arange(len(deg)), deg)
    off = rng.uniform(-2, 2, (len(cells), 2)).astype(np.float32)
    off[cells == n] = rng.uniform(0, 500, (int((cells == n).sum()), 2))
    pos = rng.uniform(0, 500, (n, 2)).astype(np.float32)
    csr = owa.build_csr(cells, off, net, len(deg), n, 10 ** 6)
    wa, grad, hpwl, out, _ = owa.wa_forward(pos, csr, 7.0, np.float32)
    g_dev, out_dev = gp_cuda.wa(_t(pos), _t(csr["cell"], torch.int32), _t(csr["off"]), _t(csr["net_start"], torch.int32),
                                _t(csr["cell_start"], torch.int32), _t(csr["cell_pin"], torch.int32), n, _t([7.0]),
                                planned=planned)
    assert rel_err(out_dev.cpu().numpy(), out) < 1e-5
    assert norm_err(g_dev.cpu().numpy(), grad) < 1e-5


def test_wa_plan_follows_the_netlist():
    """The drop-in's WA plan: nets of no pin give (0, 0) as the plan-free kernel does, every output agrees with the
    plan-free kernel's to float rounding, and a netlist changed in place (its tensors' version counters) is planned
    again: the next call sees the new pins."""
    from autoplace_amd import gp_cuda
    rng = np.random.default_rng(4)
    n = 500
    deg = rng.integers(2, 9, 300)
    deg[[3, 77, 299]] = 0                             # empty nets, one of them the last
    deg[[10, 11]] = [64, 1]
    net_start = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    cells = rng.integers(0, n + 1, int(deg.sum())).astype(np.int32)    # n = fixed pin
    off = rng.uniform(-2, 2, (len(cells), 2)).astype(np.float32)
    pos = _t(rng.uniform(0, 300, (n, 2)).astype(np.float32))
    order = np.argsort(cells, kind="stable")
    cell_start = np.searchsorted(cells[order], np.arange(n + 1)).astype(np.int32)
    t_cells, t_off, t_ns = _t(cells, torch.int32), _t(off), _t(net_start, torch.int32)
    t_cs, t_cp = _t(cell_start, torch.int32), _t(order.astype(np.int32), torch.int32)
    gamma = _t([5.0])

    def both():
        a = gp_cuda.wa(pos, t_cells, t_off, t_ns, t_cs, t_cp, n, gamma, planned=True)
        b = gp_cuda.wa(pos, t_cells, t_off, t_ns, t_cs, t_cp, n, gamma, planned=False)
        return [x.cpu().numpy() for x in a], [x.cpu().numpy() for x in b]

    (g1, o1), (g0, o0) = both()
    assert np.array_equal(o1[[3, 77, 299]], np.zeros((3, 2), np.float32))
    assert np.array_equal(o0[[3, 77, 299]], np.zeros((3, 2), np.float32))
    assert rel_err(o1, o0) < 1e-5 and norm_err(g1, g0) < 1e-5
    t_off.mul_(3.0)                                   # the same tensors, new offsets: a new plan
    (g2, o2), (h0, p0) = both()
    assert rel_err(o2, p0) < 1e-5 and norm_err(g2, h0) < 1e-5
    assert rel_err(o2, o1) > 1e-3


def test_density_map_bit_exact(small):
    from autoplace_amd import gp_cuda
    data, p, pos, smooth, ratio, bx, by = small
    nb = p.force_bin
    fixed = (p.fixed_fraction * p.target).astype(np.float32)
    m0, m1 = oden.density_maps(pos, smooth, ratio, data.num_mov, bx, by, nb, nb)
    rho, over = oden.density_post(m0, m1, fixed, p.target, nb, nb)
    rho_dev, over_dev = gp_cuda.density_map(_t(pos), _t(smooth), _t(ratio), data.num_mov, bx, by, nb, nb, _t(fixed),
                                            p.target)
    assert np.array_equal(rho_dev.cpu().numpy(), oden.mirror(rho))
    assert np.array_equal(over_dev.cpu().numpy(), over)


def _odd_nodes(seed, n, nbx, nby, bx, by, big=12):
    """Nodes that test the tile binning's edge cases: small ones, ones over tile corners, a few far wider than a
    tile (the big-node path), ones hanging over the region's edges, zero-size ones."""
    rng = np.random.default_rng(seed)
    W, H = nbx * bx, nby * by
    size = np.stack([rng.uniform(0.2, 1.8, n) * bx, rng.uniform(0.2, 1.8, n) * by], 1).astype(np.float32)
    size[:big] = np.stack([rng.uniform(20, 90, big) * bx, rng.uniform(20, 90, big) * by], 1)
    size[big:2 * big] = 0
    pos = np.stack([rng.uniform(-2 * bx, W + 2 * bx, n), rng.uniform(-2 * by, H + 2 * by, n)], 1).astype(np.float32)
    ratio = rng.uniform(0.3, 1.0, n).astype(np.float32)
    return pos, size, ratio


@pytest.mark.parametrize("case", [(20000, 15000, 96, 80, 2.0, 3.0), (20000, 20000, 64, 64, 1.0, 1.0),
                                  (3000, 1000, 7, 300, 5.0, 0.5), (1, 1, 3, 3, 1.0, 1.0), (0, 0, 16, 16, 1.0, 1.0)])
def test_density_map_edge_cases_bit_exact(case):
    """The map of the tile binning against the oracle's per-node scatter, bit for bit: big nodes, nodes outside the
    region, zero-size nodes, fillers (num_mov < num_nodes) or none, ragged grids, one node, no node."""
    from autoplace_amd import gp_cuda
    n, num_mov, nbx, nby, bx, by = case
    pos, size, ratio = _odd_nodes(n + nbx, n, nbx, nby, bx, by, big=min(12, n // 3))
    fixed = np.random.default_rng(1).uniform(0, 0.5, (nbx, nby)).astype(np.float32)
    target = 0.9
    m0, m1 = oden.density_maps(pos, size, ratio, num_mov, bx, by, nbx, nby)
    rho, over = oden.density_post(m0, m1, fixed, target, nbx, nby)
    for _ in range(2):                       # the second call reuses torch's cached workspace (stale contents)
        rho_dev, over_dev = gp_cuda.density_map(_t(pos), _t(size), _t(ratio), num_mov, bx, by, nbx, nby, _t(fixed),
                                                target)
        assert np.array_equal(rho_dev.cpu().numpy(), oden.mirror(rho))
        assert np.array_equal(over_dev.cpu().numpy(), over)


def test_density_map_caller_map_as_workspace(small):
    """ap_density_map (the reference's signature, with its u64 map): the map buffer serves as the workspace when it is
    large enough, else a stream-ordered one is taken; both give the oracle's map."""
    import ctypes
    from autoplace_amd import _lib
    data, p, pos, smooth, ratio, bx, by = small
    nb = p.force_bin
    fixed = (p.fixed_fraction * p.target).astype(np.float32)
    m0, m1 = oden.density_maps(pos, smooth, ratio, data.num_mov, bx, by, nb, nb)
    rho, over = oden.density_post(m0, m1, fixed, p.target, nb, nb)
    need = _lib.lib().ap_density_map_work_bytes(len(pos), data.num_mov, nb, nb)
    assert need > 0
    V = lambda t: ctypes.c_void_p(t.data_ptr())       # noqa: E731
    tp, ts, tr, tf = _t(pos), _t(smooth), _t(ratio), _t(fixed)      # held: their memory must outlive the call
    for words in (2 * nb * nb, (need + 7) // 8):       # the reference's map size, and one that holds the workspace
        m = torch.full((words,), -1, dtype=torch.int64, device=DEV)
        r = torch.empty((2 * nb, 2 * nb), device=DEV)
        o = torch.empty((nb, nb), device=DEV)
        st = _lib.lib().ap_density_map(V(tp), V(ts), V(tr), len(pos), data.num_mov, bx, by, nb, nb, V(m), V(tf),
                                       p.target, V(r), V(o), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        _lib.check(st, "density_map")
        assert np.array_equal(r.cpu().numpy(), oden.mirror(rho)) and np.array_equal(o.cpu().numpy(), over)


def test_density_grad_edge_cases():
    """The gather of the drop-in density_grad on nodes wider than 3 bins, outside the region, zero-size, and
    num_wl < num_nodes (fillers get no wirelength term)."""
    from autoplace_amd import gp_cuda
    n, nbx, nby, bx, by = 5000, 40, 56, 2.0, 1.5
    pos, size, ratio = _odd_nodes(7, n, nbx, nby, bx, by)
    rng = np.random.default_rng(8)
    fx, fy = rng.normal(size=(2, nbx, nby)).astype(np.float32)
    field = np.zeros((2, 2 * nbx, 2 * nby), np.float32)
    field[0, :nbx, :nby], field[1, :nbx, :nby] = fx, fy
    num_wl = 3000
    wl = rng.normal(size=(num_wl, 2)).astype(np.float32)
    pins = rng.integers(0, 6, n).astype(np.float32)
    area = size.prod(1).astype(np.float32)
    ref = oden.density_grad(pos, size, ratio, fx, fy, wl, pins, area, 0.7, 0.01, bx, by, nbx, nby)
    g = gp_cuda.density_grad(_t(pos), _t(size), _t(ratio), _t(field), _t(wl), _t(pins), _t(area), _t([0.7]),
                             _t([0.01]), bx, by, nbx, nby).cpu().numpy()
    assert norm_err(g, ref) < 1e-5


@pytest.mark.parametrize("shape", [(16, 16, 3.0, 3.0), (64, 64, 12.5, 12.5), (128, 128, 6.3, 6.3), (48, 40, 7.0, 5.5),
                                   (512, 512, 1.7, 1.7)])
def test_poisson_gemm(shape):
    from autoplace_amd import gp_cuda
    N, M, bx, by = shape
    rng = np.random.default_rng(N * 7 + M)
    rho = (rng.random((N, M)) * 2).astype(np.float32)
    ref = opo.field_mirrored_fft(oden.mirror(rho.astype(np.float64)), opo.to_field(N, M, bx, by))[:, :N, :M]
    ref32 = opo.field_mirrored_fft(oden.mirror(rho), opo.to_field(N, M, bx, by, np.float32))[:, :N, :M]
    plan = gp_cuda.PoissonPlan(N, M, bx, by, "gemm")
    f = plan.field(_t(rho)).cpu().numpy()
    e64, e32 = norm_err(f, ref), norm_err(ref32, ref)
    print("poisson gemm %dx%d: err vs fp64 %.2e (fp32 FFT oracle %.2e)" % (N, M, e64, e32))
    assert e64 < 1e-5


@pytest.mark.parametrize("shape", [(2, 2, 1.0, 1.0), (16, 16, 3.0, 3.0), (64, 32, 12.5, 4.0), (128, 128, 6.3, 6.3),
                                   (512, 512, 1.7, 1.7), (2048, 2048, 6.0, 6.0)])
def test_poisson_fft(shape):
    from autoplace_amd import gp_cuda
    N, M, bx, by = shape
    rng = np.random.default_rng(N * 3 + M)
    rho = (rng.random((N, M)) * 2).astype(np.float32)
    ref = opo.field_mirrored_fft(oden.mirror(rho.astype(np.float64)), opo.to_field(N, M, bx, by))[:, :N, :M]
    plan = gp_cuda.PoissonPlan(N, M, bx, by, "fft")
    assert plan.method == "fft"
    f = plan.field(_t(rho)).cpu().numpy()
    e64 = norm_err(f, ref)
    print("poisson fft %dx%d: err vs fp64 %.2e" % (N, M, e64))
    assert e64 < 1e-5


def test_poisson_methods_agree_on_a_cosine_mode():
    """rho = 1 + cos(pi p (n + 1/2) / N) along x has the field field_x = sin(pi p (n + 1/2) / N) / (w_p bx) with
    w_p = pi p / (N bx), and no field along y (both methods)."""
    from autoplace_amd import gp_cuda
    N = M = 64
    bx = by = 2.0
    x = np.arange(N)
    rho = np.repeat(np.cos(np.pi * 3 * (x + 0.5) / N)[:, None], M, 1).astype(np.float32) + 1.0
    fg = gp_cuda.PoissonPlan(N, M, bx, by, "gemm").field(_t(rho)).cpu().numpy()
    ff = gp_cuda.PoissonPlan(N, M, bx, by, "fft").field(_t(rho)).cpu().numpy()
    w = np.pi * 3 / (N * bx)
    expect = np.repeat((np.sin(np.pi * 3 * (x + 0.5) / N) * w / (w * w * bx))[:, None], M, 1)
    assert norm_err(ff[0], expect) < 1e-5 and norm_err(fg[0], expect) < 1e-5
    assert np.abs(ff[1]).max() < 1e-5 * np.abs(expect).max() and np.abs(fg[1]).max() < 1e-5 * np.abs(expect).max()


def test_poisson_uniform_map_has_no_field():
    from autoplace_amd import gp_cuda
    for method in ("gemm", "fft"):
        f = gp_cuda.PoissonPlan(32, 32, 1.5, 1.5, method).field(_t(np.full((32, 32), 0.7, np.float32))).cpu().numpy()
        assert np.abs(f).max() < 1e-6


def test_poisson_mirrored_layout(small):
    """poisson_field on the reference's mirrored map writes the [:, :nbx, :nby] quadrant the gather reads."""
    from autoplace_amd import gp_cuda
    N = 32
    rng = np.random.default_rng(1)
    rho = rng.random((N, N)).astype(np.float32)
    plan = gp_cuda.PoissonPlan(N, N, 2.0, 2.0, "gemm")
    field = gp_cuda.poisson_field(_t(oden.mirror(rho)), plan).cpu().numpy()[:, :N, :N]
    ref = opo.field_mirrored_fft(oden.mirror(rho.astype(np.float64)), opo.to_field(N, N, 2.0, 2.0))[:, :N, :N]
    assert norm_err(field, ref) < 1e-5


def test_density_grad_parity(small):
    from autoplace_amd import gp_cuda
    data, p, pos, smooth, ratio, bx, by = small
    nb = p.force_bin
    rng = np.random.default_rng(5)
    fx, fy = rng.normal(size=(2, nb, nb)).astype(np.float32)
    field = np.zeros((2, 2 * nb, 2 * nb), np.float32)
    field[0, :nb, :nb], field[1, :nb, :nb] = fx, fy
    wl = rng.normal(size=(data.num_mov, 2)).astype(np.float32)
    pins = p.node_pins.astype(np.float32)
    area = p.node_size.prod(1).astype(np.float32)
    for weight, precond in ((2.5e-3, 2.5e-3), (1.0, -1.0)):
        ref = oden.density_grad(pos, smooth, ratio, fx, fy, wl, pins, area, weight, precond, bx, by, nb, nb)
        g = gp_cuda.density_grad(_t(pos), _t(smooth), _t(ratio), _t(field), _t(wl), _t(pins), _t(area), _t([weight]),
                                 _t([precond]), bx, by, nb, nb).cpu().numpy()
        assert norm_err(g, ref) < 1e-5


def test_nesterov_and_select():
    from autoplace_amd import gp_cuda
    rng = np.random.default_rng(2)
    n = 10007
    u, v, g = (rng.normal(size=(n, 2)).astype(np.float32) for _ in range(3))
    lo, hi = np.full((n, 2), -1.5, np.float32), np.full((n, 2), 1.5, np.float32)
    un, vn = octl.nesterov_step(u, v, g, lo, hi, 0.37, 0.61)
    du, dv = gp_cuda.nesterov_step(_t(u), _t(v), _t(g), _t(lo), _t(hi), _t([0.37]), _t([0.61]))
    assert rel_err(du.cpu().numpy(), un) < 1e-6 and rel_err(dv.cpu().numpy(), vn) < 1e-6
    S = torch.zeros(21, dtype=torch.float64, device=DEV)
    tu, tv, tg, best = _t(u), _t(v), _t(g), _t(np.zeros((n, 2)))
    gn = _t(g * 2)
    gp_cuda.select(S, tu, tv, tg, du, dv, gn, best)          # not accepted: nothing moves
    assert torch.equal(tu, _t(u)) and torch.equal(best, _t(np.zeros((n, 2))))
    S[octl.K["accept"]] = 1
    S[octl.K["is_best"]] = 1
    gp_cuda.select(S, tu, tv, tg, du, dv, gn, best)
    assert torch.equal(tu, du) and torch.equal(tv, dv) and torch.equal(tg, gn) and torch.equal(best, dv)


@pytest.mark.parametrize("offset", [0, 1])
def test_nesterov_and_select_odd_and_unaligned(offset):
    """An odd count of coordinates, and tensors that start 4 bytes past an 8-byte boundary (the one-coordinate-per-
    lane kernels): the step equals the oracle's to the bit, select moves every coordinate."""
    from autoplace_amd import gp_cuda
    rng = np.random.default_rng(12)
    n = 20001

    def dev(a):
        buf = torch.empty(n + offset, dtype=torch.float32, device=DEV)
        buf[offset:] = torch.as_tensor(a, dtype=torch.float32)
        return buf[offset:]
    u, v, g = (rng.normal(size=n).astype(np.float32) for _ in range(3))
    lo, hi = np.full(n, -1.2, np.float32), np.full(n, 1.3, np.float32)
    un, vn = octl.nesterov_step(u, v, g, lo, hi, 0.37, 0.61)
    du, dv = gp_cuda.nesterov_step(dev(u), dev(v), dev(g), dev(lo), dev(hi), _t([0.37]), _t([0.61]))
    assert rel_err(du.cpu().numpy(), un) < 1e-6 and rel_err(dv.cpu().numpy(), vn) < 1e-6
    S = torch.zeros(21, dtype=torch.float64, device=DEV)
    S[octl.K["accept"]] = 1
    S[octl.K["is_best"]] = 1
    tu, tv, tg, best, gn = dev(u), dev(v), dev(g), dev(np.zeros(n)), dev(g * 2)
    dun, dvn = dev(un), dev(vn)
    gp_cuda.select(S, tu, tv, tg, dun, dvn, gn, best)
    assert torch.equal(tu, dun) and torch.equal(tv, dvn) and torch.equal(tg, gn) and torch.equal(best, dvn)


def test_control_matches_oracle_fp64():
    """ap_control against control_pass over a scripted sequence: S and kp bit for bit (pow may differ by an ulp)."""
    from autoplace_amd import gp_cuda
    rng = np.random.default_rng(4)
    K = octl.K
    S = [0.0] * 21
    S[K["lam"]], S[K["gam"]], S[K["alpha"]], S[K["a"]], S[K["iter"]] = 3e-4, 40.0, 0.8, 1.0, 12
    S[K["hpwl_prev"]], S[K["best_hpwl"]], S[K["best_ovf"]], S[K["loop_it"]] = 1e5, math.inf, math.inf, 12
    c = [4.0, 700.0, 4.0e4, 9.0e5, 7.0, 40, 0.07, 10, 30, 400, 2.0, 1.0, 0.0]
    kp = np.zeros(5, np.float32)
    dS = torch.tensor(S, dtype=torch.float64, device=DEV)
    dkp = torch.zeros(5, device=DEV)
    dc = torch.tensor(c, dtype=torch.float64, device=DEV)
    T = 300
    ovf = np.concatenate([np.linspace(0.9, 0.06, 150), 0.05 + 0.03 * rng.random(T - 150)]).astype(np.float32)
    hp = (1e5 * (1 + 0.03 * np.sin(np.arange(T) / 4.0))).astype(np.float32)
    moved = (40 * np.exp(-np.arange(T) / 60) + rng.random(T)).astype(np.float32)
    turned = (moved * (0.5 + rng.random(T))).astype(np.float32)
    wa = (hp * (1 - 0.1 * np.exp(-np.arange(T) / 50))).astype(np.float32)
    for t in range(T):
        octl.control_pass(S, kp, c, moved[t], turned[t], hp[t], ovf[t], wa[t])
        gp_cuda.control(dS, dkp, dc, _t([moved[t]]), _t([turned[t]]), _t([hp[t]]), _t([ovf[t]]), _t([wa[t]]))
        got = dS.cpu().numpy()
        for k, name in enumerate(octl.FIELDS):
            assert got[k] == S[k] or abs(got[k] - S[k]) <= 4e-16 * abs(S[k]), (t, name, got[k], S[k])
        assert np.array_equal(dkp.cpu().numpy(), kp) or rel_err(dkp.cpu().numpy(), kp) < 1e-7
    assert S[K["stop"]] != 0


def test_net_hpwl_exact(small):
    from autoplace_amd.wirelength import Wirelength
    data, p, pos, *_ = small
    wl = Wirelength(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov, 100, DEV)
    got = wl.hpwl(_t(pos)).item()
    ref = owa.hpwl_exact(pos, data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov)
    assert abs(got - ref) / ref < 1e-6
