"""Parity at the stated workload sizes, on the GPU, against the oracle:

  B  200,000 cells, 210,000 nets, 8 macros on 20 % of the die, overflow grid 512, force grid 1024: the drop-in
     kernels (WA, density map of both grids, Poisson, gather) and the engine's objective and map; WA and density
     gradients against central finite differences on 1,000 random cells (step 1e-3 site).
  C  2,000,000 cells, 2,100,000 nets, 16 macros on 15 %, overflow grid 512, force grid 2048: the engine's objective
     and its 2048 x 2048 map (bit for bit); the FFT and the FP32-MFMA GEMM Poisson solves at 2048 x 2048; the whole
     flow against the reference's own run of the same design (tests/golden/golden_runC.json).
  D  64 blocks of 2,000 .. 50,000 cells (log-uniform), grids 64 .. 256: batched placement against 64 sequential runs.

and the reference's own loop body (src/gpu_loop.py:58-64) composed from the six drop-in calls, 16 passes against the
oracle's trajectory.

Tolerances: the fixed-point density maps are bit-exact; FP32 quantities are within 1e-5 relative of the FP32 oracle
(2-norm for vectors); FP64 sums of the control within 1e-6."""
import json
import math
import os
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from conftest import GOLDEN, norm_err, rel_err
from oracle import control as octl
from oracle import density as oden
from oracle import place as opl
from oracle import poisson as opo
from oracle import wa as owa

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(a, dt=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)


def _config(name, grids=(0, 0)):
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators
    print("config %s: generating" % name, flush=True)
    lef, d = synth.generate_config(name, seed=0)
    data = PlaceData(lef, d)
    args = default_args(num_bin=grids[0], force_num_bin=grids[1])
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV, "fft")
    rng = np.random.default_rng(17)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    k = len(pos) // 3                                # a third of the nodes in one quarter: dense bins, long buckets
    pos[:k] = lo[:k] + 0.5 * (hi - lo)[:k] * rng.random((k, 2))
    csr = owa.build_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov,
                        args.ignore_net_degree)
    print("config %s: %d cells, %d nodes, %d pins, grids %d / %d" % (name, data.num_mov, len(pos), data.num_pins,
                                                                     prob.num_bin, prob.force_bin), flush=True)
    return SimpleNamespace(name=name, data=data, args=args, prob=prob, wl=wl, den=den, pos=pos, csr=csr)


def _oracle_level(c):
    p, n = c.prob, c.data.num_mov
    judge = None if p.force_bin == p.num_bin else {"nb": p.num_bin, "fraction": p.judge_fraction,
                                                    "size": p.node_size[:n]}
    return opl.Level(p.width, p.height, p.force_bin, p.fixed_fraction, p.target, p.node_size, n, p.node_pins,
                     p.pin_cell, p.pin_off, p.pin_net, p.num_nets, c.args.ignore_net_degree, judge)


def _engine_objective(c, pos, kp):
    """The engine's objective at pos (the product's node renumbering on) -> (gradient in the caller's node order,
    [wa, hpwl, overflow], force-grid density map [nbx, nby])."""
    from autoplace_amd.gpu_loop import Engine
    S = torch.zeros(21, dtype=torch.float64, device=DEV)
    cc = torch.zeros(13, dtype=torch.float64, device=DEV)
    v = _t(pos)
    eng = Engine(c.wl, c.den, v.clone(), v, torch.zeros_like(v), v.clone(), S, _t(kp), cc, "fft", True)
    grad = torch.empty_like(eng.v)
    stats = torch.zeros(8, dtype=torch.float64, device=DEV)
    eng.objective(eng.v, grad, stats)
    rho = eng.density()
    torch.cuda.synchronize()
    out = torch.empty_like(grad)
    out[eng.perm] = grad
    res = out.cpu().numpy(), stats.cpu().numpy()[:3], rho.cpu().numpy()
    eng.close()
    return res


def _check_engine_objective(c, kp):
    lvl = _oracle_level(c)
    g_ref, wa, hpwl, ovf = opl._evaluate_kp(lvl, c.pos, np.asarray(kp, np.float32))
    g, st, rho = _engine_objective(c, c.pos, kp)
    e = norm_err(g, g_ref)
    print("config %s engine objective: grad %.2e, wa %.2e, hpwl %.2e, overflow %.2e"
          % (c.name, e, rel_err(st[0], wa), rel_err(st[1], hpwl), rel_err(st[2], ovf)), flush=True)
    assert e < 1e-5
    assert rel_err(st[0], wa) < 1e-6 and rel_err(st[1], hpwl) < 1e-6 and rel_err(st[2], ovf) < 1e-6
    nb = c.den.nbx
    m0, m1 = oden.density_maps(c.pos, lvl.smooth, lvl.ratio, c.data.num_mov, lvl.bx, lvl.by, nb, nb)
    rho_ref, _ = oden.density_post(m0, m1, lvl.fixed, c.prob.target, nb, nb)
    assert np.array_equal(rho, rho_ref)


def _net_wa64(xs, gamma):
    """WA of one net along one axis in float64, exponents shifted by the max / min (definition of the WA model)."""
    hi, lo = xs.max(), xs.min()
    ep, en = np.exp((xs - hi) / gamma), np.exp((lo - xs) / gamma)
    return (xs * ep).sum() / ep.sum() - (xs * en).sum() / en.sum()


def _wa_fd(c, cells, gamma, h=1e-3):
    """Central differences in float64 of the WA of the optimized nets of each cell, per axis."""
    csr, n = c.csr, c.data.num_mov
    ns = csr["net_start"].astype(np.int64)
    pin_net = np.repeat(np.arange(len(ns) - 1), np.diff(ns))
    cs = csr["cell_start"].astype(np.int64)
    ext = np.concatenate([c.pos[:n].astype(np.float64), np.zeros((1, 2))])
    off = csr["off"].astype(np.float64)
    out = np.zeros((len(cells), 2))
    for k, cell in enumerate(cells):
        nets = np.unique(pin_net[csr["cell_pin"][cs[cell]:cs[cell + 1]]])
        for a in range(2):
            tot = 0.0
            for net in nets:
                pins = np.arange(ns[net], ns[net + 1])
                mine = csr["cell"][pins] == cell
                xs = ext[csr["cell"][pins], a] + off[pins, a]
                xp, xm = xs.copy(), xs.copy()
                xp[mine] += h
                xm[mine] -= h
                tot += (_net_wa64(xp, gamma) - _net_wa64(xm, gamma)) / (2 * h)
            out[k, a] = tot
    return out


def _potential64(rho, bx, by):
    """The potential of the map, phi = d E / d rho for the energy E = 1/2 sum rho phi (float64)."""
    nbx, nby = rho.shape
    Cx, _ = opo.dct_mats(nbx)
    Cy, _ = opo.dct_mats(nby)
    C = Cx @ rho @ Cy.T
    return Cx.T @ (opo.energy_weights(nbx, nby, bx, by) * C) @ Cy


def _density_fd(c, nodes, phi, smooth, ratio, bx, by, h=1e-3):
    """Central differences of the exact discrete energy for single-node moves: with the node's bins fixed over +-h,
    its charge changes linearly, the quadratic terms of E(x + h) and E(x - h) cancel, and the difference is
    sum_b (rho_b(x + h) - rho_b(x - h)) phi_b / 2h exactly."""
    nb = phi.shape[0]
    out = np.zeros((len(nodes), 2))
    flat = phi.reshape(-1)
    for k, i in enumerate(nodes):
        for a in range(2):
            q = np.repeat(c.pos[i:i + 1].astype(np.float64), 2, 0)
            q[0, a] += h
            q[1, a] -= h
            node, b, w = oden.node_bin_weights(q, np.repeat(smooth[i:i + 1], 2, 0), np.repeat(ratio[i:i + 1], 2, 0),
                                               bx, by, nb, nb, np.float64)
            out[k, a] = ((w * flat[b] * np.where(node == 0, 1.0, -1.0)).sum()) / (2 * h)
    return out


# ------------------------------------------------------------------------------------------------------- config B
@pytest.fixture(scope="module")
def cfg_b():
    return _config("B", (512, 1024))


def test_config_b_csr_and_grids(cfg_b):
    """The product's net / cell CSR equals the oracle's array for array; the grids are the stated ones."""
    c = cfg_b
    assert c.prob.num_bin == 512 and c.prob.force_bin == 1024
    assert c.data.num_mov == 200000 and c.data.num_nets == 210000
    for mine, theirs in (("cell32", "cell"), ("off", "off"), ("net_start", "net_start"), ("cell_start", "cell_start"),
                         ("cell_pin", "cell_pin")):
        assert np.array_equal(c.wl.opt[mine].cpu().numpy(), c.csr[theirs]), mine
    assert c.wl.num_opt_nets == c.csr["num_nets"]


def _clusters_match_the_oracle(c):
    """The device clustering (every round) against the oracle's numpy restatement, which the reference's own cluster
    ids pin: the same cluster of every cell and the same areas, bit for bit."""
    import time

    from autoplace_amd.cluster import AREA_SCALE, cluster_cells
    from oracle import cluster as ocl
    n = c.data.num_mov
    size = torch.tensor(c.data.size, dtype=torch.float32, device=DEV)
    area = (size.prod(1).double() * AREA_SCALE).round().long()
    torch.cuda.synchronize()
    t0 = time.time()
    cid, carea = cluster_cells(c.wl.all["cell"], c.wl.all["net_start"], area, 5.0, 0)
    torch.cuda.synchronize()
    t_dev = time.time() - t0
    ref_cid, ref_area = ocl.cluster_cells(c.data.pin_cell, c.data.pin_net, area.cpu().numpy(), 5.0, 0)
    print("config %s clustering: %d clusters of %d cells, device %.4f s" % (c.name, len(ref_area), n, t_dev),
          flush=True)
    assert np.array_equal(cid.cpu().numpy(), ref_cid)
    assert np.array_equal(carea.cpu().numpy(), ref_area)


def test_config_b_clusters_match_the_oracle(cfg_b):
    _clusters_match_the_oracle(cfg_b)


def test_config_b_dropin_kernels_vs_oracle(cfg_b):
    """gp_cuda.wa / density_map (force and overflow grid) / poisson_field / density_grad at 200k cells."""
    from autoplace_amd import gp_cuda
    c = cfg_b
    d, j = c.den, c.den.judge
    n = c.data.num_mov
    bx, by = float(d.bin_np[0]), float(d.bin_np[1])
    base = 4.0 * (bx + by)
    for gamma in (10 * base, 0.1 * base):            # the start value and the end of a run
        _, grad, hpwl, out, _ = owa.wa_forward(c.pos, c.csr, gamma, np.float32)
        _, grad64, _, out64, _ = owa.wa_forward(c.pos, c.csr, gamma, np.float64)
        wa_t, g_dev, hp_t = c.wl.forward(_t(c.pos), gamma)
        g_dev = g_dev.cpu().numpy()
        print("B WA gamma %.3g: grad vs fp32 %.2e (vs fp64 %.2e)" % (gamma, norm_err(g_dev, grad),
                                                                       norm_err(g_dev, grad64)), flush=True)
        assert norm_err(g_dev, grad) < 1e-5
        assert rel_err(float(wa_t), out64[:, 0].sum()) < 1e-5 and rel_err(float(hp_t), hpwl) < 1e-6
    smooth, ratio = d.smooth["size"].cpu().numpy(), d.smooth["ratio"].cpu().numpy()
    fixed = d.fixed_map.cpu().numpy()
    m0, m1 = oden.density_maps(c.pos, smooth, ratio, n, bx, by, d.nbx, d.nby)
    rho, over = oden.density_post(m0, m1, fixed, d.target, d.nbx, d.nby)
    rho_dev, over_dev = gp_cuda.density_map(_t(c.pos), d.smooth["size"], d.smooth["ratio"], n, bx, by, d.nbx, d.nby,
                                            d.fixed_map, d.target)
    assert np.array_equal(rho_dev.cpu().numpy(), oden.mirror(rho))
    assert np.array_equal(over_dev.cpu().numpy(), over)
    js, jr = j.smooth["size"].cpu().numpy(), j.smooth["ratio"].cpu().numpy()
    jbx, jby = float(j.bin_np[0]), float(j.bin_np[1])
    q0, q1 = oden.density_maps(c.pos[:n], js, jr, n, jbx, jby, j.nbx, j.nby)
    _, jover = oden.density_post(q0, q1, j.fixed_map.cpu().numpy(), j.target, j.nbx, j.nby)
    _, jover_dev = gp_cuda.density_map(_t(c.pos[:n]), j.smooth["size"], j.smooth["ratio"], n, jbx, jby, j.nbx,
                                       j.nby, j.fixed_map, j.target)
    assert np.array_equal(jover_dev.cpu().numpy(), jover)
    ref = opo.field_mirrored_fft(oden.mirror(rho.astype(np.float64)), opo.to_field(d.nbx, d.nby, bx, by))
    field = gp_cuda.poisson_field(rho_dev, d.plan)
    e = norm_err(field[:, :d.nbx, :d.nby].cpu().numpy(), ref[:, :d.nbx, :d.nby])
    print("B Poisson %s 1024 x 1024: field vs fp64 %.2e" % (d.plan.method, e), flush=True)
    assert e < 1e-5
    f32 = ref[:, :d.nbx, :d.nby].astype(np.float32)
    wl = np.random.default_rng(3).normal(size=(n, 2)).astype(np.float32)
    pins, area = d.pins.cpu().numpy(), d.area.cpu().numpy()
    full = np.zeros((2, 2 * d.nbx, 2 * d.nby), np.float32)
    full[:, :d.nbx, :d.nby] = f32
    for weight, precond in ((2e-3, 2e-3), (1.0, -1.0)):
        g_ref = oden.density_grad(c.pos, smooth, ratio, f32[0], f32[1], wl, pins, area, weight, precond, bx, by,
                                  d.nbx, d.nby)
        g = gp_cuda.density_grad(_t(c.pos), d.smooth["size"], d.smooth["ratio"], _t(full), _t(wl), d.pins, d.area,
                                 _t([weight]), _t([precond]), bx, by, d.nbx, d.nby).cpu().numpy()
        print("B gather (weight %g): vs oracle %.2e" % (weight, norm_err(g, g_ref)), flush=True)
        assert norm_err(g, g_ref) < 1e-5


def test_config_b_engine_objective_and_map(cfg_b):
    _check_engine_objective(cfg_b, [4.0 * 7.0, 2e-3, 2e-3, 0.0, 0.0])


OUR_SEEDS = 6


def _flow_against_reference(c, name, grids):
    """Our flow on placer seeds 0 .. 5 against the reference's own runs of the same design (its CPU path,
    golden_run<name>.json). Every run converges to overflow <= 0.07 with the reference's grids and filler count.

    A run's HPWL varies from one placer seed to another by about 1.5 % (200k cells) to 2 % (2M cells): a run is a
    chaotic process, and another seed -- or another summation order of the same sums -- ends in another local optimum
    of the same quality distribution. So the comparison is of distributions: every reference run lies within three
    of our seeds' standard deviations of our mean, and the means differ by less than 0.5 % or, when the seeds recorded
    are too few to resolve 0.5 %, by less than three standard errors of that difference (our spread for both sides)."""
    from autoplace_amd.problem import default_args
    from autoplace_amd.run_placement import place
    path = os.path.join(GOLDEN, "golden_run%s.json" % name)
    if not os.path.exists(path):
        pytest.skip("no reference run of config %s recorded" % name)
    refs = json.load(open(path))
    ref0 = refs[min(refs, key=int)]
    ours = []
    for seed in range(OUR_SEEDS):
        _, m = place(c.data, default_args(seed=seed, num_bin=grids[0], force_num_bin=grids[1]), DEV)
        print("%s flow seed %d: %d iterations, HPWL %.6e, exact overflow %.4f, GP %.3f s, stop %s"
              % (name, seed, m["iters"], m["hpwl_sites"], m["overflow"], m["gp_time_s"], m["stop_reason"]),
              flush=True)
        assert m["stop_reason"] == "converged" and m["gp_overflow"] <= 0.07
        assert m["num_fillers"] == ref0["num_fillers"] and m["force_num_bin"] == ref0["force_num_bin"]
        assert m["num_bin"] == ref0["num_bin"]
        ours.append(m["hpwl_sites"])
    ours = np.array(ours)
    theirs = np.array([refs[s]["hpwl_sites"] for s in sorted(refs, key=int)])
    spread = ours.std(ddof=1) / ours.mean()
    se = spread * math.sqrt(1.0 / len(ours) + 1.0 / len(theirs))
    ratio = float(ours.mean() / theirs.mean())
    print("%s flow: our mean HPWL %.6e over %d seeds (spread %.2f %%), reference %s over %d seed(s) "
          "(iterations %s, %.1f s of GP each on its CPU path); mean ratio %.4f, standard error %.2f %%"
          % (name, ours.mean(), len(ours), 100 * spread, np.round(theirs, 0).tolist(), len(theirs),
             [refs[s]["iters"] for s in sorted(refs, key=int)], float(np.mean([r["gp_time_s"] for r in refs.values()])),
             ratio, 100 * se), flush=True)
    for h in theirs:
        assert abs(h / ours.mean() - 1) < 3 * spread
    assert abs(ratio - 1) < max(0.005, 3 * se)


def test_config_b_flow_against_reference_run(cfg_b):
    """The whole two-level flow on the 200k-cell design at grids 512 / 1024 against the reference's own runs of the
    same design (its CPU path, golden_runB.json)."""
    _flow_against_reference(cfg_b, "B", (512, 1024))


def test_config_b_gradients_vs_finite_differences(cfg_b):
    """WA: the kernel's cell gradient against central differences (FP64, step 1e-3 site) of the WA of the cell's
    nets, on 1,000 random cells, both axes. Density: the field gather against central differences of the exact
    discrete energy 1/2 sum rho phi on 1,000 random nodes -- the gather is ePlace's approximation of that derivative
    (the field sampled over the node's bins), so direction and scale agree, not the rounding."""
    c = cfg_b
    d = c.den
    n = c.data.num_mov
    rng = np.random.default_rng(23)
    cells = rng.choice(n, 1000, replace=False)
    bx, by = float(d.bin_np[0]), float(d.bin_np[1])
    gamma = 4.0 * (bx + by)
    _, g_dev, _ = c.wl.forward(_t(c.pos), gamma)
    an = g_dev.cpu().numpy()[cells].astype(np.float64)
    fd = _wa_fd(c, cells, gamma)
    e = norm_err(an, fd)
    print("B WA gradient vs FP64 central differences on 1000 cells: %.2e" % e, flush=True)
    assert e < 1e-4
    smooth, ratio = d.smooth["size"].cpu().numpy(), d.smooth["ratio"].cpu().numpy()
    a0, a1 = oden.density_maps(c.pos, smooth, ratio, n, bx, by, d.nbx, d.nby, exact=True)
    rho, _ = oden.density_post(a0, a1, d.fixed_map.cpu().numpy().astype(np.float64), d.target, d.nbx, d.nby,
                               exact=True)
    phi = _potential64(rho, bx, by)
    nodes = rng.choice(len(c.pos), 1000, replace=False)
    g, _ = d.forward(_t(c.pos))
    gs = g.cpu().numpy()[nodes].astype(np.float64)
    fdd = _density_fd(c, nodes, phi, smooth, ratio, bx, by)
    corr = np.corrcoef(fdd.ravel(), gs.ravel())[0, 1]
    scale = float((fdd * gs).sum() / (gs * gs).sum())
    print("B density gather vs central differences of the exact energy on 1000 nodes: rel gap %.3f, correlation "
          "%.4f, scale %.3f" % (norm_err(gs, fdd), corr, scale), flush=True)
    assert corr > 0.9 and 0.8 < scale < 1.2


# ------------------------------------------------------------------------------------------------------- config C
@pytest.fixture(scope="module")
def cfg_c():
    return _config("C")


def test_config_c_engine_objective_and_map(cfg_c):
    """One engine objective at 2M cells (force grid 2048, overflow grid 512) against the oracle; the 2048 x 2048
    map bit for bit."""
    c = cfg_c
    assert c.prob.force_bin == 2048 and c.prob.num_bin == 512
    _check_engine_objective(c, [4.0 * 10.0, 1e-3, 1e-3, 0.0, 0.0])


def test_config_c_clusters_match_the_oracle(cfg_c):
    _clusters_match_the_oracle(cfg_c)


def test_config_c_free_fraction_on_the_device(cfg_c):
    for nb in (512, 1024, 2048):
        host = cfg_c.data.fixed_fraction(nb, nb)
        dev = cfg_c.data.fixed_fraction_on(DEV, nb, nb).cpu().numpy()
        assert np.abs(dev - host).max() < 1e-12, nb


def test_config_c_poisson_fft_and_gemm_2048(cfg_c):
    """Both Poisson forms at 2048 x 2048 on the design's own map against the FP64 mirrored-FFT oracle, and against
    each other."""
    from autoplace_amd import gp_cuda
    c = cfg_c
    d = c.den
    bx, by = float(d.bin_np[0]), float(d.bin_np[1])
    smooth, ratio = d.smooth["size"].cpu().numpy(), d.smooth["ratio"].cpu().numpy()
    m0, m1 = oden.density_maps(c.pos, smooth, ratio, c.data.num_mov, bx, by, d.nbx, d.nby)
    rho, _ = oden.density_post(m0, m1, d.fixed_map.cpu().numpy(), d.target, d.nbx, d.nby)
    ref = opo.field_mirrored_fft(oden.mirror(rho.astype(np.float64)), opo.to_field(d.nbx, d.nby, bx, by))
    ref = ref[:, :d.nbx, :d.nby]
    out = {}
    for method in ("fft", "gemm"):
        plan = gp_cuda.PoissonPlan(d.nbx, d.nby, bx, by, method)
        assert plan.method == method
        out[method] = plan.field(_t(rho)).cpu().numpy()
        print("C Poisson %s 2048 x 2048: vs fp64 %.2e" % (method, norm_err(out[method], ref)), flush=True)
        assert norm_err(out[method], ref) < 1e-5
    assert norm_err(out["gemm"], out["fft"]) < 2e-5


def test_config_c_flow_against_reference_run(cfg_c):
    """The whole two-level flow on the 2M-cell design against the reference's own runs of the same design (its CPU
    path, golden_runC.json)."""
    _flow_against_reference(cfg_c, "C", (0, 0))


# ------------------------------------------------------------------------------------------------------- config D
def test_config_d_64_blocks_batched_equals_sequential():
    """64 blocks, 2,000 .. 50,000 cells log-uniform, grids 64 .. 256: the batched run (grouped launches, one graph per
    run of passes, autotuned tile layouts) ends every block bit for bit where its own device loop ends."""
    from autoplace_amd import batch
    from autoplace_amd.gpu_loop import FIELDS
    blocks = batch.make_blocks(64, 2000, 50000, seed=0)
    cells = [b.num_mov for b in blocks]
    assert len(blocks) == 64 and min(cells) >= 2000 and max(cells) <= 50000
    states = [batch.start_state(b) for b in blocks]
    print("D: 64 start states, cells %d .. %d" % (min(cells), max(cells)), flush=True)
    seq = batch.run_sequential(states)
    bat = batch.run_batched(states, autotune=True)
    c_seq = np.array([r["completion_s"] for r in seq["blocks"]])
    c_bat = np.array([r["completion_s"] for r in bat["blocks"]])
    print("D: sequential %.3f s, batched %.3f s; completion p50 / p99: sequential %.3f / %.3f s, batched %.3f / %.3f s"
          % (seq["seconds"], bat["seconds"], np.percentile(c_seq, 50), np.percentile(c_seq, 99),
             np.percentile(c_bat, 50), np.percentile(c_bat, 99)), flush=True)
    assert all(r["stop"] == "converged" for r in seq["blocks"])
    for k, (ra, rb, pa, pb) in enumerate(zip(seq["blocks"], bat["blocks"], seq["positions"], bat["positions"])):
        diff = [(FIELDS[i], x, y) for i, (x, y) in enumerate(zip(ra["state"], rb["state"]))
                if x != y and FIELDS[i] not in ("accept", "is_best")]
        assert not diff, (k, diff)
        assert np.array_equal(pa, pb), k


# ------------------------------------------------------------------------------------- the reference's loop body
def test_reference_loop_body_over_the_dropins(small_design):
    """src/gpu_loop.py:58-64 restated over the six drop-in calls (nesterov_step, the objective's wa / density_map /
    Poisson / density_grad, control, select), 16 passes from one state, against the oracle's 16 passes: the control
    decisions agree and the points stay within FP32 drift."""
    from autoplace_amd import gp_cuda, gpu_loop
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import _scheduler, build_operators
    _, _, data = small_design
    args = default_args(num_bin=32, force_num_bin=64)          # an overflow grid of its own
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV, "fft")
    ps = _scheduler(den, data.num_mov, data.mov_area, float(data.cell_pins.sum()), float(prob.node_size.prod(1).sum()),
                    len(prob.node_size), args)
    ps.density_weight, ps.wa_coeff, ps.iter = 2e-3, 30.0, 40
    ps.hpwl.append(1e30)
    rng = np.random.default_rng(1)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos0 = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    S0, kp0, c0 = gpu_loop.control_state(ps, SimpleNamespace(alpha=0.05, a=1.0), 40, args)
    w, d, j = wl.opt, den, den.judge
    n = data.num_mov
    bx, by, jbx, jby = float(d.bin_np[0]), float(d.bin_np[1]), float(j.bin_np[0]), float(j.bin_np[1])
    objective = SimpleNamespace()

    def on_gpu(pos, kp):
        """Objective.on_gpu (calculator.py:37-43) of the reference's classes over the drop-ins: float32 totals
        (out.sum(0), wa_wirelength_hpwl.py:81; over.sum(), electronic_density_layer.py:110), the overflow on the
        overflow grid (:127-128), the field of the mirrored map (:133, poisson_field for torch.fft)."""
        grad_wl, out = gp_cuda.wa(pos, w["cell32"], w["off"], w["net_start"], w["cell_start"], w["cell_pin"], n,
                                  kp[0:1])
        total = out.sum(0)
        objective.wa, objective.hpwl = total[0], total[1]
        rho, _ = gp_cuda.density_map(pos, d.smooth["size"], d.smooth["ratio"], n, bx, by, d.nbx, d.nby, d.fixed_map,
                                     d.target)
        _, jover = gp_cuda.density_map(pos[:n].contiguous(), j.smooth["size"], j.smooth["ratio"], n, jbx, jby, j.nbx,
                                       j.nby, j.fixed_map, j.target)
        objective.overflow = jover.sum() * (j.bin_area / j.mov_area)
        field = gp_cuda.poisson_field(rho, d.plan)
        return gp_cuda.density_grad(pos, d.smooth["size"], d.smooth["ratio"], field, grad_wl, d.pins, d.area,
                                    kp[1:2], kp[2:3], bx, by, d.nbx, d.nby)

    objective.on_gpu = on_gpu
    pos = _t(pos0)
    kp = _t(kp0)
    g = objective.on_gpu(pos, kp)                    # the gradient at the start point
    g0 = g.cpu().numpy()
    S, c = _t(S0, torch.float64), _t(c0, torch.float64)
    u, best = pos.clone(), pos.clone()
    lo_t, hi_t = den.pos_lo, den.pos_hi
    for _ in range(16):
        u_next, v_next = gp_cuda.nesterov_step(u, pos, g, lo_t, hi_t, kp[3:4], kp[4:5])
        g_next = objective.on_gpu(v_next, kp)
        gp_cuda.control(S, kp, c, torch.dist(v_next, pos), torch.dist(g_next, g), objective.hpwl, objective.overflow,
                        objective.wa)
        gp_cuda.select(S, u, pos, g, u_next, v_next, g_next, best)
    torch.cuda.synchronize()
    lvl = opl.Level(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size,
                    data.num_mov, prob.node_pins, prob.pin_cell, prob.pin_off, prob.pin_net, prob.num_nets,
                    args.ignore_net_degree, {"nb": prob.num_bin, "fraction": prob.judge_fraction,
                                             "size": prob.node_size[:data.num_mov]})
    So, kpo = list(S0), np.array(kp0, np.float32)
    uo, vo, go = pos0.copy(), pos0.copy(), g0.copy()
    K = octl.K
    for _ in range(16):
        un, vn = octl.nesterov_step(uo, vo, go, lvl.pos_lo, lvl.pos_hi, kpo[3], kpo[4])
        gn, wa, hpwl, ovf = opl._evaluate_kp(lvl, vn, kpo)
        octl.control_pass(So, kpo, c0, octl.dist(vn, vo), octl.dist(gn, go), hpwl, ovf, wa)
        if So[K["accept"]]:
            uo, vo, go = un, vn, gn
    Sd = S.cpu().numpy()
    for name in ("iter", "loop_it", "evals", "settled", "stop", "accept"):
        assert Sd[K[name]] == So[K[name]], name
    for name in ("lam", "gam", "alpha", "hpwl_prev"):
        assert rel_err(Sd[K[name]], So[K[name]]) < 1e-4, (name, Sd[K[name]], So[K[name]])
    e = norm_err(pos.cpu().numpy(), vo)
    print("reference loop body over the drop-ins, 16 passes: v vs oracle %.2e, evals %d" % (e, Sd[K["evals"]]))
    assert e < 1e-4 and Sd[K["evals"]] == 16


def test_density_energy_of_a_cosine_mode_closed_form():
    """rho = cos(pi p (n + 1/2) / N) along x: its potential solves -lap phi = rho, phi = rho / w_p^2 (w_p = pi p /
    (N bx)), so E = 1/2 sum rho phi / bx = N M / (4 w_p^2 bx) in the field's units. The library's energy and the
    oracle's both give it."""
    from autoplace_amd import gp_cuda
    N, M, bx, by, p = 128, 64, 2.5, 1.5, 5
    x = np.arange(N)
    rho = np.repeat(np.cos(np.pi * p * (x + 0.5) / N)[:, None], M, 1)
    w = np.pi * p / (N * bx)
    expect = N * M / (4 * w * w * bx)
    e64 = opo.density_energy(rho, bx, by)
    plan = gp_cuda.PoissonPlan(N, M, bx, by, "gemm")
    e_dev = float(plan.energy(_t(rho)))
    print("cosine-mode energy: closed form %.10e, oracle %.10e, library %.10e" % (expect, e64, e_dev))
    assert rel_err(e64, expect) < 1e-12
    assert rel_err(e_dev, expect) < 1e-5
    assert math.isfinite(e_dev)
