"""The oracle and the problem preparation against vectors written by the reference implementation
(tests/golden/make_golden.py): pin indexing, grids, fillers, start, WA, density, clustering, scheduler."""
import os
import tempfile
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import GOLDEN, norm_err, rel_err
from oracle import cluster as ocl
from oracle import control as octl
from oracle import density as oden
from oracle import poisson as opo
from oracle import wa as owa


def test_pin_indexing_bit_exact(golden_small, small_design):
    _, _, data = small_design
    g = golden_small
    assert data.num_mov == int(g["num_mov"]) and data.num_nets == int(g["num_nets"])
    assert np.array_equal(data.pin_cell, g["pin_cell"])
    assert np.array_equal(data.pin_net, g["pin_net"])
    assert np.array_equal(data.pin_off, g["pin_off"])          # float64, bit for bit
    assert np.array_equal(data.size, g["size"])
    assert np.array_equal(data.cell_pins, g["cell_pins"])


def test_lefdef_round_trip_bit_exact(golden_small, small_design):
    from autoplace_amd import lefdef
    from autoplace_amd.database import PlaceData
    lef, d, _ = small_design
    with tempfile.TemporaryDirectory() as tmp:
        lp, dp = os.path.join(tmp, "s.lef"), os.path.join(tmp, "s.def")
        lefdef.write_lef(lef, lp)
        lefdef.write_def(d, dp)
        data = PlaceData(lefdef.parse_lef([lp]), lefdef.parse_def(dp))
    assert np.array_equal(data.pin_cell, golden_small["pin_cell"])
    assert np.array_equal(data.pin_net, golden_small["pin_net"])
    assert np.array_equal(data.pin_off, golden_small["pin_off"])


def test_grids_fractions_fillers_start(golden_small, small_design):
    from autoplace_amd.problem import build_problem, default_args
    _, _, data = small_design
    g = golden_small
    p = build_problem(data, default_args())
    assert p.num_bin == int(g["num_bin"]) and p.force_bin == int(g["force_bin"])
    assert np.array_equal(p.judge_fraction, g["judge_fraction"])
    assert np.array_equal(p.fixed_fraction, g["fixed_fraction"])
    assert np.array_equal(data.fixed_fraction(48, 40), g["fixed_fraction_48x40"])
    assert np.array_equal(p.node_size, g["node_size"])
    assert np.array_equal(p.pos0, g["pos0"])
    assert p.target == float(g["target"])


@pytest.mark.parametrize("tag", ["p0", "pr"])
@pytest.mark.parametrize("gam", [2, 40])
def test_wa_against_reference(golden_small, tag, gam):
    g = golden_small
    n = int(g["num_mov"])
    csr = owa.build_csr(g["pin_cell"], g["pin_off"], g["pin_net"], int(g["num_nets"]), n, 100)
    pos = g["pos0"] if tag == "p0" else g["pos_rand"]
    key = "wa_%s_g%d" % (tag, gam)
    wa, grad, hpwl, _, _ = owa.wa_forward(pos, csr, float(gam), np.float32)
    assert rel_err(wa, g[key][0]) < 2e-6
    assert rel_err(hpwl, g[key][1]) < 2e-6
    assert norm_err(grad, g[key + "_grad"]) < 1e-5
    # the float64 form is within float32 rounding of both
    wa64, grad64, hpwl64, _, _ = owa.wa_forward(pos, csr, float(gam), np.float64)
    assert norm_err(grad, grad64) < 1e-5
    assert abs(owa.hpwl_exact(pos, g["pin_cell"], g["pin_off"], g["pin_net"], int(g["num_nets"]), n)
               - float(g["hpwl_exact_" + tag])) / float(g["hpwl_exact_" + tag]) < 2e-6


def _layer(g):
    nb = int(g["force_bin"])
    bx, by = float(g["width"]) / nb, float(g["height"]) / nb
    smooth, ratio = oden.smooth_nodes(g["node_size"], bx, by)
    return nb, bx, by, smooth, ratio


def test_density_map_against_reference(golden_small):
    g = golden_small
    nb, bx, by, smooth, ratio = _layer(g)
    n = int(g["num_mov"])
    fixed = (g["fixed_fraction"] * float(g["target"])).astype(np.float32)
    m0, m1 = oden.density_maps(g["pos_rand"], smooth, ratio, n, bx, by, nb, nb)
    rho, over = oden.density_post(m0, m1, fixed, float(g["target"]), nb, nb)
    assert rel_err(oden.mirror(rho), g["rho_pr"]) < 1e-5
    mov_area = float(g["node_size"][:n].prod(1).sum())
    assert rel_err(oden.overflow(over, bx * by, mov_area), float(g["ovf_force_pr"])) < 1e-5


def test_judge_overflows_against_reference(golden_small):
    g = golden_small
    nbj = int(g["num_bin"])
    n = int(g["num_mov"])
    bx, by = float(g["width"]) / nbj, float(g["height"]) / nbj
    fixed = (g["judge_fraction"] * float(g["target"])).astype(np.float32)
    mov_area = float(g["size"].prod(1).sum())
    pos = g["pos_rand"][:n]
    sm, ra = oden.smooth_nodes(g["size"], bx, by)
    m0, m1 = oden.density_maps(pos, sm, ra, n, bx, by, nbj, nbj)
    _, over = oden.density_post(m0, m1, fixed, float(g["target"]), nbj, nbj)
    assert rel_err(oden.overflow(over, bx * by, mov_area), float(g["ovf_judge_pr"])) < 1e-5
    real = np.asarray(g["size"], np.float32)
    m0, m1 = oden.density_maps(pos, real, np.ones(n, np.float32), n, bx, by, nbj, nbj)
    _, over = oden.density_post(m0, m1, fixed, float(g["target"]), nbj, nbj)
    assert rel_err(oden.overflow(over, bx * by, mov_area), float(g["ovf_exact_pr"])) < 1e-5


@pytest.mark.parametrize("weighted", [True, False])
def test_density_gradient_against_reference(golden_small, weighted):
    g = golden_small
    nb, bx, by, smooth, ratio = _layer(g)
    n = int(g["num_mov"])
    fixed = (g["fixed_fraction"] * float(g["target"])).astype(np.float32)
    pos = g["pos_rand"]
    m0, m1 = oden.density_maps(pos, smooth, ratio, n, bx, by, nb, nb)
    rho, _ = oden.density_post(m0, m1, fixed, float(g["target"]), nb, nb)
    tf = opo.to_field(nb, nb, bx, by, np.float32)
    field = opo.field_mirrored_fft(oden.mirror(rho), tf)[:, :nb, :nb].astype(np.float32)
    pins = np.concatenate([g["cell_pins"], np.zeros(len(pos) - n)]).astype(np.float32)
    area = g["node_size"].prod(1).astype(np.float32)
    if weighted:
        csr = owa.build_csr(g["pin_cell"], g["pin_off"], g["pin_net"], int(g["num_nets"]), n, 100)
        wl = owa.wa_forward(pos, csr, 40.0)[1]
        out = oden.density_grad(pos, smooth, ratio, field[0], field[1], wl, pins, area, 3.5e-3, 3.5e-3, bx, by, nb, nb)
        ref = g["dgrad_pr"]
    else:
        out = oden.density_grad(pos, smooth, ratio, field[0], field[1], np.zeros((0, 2)), pins, area, 1.0, -1.0,
                                bx, by, nb, nb)
        ref = g["dgrad_plain_pr"]
    assert norm_err(out, ref) < 1e-5


def test_cluster_ids_bit_exact(golden_small):
    g = golden_small
    n = int(g["num_mov"])
    area = np.round(g["size"].astype(np.float32).prod(1).astype(np.float64) * 1024).astype(np.int64)
    cid, carea = ocl.cluster_cells(g["pin_cell"], g["pin_net"], area, 5.0, 0)
    assert np.array_equal(cid, g["cluster_id"])
    assert np.array_equal(carea, g["cluster_area"])
    num = len(carea)
    assert num == int(g["coarse_num"])
    csize = ocl.cluster_sizes(cid, carea, num, g["size"])
    assert np.array_equal(csize, g["coarse_size"])
    owner, off, net = ocl.coarse_pins(cid, num, g["pin_cell"], g["pin_off"], g["pin_net"])
    assert np.array_equal(owner, g["coarse_pin_cell"])
    assert np.array_equal(net, g["coarse_pin_net"])
    assert np.array_equal(off, g["coarse_pin_off"])
    assert float(np.bincount(owner, minlength=num + 1)[:num].sum()) == float(g["coarse_total_pins"])
    cpos = np.concatenate([g["pos_rand"][:num], g["pos_rand"][n:]])
    assert np.array_equal(ocl.expand(cpos, cid, csize, num, g["size"], 0), g["expand_pr"])


@pytest.mark.parametrize("rule", ["accuracy", "overflow"])
def test_scheduler_sequence_bit_exact(rule):
    """control_pass's scheduler part against ParamScheduler.step / need_to_stop, FP64 bit for bit."""
    gs = np.load(os.path.join(GOLDEN, "golden_sched.npz"), allow_pickle=False)
    inputs, rows = gs[rule + "_inputs"], gs[rule + "_rows"]
    K = octl.K
    base_gamma = 4.0 * (3.0 + 2.5)
    S = [0.0] * len(octl.FIELDS)
    S[K["lam"]], S[K["gam"]], S[K["alpha"]], S[K["a"]] = 2e-4, 10 * base_gamma, 1.0, 1.0
    S[K["iter"]], S[K["hpwl_prev"]], S[K["mu"]] = 7, 1.0e5, 1.0
    S[K["best_hpwl"]] = S[K["best_ovf"]] = float("inf")
    kp = np.zeros(5, np.float32)
    c = [base_gamma, 7.0 * np.sqrt(2500), 4000.0, 90000.0, 7.0, 20, 0.07, 5, 15, 10 ** 6, 2.0,
         float(rule == "accuracy"), 0.0]
    for t, row in enumerate(rows):
        hpwl, ovf, moved, wa = (float(v) for v in inputs[t])
        octl.control_pass(S, kp, c, moved, 1.0, hpwl, ovf, wa)
        got = [S[K["lam"]], S[K["gam"]], S[K["mu"]], S[K["motion"]], S[K["ww"]], S[K["settled"]], S[K["best_hpwl"]],
               S[K["best_ovf"]], S[K["is_best"]], S[K["improved"]], S[K["iter"]], S[K["stop"]]]
        assert got == [float(v) for v in row], (t, got, list(row))
    assert rows[-1][-1] != 0                     # the sequence reaches a stop
