"""The PyTorch restatement of the reference's CPU pass (oracle/torch_cpu.py, the bench's CPU baseline) against the
numpy oracle, which the reference's own outputs pin (test_oracle_golden.py): the objective's gradient, WA, HPWL and
overflow with and without an overflow grid of its own, and the Nesterov step."""
import numpy as np
import pytest
import torch

from conftest import norm_err, rel_err
from oracle import control as octl
from oracle import place as opl
from oracle import torch_cpu


@pytest.mark.parametrize("grids", [(0, 0), (32, 64)])
def test_torch_cpu_objective_matches_the_oracle(small_design, golden_small, grids):
    from autoplace_amd.problem import build_problem, default_args
    _, _, data = small_design
    args = default_args(num_bin=grids[0], force_num_bin=grids[1])
    prob = build_problem(data, args)
    n = data.num_mov
    judge = None if prob.force_bin == prob.num_bin else {"nb": prob.num_bin, "fraction": prob.judge_fraction,
                                                          "size": prob.node_size[:n]}
    lvl = opl.Level(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size, n,
                    prob.node_pins, prob.pin_cell, prob.pin_off, prob.pin_net, prob.num_nets, 100, judge)
    tl = torch_cpu.Level(prob)
    rng = np.random.default_rng(4)
    pos = (lvl.pos_lo + rng.random(lvl.pos_lo.shape) * (lvl.pos_hi - lvl.pos_lo)).astype(np.float32)
    kp = np.array([12.0, 3e-3, 3e-3, 0.1, 0.2], np.float32)
    g_ref, wa, hpwl, ovf = opl._evaluate_kp(lvl, pos, kp)
    with torch.no_grad():
        g, twa, thpwl, tovf = tl.objective(torch.tensor(pos), kp)
    print("torch CPU pass vs oracle (grids %s): grad %.2e, wa %.2e, hpwl %.2e, overflow %.2e"
          % (grids, norm_err(g.numpy(), g_ref), rel_err(float(twa), wa), rel_err(float(thpwl), hpwl),
             rel_err(float(tovf), ovf)))
    # float32 sums in another order, float32 FFT
    assert norm_err(g.numpy(), g_ref) < 1e-4
    assert rel_err(float(twa), wa) < 1e-5 and rel_err(float(thpwl), hpwl) < 1e-5 and rel_err(float(tovf), ovf) < 1e-4
    assert np.allclose(tl.force.lo.numpy(), lvl.pos_lo) and np.allclose(tl.force.hi.numpy(), lvl.pos_hi)
    u, v, gg = (torch.tensor(x) for x in (pos, pos[::-1].copy(), g_ref))
    un, vn = torch_cpu.nesterov_step(u, v, gg, tl.force.lo, tl.force.hi, 0.37, 0.61)
    uo, vo = octl.nesterov_step(pos, pos[::-1].copy(), g_ref, lvl.pos_lo, lvl.pos_hi, np.float32(0.37),
                                np.float32(0.61))
    assert rel_err(un.numpy(), uo) < 1e-6 and rel_err(vn.numpy(), vo) < 1e-6
