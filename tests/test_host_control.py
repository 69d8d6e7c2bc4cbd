"""Host-side control (the product's ParamScheduler, resolution rules, problem preparation) against the reference's
golden scheduler sequence and closed forms."""
import math
import os
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.mark.parametrize("rule", ["accuracy", "overflow"])
def test_param_scheduler_sequence(rule):
    from autoplace_amd.param_scheduler import ParamScheduler
    gs = np.load(os.path.join(GOLDEN, "golden_sched.npz"), allow_pickle=False)
    inputs, rows = gs[rule + "_inputs"], gs[rule + "_rows"]
    args = SimpleNamespace(wa_coeff=4.0, wa_rule=rule, start_tol=1e-3, min_iter=20, stop_overflow=0.07,
                           settle_iter=5, stall_iter=15)
    data = SimpleNamespace(bin=np.array([3.0, 2.5]), refine=2.0, total_pins=4000.0, total_area=90000.0, pitch=7.0,
                           num_nodes=2500)
    ps = ParamScheduler(data, args)
    ps.density_weight, ps.iter = 2e-4, 7
    ps.hpwl.append(1.0e5)
    code = {None: 0, "converged": 1, "stalled": 2, "diverged": 3, "handover": 5}
    for t, row in enumerate(rows):
        ps.step(*(float(v) for v in inputs[t][[0, 1, 2, 3]]))
        stop = ps.need_to_stop()
        got = [ps.density_weight, ps.wa_coeff, ps.mu, ps.motion, ps.weighted_weight, ps.settled, ps.best_hpwl,
               ps.best_overflow, float(ps.is_best), ps.improved_iter, ps.iter, code[stop]]
        assert got == [float(v) for v in row], t


def test_start_phase_rule():
    from autoplace_amd.param_scheduler import ParamScheduler
    args = SimpleNamespace(wa_coeff=4.0, wa_rule="accuracy", start_tol=1e-3)
    ps = ParamScheduler(SimpleNamespace(bin=(1.0, 1.0), refine=1.0, total_pins=1.0, total_area=1.0, pitch=1.0,
                                        num_nodes=1), args)
    wa = [1000.0 * 0.99 ** k for k in range(20)] + [990.0 * 0.99 ** 19] * 10
    ends = [ps.start_step(2 * w, w) for w in wa]
    # the window needs 5 earlier values; the fall over 5 steps is 4.9 % > 0.5 % while the WA falls 1 % a step
    assert not any(ends[:20])
    assert ends[-1]


def test_force_grid_and_smoothing_closed_forms():
    from autoplace_amd import resolution
    assert resolution.force_num_bin(10000, 1.0, 64) == 128
    assert resolution.force_num_bin(2000, 0.9, 64) == 64
    assert resolution.force_num_bin(2_860_000, 0.85, 512) == 2048
    # gamma moves by at most mu per iteration and stays in [0.01, 100] x unit
    g = resolution.smoothing(10.0, 100.0, 99.0, 0.5, 2.0, 4.0, 1.05)
    assert abs(g - 10.0 * 1.05) < 1e-12
    assert resolution.smoothing(1e9, 100.0, 50.0, 0.5, 1.0, 4.0, 1.05) == 400.0
    assert resolution.growth(0.0, 1.0, 1.0, True) == 1.05 * 0.98
    assert resolution.growth(1e9, 1.0, 1.0, False) == 1.05 * 0.95
