"""Problem preparation of a run: grids, fixed fractions, fillers, start positions (run_placement.py:120-178).

`build_problem(data, args)` returns everything the placement levels need as host arrays; it is deterministic
(numpy and torch's seeded CPU generator) and runs once before the GP timer starts.
"""
from types import SimpleNamespace

import numpy as np

from .initializer import choose_num_bin, compute_fillers, init_positions
from .resolution import force_num_bin


def default_args(**kw):
    """The reference's option defaults (main.py:19-47) as a namespace; keywords override."""
    a = SimpleNamespace(lef=[], def_file="", out_dir="", eval_def="", write_def=True, gpu=0,
                        seed=0, target_density=1.0, num_bin=0, force_num_bin=0, stop_overflow=0.07, inner_iter=10000,
                        settle_iter=30, stall_iter=200, min_iter=100, ignore_net_degree=100, start_tol=1e-3,
                        density_weight=1e-3, wa_coeff=4.0, wa_rule="accuracy", cluster_ratio=5.0, lr=0.01,
                        log_freq=100)
    for k, v in kw.items():
        if not hasattr(a, k):
            raise KeyError("unknown option %s" % k)
        setattr(a, k, v)
    return a


def build_problem(data, args, log=None):
    """-> namespace(num_bin, force_bin, judge_fraction, fixed_fraction, node_size, node_pins, target, pos0, ...).
    May raise args.target_density to the utilization (initializer.py:27-30)."""
    num_bin = choose_num_bin(data, args.num_bin)
    judge_fraction = data.fixed_fraction(num_bin, num_bin)
    free = 1.0 - float(judge_fraction.mean())
    force_bin = args.force_num_bin or force_num_bin(data.num_mov, free, num_bin)
    fixed_fraction = judge_fraction if force_bin == num_bin else data.fixed_fraction(force_bin, force_bin)
    bin_area = data.width / force_bin * data.height / force_bin
    fillers = compute_fillers(data, fixed_fraction, bin_area, args, log)
    node_size = np.concatenate([data.size, fillers])
    node_pins = np.concatenate([data.cell_pins, np.zeros(len(node_size) - data.num_mov)])
    pos0 = init_positions(data, node_size, args.seed).numpy()
    # the scheduler's totals (node areas summed in float64 as one contiguous array)
    total_area = float((node_size[:, 0] * node_size[:, 1]).sum())
    return SimpleNamespace(num_bin=num_bin, force_bin=force_bin, judge_fraction=judge_fraction,
                           fixed_fraction=fixed_fraction, node_size=node_size, node_pins=node_pins,
                           target=args.target_density, pos0=pos0, num_mov=data.num_mov, num_fillers=len(fillers),
                           width=data.width, height=data.height, mov_area=data.mov_area, cell_pins=data.cell_pins,
                           pin_cell=data.pin_cell, pin_off=data.pin_off, pin_net=data.pin_net, num_nets=data.num_nets,
                           fraction_of=lambda nb: data.fixed_fraction(nb, nb), data=data,
                           total_area=total_area, total_pins=float(data.cell_pins.sum()))
