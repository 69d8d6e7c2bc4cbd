"""Writes the golden fixtures in tests/golden/ from the reference implementation (AutoPlace, Python CPU path).

Run in a container that holds the reference at /root/reference (never on the GPU host; the fixtures travel, the
reference does not):
    python tests/golden/make_golden.py [--full-run]

Designs come from autoplace_amd.synth, written as LEF / DEF by autoplace_amd.lefdef and read by the reference's own
reader. What is recorded (inputs and outputs only, no reference code):
    golden_small.npz   1,500-cell design with 2 macros and 12 design pins: pin indexing, fixed fractions, grids,
                       fillers, start positions, WA value / gradient / HPWL, density map / overflow / gradient,
                       tension unit, cluster ids and areas, the coarse pins, the expansion
    golden_sched.npz   ParamScheduler.step / need_to_stop over a scripted input sequence (both --wa_rule values)
    golden_runA.json   --gpu -1 runs of the 10k-cell design, placer seeds 0-5: metrics (with --full-run)
    golden_runA.npz    the final cell positions of the seed-0 run (with --full-run)
    golden_runB.json   --gpu -1 run of the 200k-cell design (--num_bin 512 --force_num_bin 1024), placer seed 0
                       (with --run B; more placer seeds with --seeds 1,2,...)
    golden_runC.json   --gpu -1 run of the 2M-cell design (auto grids: 512 / 2048), placer seed 0 (with --run C)
"""
import argparse
import json
import os
import sys
import tempfile
from types import SimpleNamespace

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, REPO)

from autoplace_amd import lefdef, synth  # noqa: E402


def write_design(lef, d, folder, name):
    lp, dp = os.path.join(folder, name + ".lef"), os.path.join(folder, name + ".def")
    lefdef.write_lef(lef, lp)
    lefdef.write_def(d, dp)
    return lp, dp


def small_golden(tmp):
    import torch
    from src.cluster import cluster_cells, coarse_problem, expand
    from src.core import ElectronicDensityLayer, Wirelength
    from src.database import PlaceData
    from src.initializer import choose_num_bin, compute_fillers, init_positions
    from src.resolution import force_num_bin
    from utils.io_parser import parse_def, parse_lef
    import logging

    lef, d = synth.generate(1500, num_macros=2, macro_frac=0.1, seed=7, io_pins=12, name="small")
    lp, dp = write_design(lef, d, tmp, "small")
    data = PlaceData(parse_lef([lp]), parse_def(dp))
    out = {"pin_cell": data.pin_cell, "pin_off": data.pin_off, "pin_net": data.pin_net, "num_mov": data.num_mov,
           "num_nets": data.num_nets, "size": data.size, "cell_pins": data.cell_pins,
           "width": data.width, "height": data.height, "row_height": data.row_height}
    args = SimpleNamespace(num_bin=0, target_density=1.0, seed=0, cluster_ratio=5.0, ignore_net_degree=100,
                           wa_rule="accuracy")
    num_bin = choose_num_bin(data, 0)
    judge_fraction = data.fixed_fraction(num_bin, num_bin)
    free = 1.0 - float(judge_fraction.mean())
    force_bin = force_num_bin(data.num_mov, free, num_bin)
    fixed_fraction = data.fixed_fraction(force_bin, force_bin)
    log = logging.getLogger("golden")
    fillers = compute_fillers(data, fixed_fraction, data.width / force_bin * data.height / force_bin, args, log)
    node_size = np.concatenate([data.size, fillers])
    pos0 = init_positions(data, node_size, args, "cpu")
    out.update(num_bin=num_bin, force_bin=force_bin, judge_fraction=judge_fraction, fixed_fraction=fixed_fraction,
               fixed_fraction_48x40=data.fixed_fraction(48, 40), node_size=node_size, pos0=pos0.numpy(),
               target=args.target_density)

    wl = Wirelength(torch.tensor(data.pin_cell), torch.tensor(data.pin_off, dtype=torch.float32),
                    torch.tensor(data.pin_net), data.num_nets, data.num_mov, 100)
    node_pins = np.concatenate([data.cell_pins, np.zeros(len(node_size) - data.num_mov)])
    den = ElectronicDensityLayer(data.width, data.height, force_bin, force_bin, fixed_fraction, args.target_density,
                                 node_size, data.num_mov, "cpu", node_pins)
    gen = np.random.default_rng(11)
    pos_rand = np.concatenate([gen.uniform(2, data.width - 2, (len(node_size), 1)),
                               gen.uniform(5, data.height - 5, (len(node_size), 1))], 1).astype(np.float32)
    # inside the bounds every optimizer step keeps the nodes in (electronic_density_layer.py:54-55)
    pos_rand = torch.tensor(pos_rand).clamp_(min=den.pos_lo, max=den.pos_hi).numpy()
    out["pos_rand"] = pos_rand
    for tag, pos in (("p0", pos0), ("pr", torch.tensor(pos_rand))):
        for gam in (2.0, 40.0):
            wa, grad, hpwl = wl.forward(pos, gam)
            key = "wa_%s_g%d" % (tag, int(gam))
            out[key] = np.array([float(wa), float(hpwl)])
            out[key + "_grad"] = grad.numpy()
        out["hpwl_exact_" + tag] = float(wl.hpwl(pos))

    judge = ElectronicDensityLayer(data.width, data.height, num_bin, num_bin, judge_fraction, args.target_density,
                                   data.size, data.num_mov, "cpu", field=False)
    pr = torch.tensor(pos_rand)
    rho, ovf = den._density(pr, den.smooth, data.num_mov)
    out["rho_pr"] = rho.numpy()
    out["ovf_force_pr"] = float(ovf)
    _, grad_wl, _ = wl.forward(pr, 40.0)
    g, _ = den.forward(pr, grad_wl, 3.5e-3, 3.5e-3)
    out["dgrad_pr"] = g.numpy()
    g1, _ = den.forward(pr)
    out["dgrad_plain_pr"] = g1.numpy()
    out["ovf_judge_pr"] = float(judge.smoothed_overflow(pr))
    out["ovf_exact_pr"] = float(judge.exact_overflow(pr))
    den.judge = judge
    from src.calculator import Objective
    ps = SimpleNamespace(wa_coeff=40.0, density_weight=0.0)
    obj = Objective(wl, den, ps)
    out["tension_unit_pr"] = obj.tension_unit(pr)

    area = (torch.tensor(data.size, dtype=torch.float32).prod(1).double() * 1024).round().long()
    cid, carea = cluster_cells(wl.all["cell"], wl.all["net"], area, 5.0, 0)
    out["cluster_id"] = cid.numpy()
    out["cluster_area"] = carea.numpy()
    coarse = coarse_problem(data, node_size, wl, den, fixed_fraction, args)
    out["coarse_num"] = coarse.num
    out["coarse_size"] = coarse.size.numpy()
    out["coarse_nb"] = coarse.density.nbx
    out["coarse_total_pins"] = coarse.total_pins
    cpins = coarse.wirelength.pins()
    out["coarse_pin_cell"] = cpins["cell"].numpy()
    out["coarse_pin_off"] = cpins["off"].numpy()
    out["coarse_pin_net"] = cpins["net"].numpy()
    cpos = torch.cat([pr[:coarse.num], pr[data.num_mov:]])
    out["expand_pr"] = expand(cpos, coarse, node_size, 0).numpy()
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **{k: np.asarray(v) for k, v in out.items()})


def sched_golden():
    from src.param_scheduler import ParamScheduler

    out = {}
    for rule in ("accuracy", "overflow"):
        args = SimpleNamespace(wa_coeff=4.0, wa_rule=rule, start_tol=1e-3, min_iter=20, stop_overflow=0.07,
                               settle_iter=5, stall_iter=15)
        data = SimpleNamespace(bin=np.array([3.0, 2.5]), refine=2.0, total_pins=4000.0, total_area=90000.0,
                               pitch=7.0, num_nodes=2500)
        ps = ParamScheduler(data, args, None)
        ps.density_weight, ps.iter = 2e-4, 7
        ps.hpwl.append(1.0e5)
        gen = np.random.default_rng(5)
        T = 160
        ovf = np.concatenate([np.linspace(0.9, 0.05, 110), 0.05 + 0.01 * gen.random(50)])
        hpwl = 1.0e5 * (1 + 0.02 * np.sin(np.arange(T) / 3.0)) * (1 + np.arange(T) / 400.0)
        moved = 50.0 * np.exp(-np.arange(T) / 40.0) * (1 + 0.3 * gen.random(T))
        wa = hpwl * (1 - np.clip(0.2 * np.exp(-np.arange(T) / 30.0) + 0.001 * gen.random(T), 1e-7, None))
        rows = []
        for t in range(T):
            ps.step(float(hpwl[t]), float(ovf[t]), float(moved[t]), float(wa[t]))
            stop = ps.need_to_stop()
            rows.append([ps.density_weight, ps.wa_coeff, ps.mu, ps.motion, ps.weighted_weight, ps.settled,
                         ps.best_hpwl, ps.best_overflow, float(ps.is_best), ps.improved_iter, ps.iter,
                         {None: 0, "converged": 1, "stalled": 2, "diverged": 3, "handover": 5}[stop]])
            if stop:
                break
        out[rule + "_inputs"] = np.stack([hpwl, ovf, moved, wa], 1)
        out[rule + "_rows"] = np.array(rows)
    np.savez_compressed(os.path.join(HERE, "golden_sched.npz"), **out)


def full_run(tmp, seeds=(0, 1, 2, 3, 4, 5)):
    from src.run_placement import run_placement_main
    import logging

    lef, d = synth.generate_config("A", seed=0)
    lp, dp = write_design(lef, d, tmp, "synth_A")
    runs = {}
    for seed in seeds:
        runs[str(seed)] = _one_run(lp, dp, seed)
    with open(os.path.join(HERE, "golden_runA.json"), "w") as f:
        json.dump(runs, f, indent=1, sort_keys=True)


def config_run(tmp, name, seeds=(0,)):
    """Reference runs of a scale config (B at its stated grids, C at the automatic ones), one per placer seed, merged
    into the fixture's earlier seeds."""
    grids = {"B": (512, 1024), "C": (0, 0)}[name]
    lef, d = synth.generate_config(name, seed=0)
    lp, dp = write_design(lef, d, tmp, "synth_" + name)
    del lef, d
    path = os.path.join(HERE, "golden_run%s.json" % name)
    for seed in seeds:
        m = _one_run(lp, dp, seed, num_bin=grids[0], force_num_bin=grids[1], keep_pos=False)
        runs = json.load(open(path)) if os.path.exists(path) else {}
        runs[str(seed)] = m
        with open(path, "w") as f:
            json.dump(runs, f, indent=1, sort_keys=True)


def _one_run(lp, dp, seed, num_bin=0, force_num_bin=0, keep_pos=True):
    from src.run_placement import run_placement_main
    import logging

    args = SimpleNamespace(lef=[lp], def_file=dp, out_dir="", eval_def="", write_def=False, seed=seed, gpu=-1,
                           target_density=1.0, num_bin=num_bin, force_num_bin=force_num_bin, stop_overflow=0.07,
                           inner_iter=10000,
                           settle_iter=30, stall_iter=200, min_iter=100, ignore_net_degree=100, start_tol=1e-3,
                           density_weight=1e-3, wa_coeff=4.0, wa_rule="accuracy", cluster_ratio=5.0, lr=0.01,
                           log_freq=100)
    logger = logging.getLogger("golden_run")
    logger.addHandler(logging.StreamHandler(sys.stdout))
    logger.setLevel(logging.INFO)
    import src.run_placement as rp
    keep = {}
    orig = rp.evaluate_placement

    def spy(pos, wirelength, density, data):
        keep["pos"] = pos[:data.num_mov].detach().cpu().numpy().copy()
        return orig(pos, wirelength, density, data)

    rp.evaluate_placement = spy
    m = run_placement_main(args, logger)
    m = {k: v for k, v in m.items() if isinstance(v, (int, float, str, bool))}
    if seed == 0 and keep_pos:
        np.savez_compressed(os.path.join(HERE, "golden_runA.npz"), pos=keep["pos"].astype(np.float32))
    return m


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--full-run", action="store_true")
    p.add_argument("--run", choices=("B", "C"), help="only the reference runs of this scale config")
    p.add_argument("--seeds", default="0", help="placer seeds of --run, comma separated")
    a = p.parse_args()
    sys.path.insert(0, REF)
    import torch
    torch.set_num_threads(8)
    with tempfile.TemporaryDirectory() as tmp:
        if a.run:
            config_run(tmp, a.run, [int(x) for x in a.seeds.split(",")])
            return
        small_golden(tmp)
        sched_golden()
        if a.full_run:
            full_run(tmp)


if __name__ == "__main__":
    main()
