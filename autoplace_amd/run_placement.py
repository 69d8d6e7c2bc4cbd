"""The placement flow on the device: problem -> two-level global placement -> evaluation
(src/run_placement.py:34-200 of the reference; LEF / DEF reading and writing in lefdef.py).
"""
import math
import time
from types import SimpleNamespace

import numpy as np
import torch

from . import gpu_loop
from .calculator import Objective
from .cluster import ExpansionDraws, coarse_problem, expand
from .density import ElectronicDensityLayer
from .initializer import choose_num_bin
from .nesterov_optimizer import NesterovOptimizer, estimate_initial_learning_rate
from .param_scheduler import ParamScheduler
from .problem import build_problem
from .resolution import refine
from .wirelength import Wirelength


def _sync(device):
    torch.cuda.synchronize(device)


def _scheduler(density, num_mov, mov_area, total_pins, total_area, num_nodes, args):
    """pitch: the spacing of the cells once spread at the target density (run_placement.py:34-38)."""
    return ParamScheduler(SimpleNamespace(bin=density.bin_np, total_pins=total_pins, total_area=total_area,
                                          num_nodes=num_nodes, pitch=math.sqrt(mov_area / density.target / num_mov),
                                          refine=refine(density)), args)


def _place_level(pos, objective, ps, args, log, first_iter=0, opts=None, phase=0.0):
    """One level (run_placement.py:41-76): host loop through the wirelength-only start, then the device loop.
    -> namespace(reason, iters, start_iters, hpwl, overflow)."""
    opts = opts or {}
    log = gpu_loop.as_log(log)
    bounds = (objective.density.pos_lo, objective.density.pos_hi)
    lr, start = estimate_initial_learning_rate(pos, objective, args.lr, bounds, keep=True)
    optimizer = NesterovOptimizer(pos, objective, lr, bounds, stats=lambda: (objective.hpwl, objective.overflow,
                                                                               objective.wa), start=start)
    reason, iteration, start_iters = "max_iter", first_iter, 0
    hpwl = overflow = float("nan")
    engine = None
    if ps.density_weight == 0 and gpu_loop.usable(optimizer) and opts.get("device_start", True):
        # the wirelength-only start as engine passes; the level's engine then goes on with the density phase
        gpu_loop.mark(opts.get("timeline"), "start set-up", pos.device)
        engine, done, count = gpu_loop.run_start_on_gpu(
            optimizer, objective, ps, args, log, first_iter, use_graph=opts.get("graph", True),
            poisson=opts.get("poisson", "auto"), reorder=opts.get("reorder", True), timeline=opts.get("timeline"))
        first_iter += count
        iteration = first_iter - 1
        if count:
            hpwl, wa = ps.hpwl[-1], ps.wa[-1]
        if done == 2:
            engine.close()
            return SimpleNamespace(reason="diverged", iters=first_iter, start_iters=0, hpwl=hpwl, overflow=overflow)
        if done == 1:
            ps.density_weight, start_iters = args.density_weight * objective.tension_unit(pos), first_iter
            if log:
                log("wirelength-only start over after %d iterations (hpwl %.4E): density weight %.4E"
                    % (start_iters, hpwl, ps.density_weight))
        else:
            engine.close()
            engine = None
    for iteration in range(first_iter, args.inner_iter):
        if ps.density_weight > 0 and gpu_loop.usable(optimizer):
            gpu_loop.mark(opts.get("timeline"), "start phase", pos.device)
            reason, iters, hpwl, overflow = gpu_loop.run_level_on_gpu(
                optimizer, objective, ps, args, log, iteration, use_graph=opts.get("graph", True),
                poisson=opts.get("poisson", "auto"), reorder=opts.get("reorder", True), hook=opts.get("hook"),
                controller=opts.get("controller"), phase=phase, autotune=opts.get("autotune"),
                timeline=opts.get("timeline"), engine=engine)
            return SimpleNamespace(reason=reason, iters=iters, start_iters=start_iters, hpwl=hpwl, overflow=overflow)
        optimizer.step()
        if optimizer.diverged:
            reason = "diverged"
            break
        hpwl, overflow, wa = optimizer.fetched
        if log and iteration % args.log_freq == 0:
            log("iter: %d | hpwl: %.4E overflow: %.4f density_weight: %.4E wa_coeff: %.4E"
                % (iteration, hpwl, overflow, ps.density_weight, ps.wa_coeff))
        if ps.density_weight == 0:
            if ps.start_step(hpwl, wa):
                ps.density_weight, start_iters = args.density_weight * objective.tension_unit(pos), iteration + 1
                if log:
                    log("wirelength-only start over after %d iterations (hpwl %.4E): density weight %.4E"
                        % (start_iters, hpwl, ps.density_weight))
            continue
        raise RuntimeError("unreachable: the scheduled part runs on the device")
    return SimpleNamespace(reason=reason, iters=iteration + 1, start_iters=start_iters, hpwl=hpwl, overflow=overflow)


def global_placement_main(prob, wirelength, density, pos, args, log=None, opts=None):
    """Optimize pos in place (run_placement.py:79-117). -> dict of GP statistics (gp_time_s from the synced start)."""
    device, n = pos.device, prob.num_mov
    data, node_size, fixed_fraction = prob.data, prob.node_size, prob.fixed_fraction
    opts = opts or {}
    log = gpu_loop.as_log(log)
    _sync(device)
    start = time.time()
    ps = _scheduler(density, n, data.mov_area, prob.total_pins, prob.total_area, len(node_size), args)
    objective = Objective(wirelength, density, ps)
    stats, evals, first_iter, start_iters = {}, 0, 0, 0
    if args.cluster_ratio > 1:
        coarse = coarse_problem(data, node_size, wirelength, density, fixed_fraction, args,
                                poisson=opts.get("poisson", "auto"))
        _sync(device)
        gpu_loop.mark(opts.get("timeline"), "clusters", device)
        stats = {"clusters": coarse.num, "cluster_time_s": time.time() - start, "coarse_num_bin": coarse.density.nbx}
        if log:
            log("two levels: %d clusters of %d cells (%.2fs), %d x %d bins"
                % (coarse.num, n, stats["cluster_time_s"], coarse.density.nbx, coarse.density.nby))
        cargs = SimpleNamespace(**vars(args))
        cargs.settle_iter, cargs.min_iter = 0, 0
        cps = _scheduler(coarse.density, coarse.num, data.mov_area, coarse.total_pins, coarse.total_area,
                         coarse.num_nodes, cargs)
        cps.handover_gamma = ps.wa_coeff
        cobjective = Objective(coarse.wirelength, coarse.density, cps)
        cpos = torch.cat([pos[:coarse.num], pos[n:]])
        draws = ExpansionDraws(n, args.seed)             # drawn on the host beside the clusters' placement
        level = _place_level(cpos, cobjective, cps, cargs, log, 0, opts, phase=1.0)
        gpu_loop.mark(opts.get("timeline"), "coarse level", device)
        pos.copy_(expand(cpos, coarse, node_size, args.seed, draws)).clamp_(min=density.pos_lo, max=density.pos_hi)
        weight = cps.density_weight * objective.tension_unit(pos) / cobjective.tension_unit(cpos)
        overflow = density.forward(pos)[1].item() if args.wa_rule == "overflow" else None
        ps.resume(cps, weight, wirelength.forward(pos, ps.wa_coeff)[2].item(), overflow)
        gpu_loop.mark(opts.get("timeline"), "expansion", device)
        if log:
            log("clusters placed after %d iterations (stop: %s, hpwl %.4E, overflow %.4f): density weight %.4E"
                % (level.iters, level.reason, level.hpwl, level.overflow, weight))
        evals, first_iter, start_iters = cobjective.evals, level.iters, level.start_iters
        stats["coarse_iters"] = level.iters
        stats["coarse_spills"] = getattr(cobjective, "bucket_spills", None)
        stats["cluster_ids"] = coarse.cid
        del coarse, cobjective, cps, cpos
    level = _place_level(pos, objective, ps, args, log, first_iter, opts)
    _sync(device)
    gp_time, iters = time.time() - start, level.iters
    if opts.get("timeline") is not None:
        opts["timeline"].insert(0, ("start", start))
        opts["timeline"].append(("cell level", start + gp_time))
    if log:
        log("GP stop (%s): #iters %d hpwl: %.6E overflow: %.4f GP time: %.4fs (%d gradient calls)"
            % (level.reason, iters, level.hpwl, level.overflow, gp_time, evals + objective.evals))
    stats.update({"stop_reason": level.reason, "iters": iters, "start_iters": start_iters or level.start_iters,
                  "gp_time_s": gp_time, "per_iter_s": gp_time / max(iters, 1), "grad_calls": evals + objective.evals,
                  "gp_hpwl": level.hpwl, "gp_overflow": level.overflow,
                  "spills": getattr(objective, "bucket_spills", None)})
    return stats


def build_operators(prob, args, device, poisson="auto"):
    """Wirelength, density (force grid) with its overflow grid, as run_placement.py:163-170."""
    data = prob.data
    wirelength = Wirelength(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov,
                            args.ignore_net_degree, device)
    density = ElectronicDensityLayer(data.width, data.height, prob.force_bin, prob.force_bin, prob.fixed_fraction,
                                     args.target_density, prob.node_size, data.num_mov, device, prob.node_pins,
                                     poisson=poisson)
    if prob.force_bin != prob.num_bin:
        density.judge = ElectronicDensityLayer(data.width, data.height, prob.num_bin, prob.num_bin,
                                               prob.judge_fraction, args.target_density, data.size, data.num_mov,
                                               device, field=False)
    return wirelength, density


def evaluate_placement(pos, wirelength, density, data):
    """Exact HPWL (every net) and exact overflow (real sizes, overflow grid) (evaluator.py:4-8)."""
    hpwl = wirelength.hpwl(pos).item()
    return {"hpwl_sites": hpwl, "hpwl_um": hpwl * data.unit_dbu / float(data.d.dbu),
            "overflow": density.exact_overflow(pos[:data.num_mov]).item()}


def place(data, args, device="cuda", log=None, opts=None):
    """The whole run on a PlaceData: problem, two-level GP, evaluation. -> (positions [nodes, 2] on the device,
    metrics dict)."""
    device = torch.device(device)
    if device.type == "cuda" and device.index is not None:
        torch.cuda.set_device(device)
    opts = opts or {}
    log = gpu_loop.as_log(log)
    prob = build_problem(data, args, log)
    wirelength, density = build_operators(prob, args, device, opts.get("poisson", "auto"))
    metrics = {"design": data.name, "num_movable": data.num_mov, "num_nets": data.num_nets,
               "num_pins": data.num_pins, "num_fillers": prob.num_fillers, "num_bin": prob.num_bin,
               "force_num_bin": prob.force_bin, "target_density": args.target_density, "seed": args.seed,
               "device": str(device), "poisson": density.plan.method}
    pos = torch.tensor(prob.pos0, dtype=torch.float32, device=device)
    metrics.update(global_placement_main(prob, wirelength, density, pos, args, log, opts))
    if not bool(torch.isfinite(pos[:data.num_mov]).all()):
        raise RuntimeError("the placement is not finite (diverged before any usable point)")
    metrics.update(evaluate_placement(pos, wirelength, density, data))
    return pos, metrics


def evaluate_def(data, placed_def, args, device="cuda"):
    """Only evaluate: the cells where another DEF of the same netlist places them (run_placement.py:145-152,175-176):
    exact HPWL over every net (FP64 device sum) and exact overflow (real sizes, no fillers) on the overflow grid.
    -> (positions [num_mov, 2] on the device, metrics dict)."""
    device = torch.device(device)
    num_bin = choose_num_bin(data, args.num_bin)
    wirelength = Wirelength(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov,
                            args.ignore_net_degree, device)
    density = ElectronicDensityLayer(data.width, data.height, num_bin, num_bin, data.fixed_fraction(num_bin, num_bin),
                                     args.target_density, data.size, data.num_mov, device, field=False)
    pos = torch.tensor(data.placed_pos(placed_def), dtype=torch.float32, device=device)
    metrics = {"design": data.name, "num_movable": data.num_mov, "num_nets": data.num_nets,
               "num_pins": data.num_pins, "num_fillers": 0, "num_bin": num_bin, "force_num_bin": num_bin,
               "target_density": args.target_density, "device": str(device), "evaluated": placed_def.path}
    metrics.update(evaluate_placement(pos, wirelength, density, data))
    return pos, metrics


def run_placement_main(args, log=None, opts=None):
    """The run from files (run_placement.py:120-198): read the LEF(s) and the DEF, place -- or, with args.eval_def,
    only evaluate that placed DEF --, and with args.out_dir write gp.def (args.write_def) and metrics.json there.

    args: the reference's options (main.py:19-47): lef (list), def_file, out_dir, eval_def, write_def, gpu and the
    placement options of problem.default_args. The placement runs on the HIP device args.gpu; there is no CPU path
    (gpu < 0 is an error). -> metrics dict."""
    import json
    import os

    from . import lefdef
    from .database import PlaceData

    log = gpu_loop.as_log(log)
    if getattr(args, "gpu", 0) < 0:
        raise RuntimeError("this placer runs on a HIP device: --gpu %d (the reference's CPU path) is not part of it"
                           % args.gpu)
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device: the placer's kernels need one")
    start = time.time()
    device = torch.device("cuda:%d" % getattr(args, "gpu", 0))
    torch.cuda.set_device(device)
    design = lefdef.parse_def(args.def_file)
    data = PlaceData(lefdef.parse_lef(args.lef), design)
    parse_time = time.time() - start
    if log:
        log("design %s: %d movable, %d fixed, %d nets, %d pins, %d rows, %d blockages (read in %.2fs)"
            % (data.name, data.num_mov, int(data.fixed.sum()), data.num_nets, data.num_pins, len(design.rows),
               len(design.blockages), parse_time))
        if design.soft_blockages:
            log("warning: %d SOFT / PARTIAL placement blockages are not used" % design.soft_blockages)
        if data.unplaced_io_pins:
            log("warning: %d connections to design pins without a location are not used" % data.unplaced_io_pins)
    evaluate_only = bool(getattr(args, "eval_def", ""))
    if evaluate_only:
        pos, metrics = evaluate_def(data, lefdef.parse_def(args.eval_def), args, device)
    else:
        pos, metrics = place(data, args, device, log, opts)
        if metrics["stop_reason"] != "converged" and log:
            log("warning: global placement ended '%s', not converged (stop overflow %.3f)"
                % (metrics["stop_reason"], args.stop_overflow))
    metrics["parse_time_s"] = parse_time
    metrics.pop("cluster_ids", None)
    if log:
        log("exact HPWL: %.6E sites = %.1f um, exact overflow: %.4f"
            % (metrics["hpwl_sites"], metrics["hpwl_um"], metrics["overflow"]))
    out_dir = getattr(args, "out_dir", "")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        if not evaluate_only and getattr(args, "write_def", True):
            t = time.time()
            lefdef.write_placed_def(design, os.path.join(out_dir, "gp.def"), data.mov_index,
                                    data.to_def_xy(pos[:data.num_mov].cpu().numpy()))
            metrics["write_time_s"] = time.time() - t
        metrics["total_time_s"] = time.time() - start
        with open(os.path.join(out_dir, "metrics.json"), "w") as f:
            json.dump({k: v for k, v in metrics.items() if isinstance(v, (int, float, str, bool, list))}, f, indent=2)
    return metrics
