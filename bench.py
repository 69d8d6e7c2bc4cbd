"""Benchmark: gradient evaluations per second of the GP iteration on the 2M-cell synthetic design.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C] [--full] [--no-cpu]

Workload: the 2,000,000-cell design (2,100,000 nets, ~8.2M pins, 16 fixed macros on 15 % of the die, seed 0),
force grid 2048 x 2048, overflow grid 512 x 512, with the density term on (cells spread uniformly inside their
bounds, density weight from the tension unit, gamma at its start value). A step is one pass of the GP iteration on
the device: Nesterov step, WA, both density scatters, Poisson solve, gather, control -- a whole gradient
evaluation (src/gpu_loop.py:59-64 of the reference). Inputs are resident in HBM before the timed region.

One JSON line: value = grad evals / s; roofline of the dominant kernel class, measured with HIP events on the
engine's stream over a replay of the same K passes; cpu_baseline = the CPU oracle timed on a bounded sample.
--full adds a full two-level GP run to overflow <= 0.07 (GP seconds, iterations, HPWL).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12            # B/s
PEAK_FP32_MFMA = 157.3e12    # flop/s


def algorithmic_bytes(cls, dims):
    """Compulsory HBM bytes of one launch of a kernel class (every array element once, pos gathered once per pin)."""
    N, M, P, nets = dims["nodes"], dims["mov"], dims["pins"], dims["nets"]
    nb2, nj2 = dims["nb"] ** 2, dims["nj"] ** 2
    return {
        # u, v, g, lo, hi read; u', v' written (8 B each per node)
        "nesterov": 56 * N,
        # pin cell 4 + offset 8 + position 8 read, pin gradient 8 written per pin; net_start per net
        "wa": 28 * P + 4 * nets,
        # position 8 + size 8 + ratio 4 per node (force grid: all nodes, overflow grid: cells), map RMW per bin 16
        "density_map": 20 * N + 20 * M + 16 * (2 * nb2 + nj2),
        # maps read 8 B/bin/plane, fixed 4, rho written 4 (force); judge map 8 + fixed 4
        "density_post": (2 * 8 + 4 + 4) * nb2 + 12 * nj2,
        # position, size, ratio, pins, area, g_prev read, g written per node; pin gradients via the cell CSR;
        # field planes 8 B/bin
        "gather": (8 + 8 + 4 + 4 + 4 + 8 + 8) * N + (4 + 8) * P + 4 * M + 8 * nb2,
    }.get(cls)


def poisson_flops(nb, method):
    return 12.0 * nb ** 3 if method == "gemm" else None


def make_workload(config, seed=0, log=print):
    import numpy as np
    import torch

    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import _scheduler, build_operators
    from autoplace_amd.calculator import Objective

    t0 = time.time()
    lef, d = synth.generate_config(config, seed=seed)
    data = PlaceData(lef, d)
    args = default_args()
    prob = build_problem(data, args)
    log("design %s: %d cells, %d nets, %d pins, %d fillers, grids %d / %d (%.1fs)"
        % (config, data.num_mov, data.num_nets, data.num_pins, prob.num_fillers, prob.force_bin, prob.num_bin,
           time.time() - t0))
    dev = torch.device("cuda:0")
    wl, den = build_operators(prob, args, dev)
    rng = np.random.default_rng(seed + 1)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = torch.tensor(lo + rng.random(lo.shape).astype(np.float32) * (hi - lo), device=dev)
    ps = _scheduler(den, data.num_mov, data.mov_area, float(data.cell_pins.sum()), float(prob.node_size.prod(1).sum()),
                    len(prob.node_size), args)
    obj = Objective(wl, den, ps)
    ps.density_weight = args.density_weight * obj.tension_unit(pos)
    ps.hpwl.append(float("inf"))
    ps.iter = 50
    dims = {"nodes": len(prob.node_size), "mov": data.num_mov, "pins": int(wl.opt["cell32"].numel()),
            "nets": wl.num_opt_nets, "nb": prob.force_bin, "nj": prob.num_bin if prob.num_bin != prob.force_bin else 0}
    return data, args, prob, wl, den, pos, ps, obj, dims


def gpu_bench(config, steps, warmup, log):
    import torch

    from autoplace_amd import gpu_loop
    from autoplace_amd.nesterov_optimizer import NesterovOptimizer

    data, args, prob, wl, den, pos, ps, obj, dims = make_workload(config, log=log)
    args.inner_iter = 10 ** 9
    bounds = (den.pos_lo, den.pos_hi)
    opt = NesterovOptimizer(pos, obj, 0.01, bounds)
    S, kp, c = gpu_loop.control_state(ps, opt, 50, args)
    dev = pos.device
    S = torch.tensor(S, dtype=torch.float64, device=dev)
    kp = torch.tensor(kp, dtype=torch.float32, device=dev)
    c = torch.tensor(c, dtype=torch.float64, device=dev)
    u, g, best = opt.u.contiguous(), opt.g.contiguous(), pos.clone()
    eng = gpu_loop.Engine(wl, den, u, pos, g, best, S, kp, c)
    chunk = 16 if steps % 16 == 0 and warmup % 16 == 0 else 1
    for _ in range(max(warmup // chunk, 1 if warmup else 0)):
        eng.run(chunk, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        eng.run(chunk, True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = eng.profile(steps)
    torch.cuda.synchronize()
    stop = int(S[gpu_loop.K["stop"]].item())
    eng.close()
    return elapsed, prof, dims, den.plan.method, stop


def cpu_baseline(config, budget_s=25.0, log=print):
    """The CPU oracle (numpy, single thread) on a bounded sample of the same workload: whole passes of the
    iteration at the same state, as many as fit ~budget_s. -> dict."""
    import numpy as np

    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import build_problem, default_args
    from oracle import control as octl
    from oracle import place as opl

    lef, d = synth.generate_config(config, seed=0)
    data = PlaceData(lef, d)
    args = default_args()
    prob = build_problem(data, args)
    n = data.num_mov
    judge = None if prob.force_bin == prob.num_bin else {"nb": prob.num_bin, "fraction": prob.judge_fraction,
                                                          "size": prob.node_size[:n]}
    lvl = opl.Level(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size, n,
                    prob.node_pins, prob.pin_cell, prob.pin_off, prob.pin_net, prob.num_nets, args.ignore_net_degree,
                    judge)
    rng = np.random.default_rng(1)
    pos = (lvl.pos_lo + rng.random(lvl.pos_lo.shape).astype(np.float32) * (lvl.pos_hi - lvl.pos_lo))
    kp = np.array([10 * args.wa_coeff * (lvl.bx + lvl.by), 1e-4, 1e-4, 0.05, 0.0], np.float32)
    u, v, g = pos.copy(), pos.copy(), opl._evaluate_kp(lvl, pos, kp)[0]
    done, t0 = 0, time.time()
    while True:
        un, vn = octl.nesterov_step(u, v, g, lvl.pos_lo, lvl.pos_hi, kp[3], kp[4])
        gn, wa, hpwl, ovf = opl._evaluate_kp(lvl, vn, kp)
        octl.dist(vn, v), octl.dist(gn, g)
        u, v, g = un, vn, gn
        done += 1
        if time.time() - t0 > budget_s * 0.5 or done >= 50:
            break
    el = time.time() - t0
    return {"value": done / el, "unit": "grad evals/s", "cores": 1, "kind": "port",
            "sample": "%d full passes (Nesterov + objective + norms) of the CPU oracle at the %s workload, numpy "
                      "single thread, %.1f s" % (done, config, el)}


def full_run(config, log):
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import default_args
    from autoplace_amd.run_placement import place

    lef, d = synth.generate_config(config, seed=0)
    data = PlaceData(lef, d)
    _, m = place(data, default_args(), "cuda:0", log=log)
    m.pop("cluster_ids", None)
    return {k: m[k] for k in ("gp_time_s", "iters", "grad_calls", "stop_reason", "hpwl_sites", "overflow",
                              "gp_overflow", "clusters", "coarse_iters", "start_iters")}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=64)
    p.add_argument("--warmup", type=int, default=16)
    p.add_argument("--config", default="C")
    p.add_argument("--full", action="store_true", help="also run the whole GP to overflow 0.07")
    p.add_argument("--no-cpu", action="store_true")
    a = p.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the GP iteration of one design runs on one device here "
              "(--gpus %d is not supported)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    log = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    elapsed, prof, dims, method, stop = gpu_bench(a.config, a.steps, a.warmup, log)
    ms_step = elapsed / a.steps * 1e3
    log("per-pass kernel ms: " + ", ".join("%s %.3f" % kv for kv in prof.items()))
    dom = max(prof, key=prof.get)
    if dom == "poisson" and method == "gemm":
        fl = poisson_flops(dims["nb"], method)
        roof = {"kernel": "poisson (fp32 MFMA GEMM DCT)", "bound": "mfma", "achieved": fl / (prof[dom] * 1e-3) / 1e12,
                "peak": PEAK_FP32_MFMA / 1e12, "unit": "TFLOP/s", "traffic": None}
    else:
        by = algorithmic_bytes(dom, dims)
        roof = {"kernel": dom, "bound": "hbm", "achieved": by / (prof[dom] * 1e-3) / 1e9 if by else None,
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "traffic": None}
    if roof["achieved"] is not None:
        roof["frac"] = roof["achieved"] / roof["peak"]
    out = {"metric": "grad evals/s @2M cells (GP iteration, 1 GPU)", "value": a.steps / elapsed,
           "unit": "grad evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (generator seed 0)",
           "config": {"workload": "synthetic %s: %d cells, %d pins, force grid %d, overflow grid %d, poisson %s"
                      % (a.config, dims["mov"], dims["pins"], dims["nb"], dims["nj"] or dims["nb"], method),
                      "parallelism": "single GPU"},
           "roofline": roof, "kernel_ms": prof, "stopped": stop}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.config, log=log)
    if a.full:
        out["full_run"] = full_run(a.config, log)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
