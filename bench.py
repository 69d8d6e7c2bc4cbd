"""Benchmark of the GP iteration on the 2M-cell synthetic design, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C] [--no-cpu] [--no-full]

Workload: the 2,000,000-cell design (2,100,000 nets, ~8.2M pins, 16 fixed macros on 15 % of the die, generator
seed 0, placer seed 0): force grid 2048 x 2048, overflow grid 512 x 512, default options.

1. A whole two-level global placement to overflow <= 0.07 (untimed by this script's step clock; its own GP clock
   is reported): GP seconds, iterations, gradient evaluations, HPWL, overflow -- and it captures the state the
   cell level's device loop starts from (cells just expanded from their clusters, overflow ~0.5).
2. The timed region: from that captured state, W warm-up passes, then exactly K passes of the GP iteration on the
   device (Nesterov step, WA, both density scatters, Poisson solve, gather, control; replayed from a HIP graph),
   bracketed by synchronisations. value = K / time = gradient evaluations per second.
3. Per-kernel times: the same K passes replayed with HIP events on the engine's stream; the dominant kernel's
   roofline fraction uses its algorithmic bytes (every array element once, node data gathered once per pin).
4. cpu_baseline: the CPU oracle (numpy, one thread) running passes from the same captured state, bounded to ~25 s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12            # B/s
PEAK_FP32_MFMA = 157.3e12    # flop/s


def algorithmic_bytes(cls, dims):
    """Compulsory HBM bytes of one launch of a kernel class."""
    N, M, P, nets = dims["nodes"], dims["mov"], dims["pins"], dims["nets"]
    nb2, nj2 = dims["nb"] ** 2, dims["nj"] ** 2
    return {
        # u, v, g, lo, hi read; u', v' written (8 B each per node)
        "nesterov": 56 * N,
        # per pin: cell 4 + offset 8 + position 8 + slot 4 read, gradient 8 written; net_start per net
        "wa": 32 * P + 4 * nets,
        # position 8 + size 8 + ratio 4 per node and grid (force: all nodes, overflow: cells); map written 8 B/bin
        "density_map": 20 * N + 20 * M + 8 * (2 * nb2 + nj2),
        # maps read 8 B/bin/plane, fixed 4, rho written 4 (force); overflow map 8 + fixed 4
        "density_post": (2 * 8 + 4 + 4) * nb2 + 12 * nj2,
        # per node: position, size, ratio, pins, area, g_prev read, g written; pin gradients 8 B/pin, cell_start;
        # field planes 8 B/bin
        "gather": (8 + 8 + 4 + 4 + 4 + 8 + 8) * N + 8 * P + 4 * M + 8 * nb2,
    }.get(cls)


def make_design(config, seed=0):
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    lef, d = synth.generate_config(config, seed=seed)
    return PlaceData(lef, d)


def full_run(data, log):
    """The whole GP run; captures the state of the last level's device loop at its start."""
    import torch

    from autoplace_amd.problem import default_args
    from autoplace_amd.run_placement import place

    snap = {}

    def hook(objective, u, v, g, S, kp, c):
        snap.clear()
        snap.update(objective=objective, u=u.clone(), v=v.clone(), g=g.clone(), S=S.clone(), kp=kp.clone(),
                    c=c.clone())

    args = default_args()
    _, m = place(data, args, "cuda:0", log=log, opts={"hook": hook})
    m.pop("cluster_ids", None)
    torch.cuda.synchronize()
    keys = ("gp_time_s", "iters", "grad_calls", "stop_reason", "hpwl_sites", "overflow", "gp_overflow", "clusters",
            "coarse_iters", "start_iters", "force_num_bin", "num_bin", "num_fillers", "poisson")
    out = {k: m[k] for k in keys if k in m}
    out["grad_evals_per_s"] = m["grad_calls"] / m["gp_time_s"]
    return out, snap


def gpu_passes(snap, steps, warmup):
    import torch

    from autoplace_amd import gpu_loop

    obj = snap["objective"]
    wl, den = obj.wirelength, obj.density
    S, kp, c = snap["S"].clone(), snap["kp"].clone(), snap["c"].clone()
    c[9] = 1e12                                      # inner_iter: no max_iter stop inside the timed window
    u, v, g = snap["u"].clone(), snap["v"].clone(), snap["g"].clone()
    eng = gpu_loop.Engine(wl, den, u, v, g, v.clone(), S, kp, c)
    chunk = 16 if steps % 16 == 0 and warmup % 16 == 0 else 1
    for _ in range(warmup // chunk):
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
    d = den
    dims = {"nodes": d.num_nodes, "mov": d.num_mov, "pins": int(wl.opt["cell32"].numel()), "nets": wl.num_opt_nets,
            "nb": d.nbx, "nj": d.judge.nbx if d.judge is not None else 0}
    method = d.plan.method
    eng.close()
    return elapsed, prof, dims, method, stop


def cpu_baseline(data, snap, budget_s=25.0):
    """Passes of the CPU oracle from the captured state (numpy, one thread), bounded to ~budget_s."""
    import numpy as np

    from autoplace_amd.problem import build_problem, default_args
    from oracle import control as octl
    from oracle import place as opl

    args = default_args()
    prob = build_problem(data, args)
    n = data.num_mov
    judge = None if prob.force_bin == prob.num_bin else {"nb": prob.num_bin, "fraction": prob.judge_fraction,
                                                          "size": prob.node_size[:n]}
    lvl = opl.Level(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size, n,
                    prob.node_pins, prob.pin_cell, prob.pin_off, prob.pin_net, prob.num_nets, args.ignore_net_degree,
                    judge)
    kp = snap["kp"].cpu().numpy().astype(np.float32)
    u, v, g = (snap[k].cpu().numpy() for k in ("u", "v", "g"))
    done, t0 = 0, time.time()
    while True:
        un, vn = octl.nesterov_step(u, v, g, lvl.pos_lo, lvl.pos_hi, kp[3], kp[4])
        gn, wa, hpwl, ovf = opl._evaluate_kp(lvl, vn, kp)
        octl.dist(vn, v), octl.dist(gn, g)
        u, v, g = un, vn, gn
        done += 1
        if time.time() - t0 > budget_s * 0.6 or done >= 50:
            break
    el = time.time() - t0
    return {"value": done / el, "unit": "grad evals/s", "cores": 1, "kind": "port",
            "sample": "%d passes (Nesterov step + objective + norms) of the numpy CPU oracle from the cell level's "
                      "start state of the 2M-cell design, one thread, %.1f s" % (done, el)}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=64)
    p.add_argument("--warmup", type=int, default=16)
    p.add_argument("--config", default="C")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--save-snapshot", default="", help="write the captured state (npz) for kernel studies")
    a = p.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the GP iteration of one design runs on one device here "
              "(--gpus %d is not supported)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    log = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    t0 = time.time()
    data = make_design(a.config)
    log("design %s: %d cells, %d nets, %d pins (%.1fs)" % (a.config, data.num_mov, data.num_nets, data.num_pins,
                                                           time.time() - t0))
    run, snap = full_run(data, log)
    log("full run: %s" % json.dumps(run))
    if a.save_snapshot:
        import numpy as np
        np.savez(a.save_snapshot, **{k: snap[k].cpu().numpy() for k in ("u", "v", "g", "S", "kp", "c")})
    elapsed, prof, dims, method, stop = gpu_passes(snap, a.steps, a.warmup)
    log("per-pass kernel ms: " + ", ".join("%s %.3f" % kv for kv in prof.items()))
    dom = max(prof, key=prof.get)
    if dom == "poisson" and method == "gemm":
        roof = {"kernel": "poisson (fp32 MFMA GEMM DCT)", "bound": "mfma",
                "achieved": 12.0 * dims["nb"] ** 3 / (prof[dom] * 1e-3) / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                "unit": "TFLOP/s", "traffic": None}
    else:
        by = algorithmic_bytes(dom, dims)
        roof = {"kernel": dom, "bound": "hbm", "achieved": by / (prof[dom] * 1e-3) / 1e9 if by else None,
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "traffic": None}
    if roof["achieved"] is not None:
        roof["frac"] = roof["achieved"] / roof["peak"]
    out = {"metric": "grad evals/s @2M cells (GP iteration, 1 GPU)", "value": a.steps / elapsed,
           "unit": "grad evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": elapsed / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (generator seed 0), state captured from a real run",
           "config": {"workload": "synthetic %s: %d cells, %d nets, %d pins, force grid %d, overflow grid %d, "
                                  "poisson %s" % (a.config, data.num_mov, data.num_nets, data.num_pins, dims["nb"],
                                                  dims["nj"] or dims["nb"], method),
                      "parallelism": "single GPU"},
           "roofline": roof, "kernel_ms": prof, "stopped_in_window": stop, "full_run": run}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(data, snap)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
