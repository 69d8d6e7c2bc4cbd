"""Benchmark of the GP iteration on the 2M-cell synthetic design, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C] [--no-cpu]
                    [--save-snapshot F] [--snapshot F] [--copy-calibration] [--tile-layout F,G,J,H]

Workload: the 2,000,000-cell design (2,100,000 nets, ~8.2M pins, 16 fixed macros on 15 % of the die, generator
seed 0, placer seed 0): force grid 2048 x 2048, overflow grid 512 x 512, default options.

1. A whole two-level global placement to overflow <= 0.07 (untimed by this script's step clock; its own GP clock
   is reported): GP seconds, iterations, gradient evaluations, HPWL, overflow -- and it captures the state the
   cell level's device loop starts from (cells just expanded from their clusters). A 20,000-cell placement runs
   first, so that the process's first kernel loads are not inside the GP clock.
2. W untimed warm-up passes from that state: W/4 with direct launches, every kernel class in order on one stream with
   HIP events between them (the per-class breakdown `kernel_ms`); the rest as graph runs, as the timed region runs
   them, with HIP events around one single-kernel class per run. The class with the longest time per pass in those
   runs is the roofline kernel.
3. The timed region: exactly K passes of the GP iteration on the device (Nesterov step, WA, density binning of both
   grids, Poisson solve, gather, control), one replay of a HIP graph captured beforehand, bracketed by
   synchronisations. value = K / time = gradient evaluations per second. In the graph WA and the overflow grid run
   on a second stream beside the force grid and the Poisson solve. The graph carries a HIP event before and after the
   roofline kernel in every pass (on the stream that kernel runs on): its average duration over the K passes gives
   `roofline.achieved` = its algorithmic bytes per launch / that duration.
4. cpu_baseline: the CPU oracle (numpy, one thread) running passes from the same captured state, bounded to ~25 s.

--snapshot F starts from a state written by --save-snapshot (no full run, no CPU baseline): the form used under
rocprofv3, together with --copy-calibration (three 256 MiB device copies first, the scale of the FETCH_SIZE /
WRITE_SIZE counters) and --tile-layout F,G,J,H (the autotuner's choice of tile side and tiles per workgroup for
the force and the overflow grid, replayed, so that its trial launches stay out of the profile).
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
TIME_EVERY = 8               # the roofline kernel's events: in passes 0, 8, 16, ... of the timed run

# kernel classes whose events in graph runs bracket one kernel (the binning class brackets the force grid's fill, the
# overflow grid's running on the other stream; the Poisson solve and the control launch several)
SINGLE_KERNEL = {"nesterov": "eng_nesterov", "wa": "wa_wave_kernel", "wa_cells": "eng_cell_wl",
                 "density_bin": "eng_bin_fill", "density_tiles": "eng_bin_tiles", "overflow_tiles": "eng_over_tiles",
                 "gather": "eng_gather"}


def algorithmic_bytes(cls, dims):
    """Compulsory HBM bytes of one launch of a kernel class: every array element read or written once, positions
    gathered once per pin."""
    N, M, P, nets = dims["nodes"], dims["mov"], dims["pins"], dims["nets"]
    nb2, nj2 = dims["nb"] ** 2, dims["nj"] ** 2
    return {
        # u, v, g, lo, hi read; u', v' written (8 B each per node)
        "nesterov": 56 * N,
        # per pin: cell 4 + offset 8 + position 8 read, gradient 8 written (the cell sums are wa_cells');
        # per net: its start 4
        "wa": 28 * P + 4 * nets,
        # per pin: its gradient 8 read; per cell: cell_start 4 read, its gradient 8 written
        "wa_cells": 8 * P + 12 * M,
        # the force grid's fill: position 8 + size 8 + ratio 4 read per node (the bucket entries it writes and the
        # tiles read back are intermediates)
        "density_bin": 20 * N,
        # the force map: fixed 4 read and density 4 written per bin
        "density_tiles": 8 * nb2,
        # the overflow map: fixed 4 read per bin
        "overflow_tiles": 4 * nj2,
        # per node: position 8, size 8, ratio 4, pins 4, area 4, previous gradient 8 read, gradient 8 written;
        # per cell: its WA gradient 8; the field 8 per bin
        "gather": 44 * N + 8 * M + 8 * nb2,
    }.get(cls)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "engine_2M_cells_pmc_traffic.md")


def profile_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE summary of the 2M-cell design
    (copy-calibrated), or None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == "ap::" + kernel:
            return float(cells[4]) * 1e6
    return None


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
        snap.update(wl=objective.wirelength, den=objective.density, u=u.clone(), v=v.clone(), g=g.clone(),
                    S=S.clone(), kp=kp.clone(), c=c.clone())

    args = default_args()
    # a 20,000-cell placement first, untimed: the process's first torch and HIP kernel loads (clustering, sorting)
    # happen there, as in a placer that has run before, and not inside the timed run
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    lef, d = synth.generate(20000, seed=999, name="warmup")
    place(PlaceData(lef, d), default_args(), "cuda:0")
    torch.cuda.synchronize()
    _, m = place(data, args, "cuda:0", log=log, opts={"hook": hook})
    m.pop("cluster_ids", None)
    torch.cuda.synchronize()
    keys = ("gp_time_s", "iters", "grad_calls", "stop_reason", "hpwl_sites", "overflow", "gp_overflow", "clusters",
            "coarse_iters", "start_iters", "force_num_bin", "num_bin", "num_fillers", "poisson", "coarse_spills",
            "spills")
    out = {k: m[k] for k in keys if k in m}
    out["grad_evals_per_s"] = m["grad_calls"] / m["gp_time_s"]
    return out, snap


def load_snapshot(data, path):
    """State written by --save-snapshot, with the cell level's operators rebuilt as the full run builds them."""
    import numpy as np
    import torch

    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators

    z = np.load(path)
    args = default_args()
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, torch.device("cuda:0"))
    dt = {"S": torch.float64, "c": torch.float64}
    snap = {k: torch.tensor(z[k], dtype=dt.get(k, torch.float32), device="cuda:0") for k in
            ("u", "v", "g", "S", "kp", "c")}
    snap.update(wl=wl, den=den)
    return snap


def copy_calibration():
    """Three plain 256 MiB device-to-device copies: the FETCH_SIZE / WRITE_SIZE reading of known traffic."""
    import torch
    src = torch.ones(64 * 1024 * 1024, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def gpu_passes(snap, steps, warmup, layout=None):
    import torch

    from autoplace_amd import gpu_loop

    wl, den = snap["wl"], snap["den"]
    S, kp, c = snap["S"].clone(), snap["kp"].clone(), snap["c"].clone()
    c[9] = 1e12                                      # inner_iter: no max_iter stop inside the timed window
    u, v, g = snap["u"].clone(), snap["v"].clone(), snap["g"].clone()
    eng = gpu_loop.Engine(wl, den, u, v, g, v.clone(), S, kp, c)
    if layout:
        eng.set_tile_layout(layout[:2], layout[2:])  # a choice of the autotuner, replayed (profiling runs)
        lay = (tuple(layout[:2]), tuple(layout[2:]))
    else:
        lay = eng.autotune(5)                        # as the flow does for levels of 100k nodes and more
    print("density tile layout (%s): force grid side %d x %d per workgroup, overflow grid side %d x %d"
          % ((("given" if layout else "autotuned"),) + lay[0] + lay[1]), file=sys.stderr, flush=True)
    # the warm-up: a quarter of the passes profiled (every class in order on one stream: the per-class breakdown),
    # the rest as graph runs with events around one single-kernel class each (the passes as the timed region runs
    # them, the two streams side by side); the roofline kernel is the one with the longest time in those runs
    n_prof = max(warmup // 4, 1)
    prof = eng.profile(n_prof)
    per = max((warmup - n_prof) // len(SINGLE_KERNEL), 1)
    in_graph_ms = {}
    for cls in SINGLE_KERNEL:
        eng.time_class(cls)
        eng.run(per, True)
        in_graph_ms[cls] = eng.class_ms()[0]
    dom = max(in_graph_ms, key=in_graph_ms.get)
    print("single-kernel classes in graph runs (ms per pass): "
          + ", ".join("%s %.4f" % kv for kv in in_graph_ms.items()), file=sys.stderr, flush=True)
    # event-record nodes around the roofline kernel in every TIME_EVERY-th pass of the timed graph: an event record
    # holds its queue for a few microseconds, which the timed passes would otherwise carry in every pass
    eng.time_every(TIME_EVERY)
    eng.time_class(dom)
    try:
        eng.prepare(steps)                           # graph with event-record nodes around the roofline kernel
        in_graph = True
    except RuntimeError as e:
        print("event-record nodes in the graph failed (%s): the roofline kernel is timed over K direct passes after "
              "the timed region" % e, file=sys.stderr, flush=True)
        eng.time_class(None)
        eng.prepare(steps)
        in_graph = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(steps, True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not in_graph:
        eng.time_class(dom)
        eng.run(steps, False)
    dom_ms, n = eng.class_ms()
    assert n == (steps + TIME_EVERY - 1) // TIME_EVERY
    stop = int(S[gpu_loop.K["stop"]].item())
    d = den
    dims = {"nodes": d.num_nodes, "mov": d.num_mov, "pins": int(wl.opt["cell32"].numel()), "nets": wl.num_opt_nets,
            "nb": d.nbx, "nj": d.judge.nbx if d.judge is not None else 0}
    method = d.plan.method
    print("density entries that took the spill path since the engine's set-up (force grid, overflow grid): %s"
          % (eng.bucket_spills(),), file=sys.stderr, flush=True)
    eng.close()
    return elapsed, prof, dom, dom_ms, dims, method, stop, in_graph, in_graph_ms


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
    p.add_argument("--warmup", type=int, default=64)
    p.add_argument("--config", default="C")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--save-snapshot", default="", help="write the captured state (npz)")
    p.add_argument("--snapshot", default="", help="start from a saved state: no full run, no CPU baseline")
    p.add_argument("--copy-calibration", action="store_true", help="three 256 MiB device copies first")
    p.add_argument("--tile-layout", default="", help="F,G,J,H: force side, group, overflow side, group "
                                                       "(skip the autotuner)")
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
    if a.copy_calibration:
        copy_calibration()
    if a.snapshot:
        run, snap = None, load_snapshot(data, a.snapshot)
    else:
        run, snap = full_run(data, log)
        log("full run: %s" % json.dumps(run))
    if a.save_snapshot:
        import numpy as np
        np.savez(a.save_snapshot, **{k: snap[k].cpu().numpy() for k in ("u", "v", "g", "S", "kp", "c")})
    layout = tuple(int(x) for x in a.tile_layout.split(",")) if a.tile_layout else None
    elapsed, prof, dom, dom_ms, dims, method, stop, in_graph, in_graph_ms = gpu_passes(snap, a.steps, a.warmup,
                                                                                       layout)
    log("per-pass kernel ms (warm-up, direct launches): " + ", ".join("%s %.3f" % kv for kv in prof.items()))
    log("roofline kernel %s (%s): %.4f ms per launch over the timed passes" % (dom, SINGLE_KERNEL[dom], dom_ms))
    by = algorithmic_bytes(dom, dims)
    roof = {"kernel": SINGLE_KERNEL[dom], "bound": "hbm", "achieved": by / (dom_ms * 1e-3) / 1e9,
            "peak": PEAK_HBM / 1e9, "unit": "GB/s", "traffic": None, "algorithmic_bytes": by, "launch_ms": dom_ms,
            "events": ("event-record nodes in the timed graph, every %d-th pass" % TIME_EVERY if in_graph
                       else "direct launches after the timed region, every %d-th pass" % TIME_EVERY)}
    roof["frac"] = roof["achieved"] / roof["peak"]
    if a.config == "C":
        roof["traffic"] = profile_traffic(SINGLE_KERNEL[dom])
        if roof["traffic"] is not None:
            roof["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT)
    out = {"metric": "grad evals/s @2M cells", "value": a.steps / elapsed,
           "unit": "grad evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": elapsed / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (generator seed 0), state captured from a real run",
           "config": {"workload": "synthetic %s: %d cells, %d nets, %d pins, force grid %d, overflow grid %d, "
                                  "poisson %s" % (a.config, data.num_mov, data.num_nets, data.num_pins, dims["nb"],
                                                  dims["nj"] or dims["nb"], method),
                      "parallelism": "single GPU"},
           "roofline": roof, "kernel_ms": prof, "kernel_ms_in_graph": in_graph_ms, "stopped_in_window": stop}
    if run is not None:
        out["full_run"] = run
    if not a.no_cpu and not a.snapshot:
        out["cpu_baseline"] = cpu_baseline(data, snap)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
