"""Training and held-out evaluation of the learned schedule controller (controller.py), on the GPU.

    python -m autoplace_amd.train_controller [--designs 16] [--rollouts 6] [--rounds 8] [--out PATH] ...

1. Training designs: the synthetic generator, cells log-uniform in [--cells-lo, --cells-hi], nets 1.05 x cells,
   seeds 1, 2, ...
2. Per design, the plain rule once: a network with a zero output layer (exactly the rule), logged. Its iterations
   I0 and exact HPWL H0 are the design's reference.
3. Per round, --rollouts runs per design, each with the current network's output layer perturbed (antithetic
   pairs of Gaussian perturbations, --psigma on the output biases, --psigma / 8 on the output weights): every run's
   actions are then a smooth function of its state, which the regression below can reproduce. (--explore action
   adds per-pass Gaussian noise to the outputs instead: --sigma, a new draw every --hold passes.) Reward
   r = -I / I0 - w max(0, H / H0 - 1) (w = --hpwl-weight); -3 for a run that does not converge.
4. Reward-weighted regression: per design, weights softmax(r / --beta) over its runs (the rule's run included, as
   the imitation term), spread evenly over each run's logged passes; the network is fitted to the outputs each
   pass applied (Adam on the device, the library's gradient kernels).
   After each round the fitted policy runs without exploration on every training design; the best-scoring policy
   (the rule, -1, if none beats it) is the one kept.
5. Held-out evaluation: --eval designs of --eval-cells cells, seeds 1001, 1002, ...; the trained network (no noise)
   against the plain rule on the same GPU: iterations, exact HPWL, GP time, stop reason.

Prints one line per run (progress) and writes the weights (+ the evaluation) to --out and a JSON report.
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import controller, synth
from .database import PlaceData
from .problem import default_args
from .run_placement import place


def make(cells, seed):
    lef, d = synth.generate(int(cells), seed=seed, name="ctl%d" % seed)
    return PlaceData(lef, d)


def run(data, params, sigma=0.0, seed=0, hold=1, log_cap=0):
    ctl = controller.options(params, sigma, seed, hold, 50, log_cap)
    _, m = place(data, default_args(), "cuda:0", log=None, opts={"controller": ctl})
    logs = [lv["log"] for lv in ctl.get("levels", []) if lv.get("log") is not None]
    fallback = any(lv["fallback"] for lv in ctl.get("levels", []))
    return m, (np.concatenate(logs) if logs else np.zeros((0, 16), np.float32)), fallback


def output_noise(params, sigma, rng):
    """A Gaussian perturbation of the output layer: sigma on the two biases, sigma / 8 on the weights."""
    n = params.numel()
    out_w = n - controller.OUT - controller.OUT * controller.HID
    noise = np.zeros(n, np.float32)
    noise[out_w:n - controller.OUT] = rng.normal(0, sigma / 8, controller.OUT * controller.HID)
    noise[n - controller.OUT:] = rng.normal(0, sigma, controller.OUT)
    return torch.tensor(noise, device=params.device)


HPWL_WEIGHT = 50.0


def reward(m, ref):
    """-iterations / the rule's, minus HPWL_WEIGHT x the relative HPWL loss; -3 for a run that does not converge."""
    if m.get("stop_reason") != "converged" or not math.isfinite(m["hpwl_sites"]):
        return -3.0
    return -m["iters"] / ref["iters"] - HPWL_WEIGHT * max(0.0, m["hpwl_sites"] / ref["hpwl_sites"] - 1.0)


def main(argv=None):
    global HPWL_WEIGHT
    ap = argparse.ArgumentParser()
    ap.add_argument("--designs", type=int, default=16)
    ap.add_argument("--cells-lo", type=float, default=20000)
    ap.add_argument("--cells-hi", type=float, default=200000)
    ap.add_argument("--hpwl-weight", type=float, default=HPWL_WEIGHT)
    ap.add_argument("--rollouts", type=int, default=6)
    ap.add_argument("--rounds", type=int, default=8)
    ap.add_argument("--explore", choices=("param", "action"), default="param")
    ap.add_argument("--psigma", type=float, default=0.0015)
    ap.add_argument("--sigma", type=float, default=0.015)
    ap.add_argument("--hold", type=int, default=40)
    ap.add_argument("--beta", type=float, default=0.05)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--eval", type=int, default=20)
    ap.add_argument("--eval-cells", type=int, default=500000)
    ap.add_argument("--out", default=controller.WEIGHTS)
    ap.add_argument("--init", default="", help="start from these weights (with --designs 0: evaluate them only)")
    ap.add_argument("--report", default="")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    say = lambda s: print(s, flush=True)   # noqa: E731
    HPWL_WEIGHT = a.hpwl_weight
    t_start = time.time()
    rng = np.random.default_rng(a.seed)
    params = controller.init_params(a.seed)
    zero = params.clone()
    prior = {}
    if a.init:
        params, meta = controller.load(a.init)
        prior = json.loads(meta.get("report", "{}")) if isinstance(meta.get("report"), str) else {}
    cells = np.exp(rng.uniform(math.log(a.cells_lo), math.log(a.cells_hi), a.designs)).astype(int)
    designs = []
    for k in range(a.designs):
        data = make(cells[k], k + 1)
        m, log, _ = run(data, zero, log_cap=4096)
        designs.append({"data": data, "ref": m, "runs": [(log, reward(m, m))]})
        say("design %d: %d cells, rule: %d iterations, HPWL %.6e, %s" % (k + 1, data.num_mov, m["iters"],
                                                                     m["hpwl_sites"], m["stop_reason"]))
    history = []
    best_params, best_score = zero.clone(), -1.0            # the rule scores -1 on every design
    if a.init and designs:
        score = float(np.mean([reward(run(d["data"], params)[0], d["ref"]) for d in designs]))
        say("starting policy (%s) scores %.4f" % (a.init, score))
        if score > best_score:
            best_params, best_score = params.clone(), score
    for rnd in range(a.rounds if designs else 0):
        rs = []
        for k, d in enumerate(designs):
            noise = None
            for r in range(a.rollouts):
                seed = 1000003 * (rnd + 1) + 1009 * k + r
                if a.explore == "param":
                    if r % 2 == 0:
                        noise = output_noise(params, a.psigma, np.random.default_rng(seed))
                    m, log, fb = run(d["data"], params + (noise if r % 2 == 0 else -noise), 0.0, 0, 1, 4096)
                else:
                    m, log, fb = run(d["data"], params, a.sigma, seed, a.hold, 4096)
                rw = reward(m, d["ref"])
                d["runs"].append((log, rw))
                rs.append(rw)
                say("round %d design %d rollout %d: %d iterations (rule %d), HPWL ratio %.4f, %s%s, reward %.3f"
                    % (rnd, k + 1, r, m["iters"], d["ref"]["iters"], m["hpwl_sites"] / d["ref"]["hpwl_sites"],
                       m["stop_reason"], " (fallback)" if fb else "", rw))
        xs, ts, ws = [], [], []
        for d in designs:
            rews = np.array([rw for _, rw in d["runs"]])
            wts = np.exp((rews - rews.max()) / a.beta)
            wts /= wts.sum()
            for (log, _), wt in zip(d["runs"], wts):
                if len(log) == 0 or wt < 1e-6:
                    continue
                xs.append(log[:, :12])
                ts.append(log[:, 12:14])
                ws.append(np.full(len(log), wt / len(log), np.float32))
        x = torch.tensor(np.concatenate(xs), device="cuda:0")
        t = torch.tensor(np.concatenate(ts), device="cuda:0")
        w = torch.tensor(np.concatenate(ws), device="cuda:0")
        loss = controller.fit(params, x, t, w, epochs=a.epochs, seed=a.seed + rnd)
        # the fitted policy itself, without exploration, on every training design
        score = float(np.mean([reward(run(d["data"], params)[0], d["ref"]) for d in designs]))
        if score > best_score:
            best_params, best_score = params.clone(), score
        history.append({"round": rnd, "mean_reward": float(np.mean(rs)), "samples": int(x.shape[0]), "loss": loss,
                        "policy_reward": score})
        say("round %d: mean rollout reward %.4f over %d rollouts, fit on %d passes, weighted loss %.3e; the fitted "
            "policy scores %.4f (best %.4f; the rule -1)" % (rnd, np.mean(rs), len(rs), x.shape[0], loss, score,
                                                             best_score))
    if designs:
        params = best_params
    evals = []
    for k in range(a.eval):
        data = make(a.eval_cells, 1001 + k)
        rule, _, _ = run(data, None)
        learned, _, fb = run(data, params)
        row = {"seed": 1001 + k, "cells": data.num_mov,
               "rule": {q: rule[q] for q in ("iters", "hpwl_sites", "gp_time_s", "stop_reason", "overflow")},
               "learned": {q: learned[q] for q in ("iters", "hpwl_sites", "gp_time_s", "stop_reason", "overflow")},
               "fallback": fb}
        evals.append(row)
        say("held-out %d (%d cells): rule %d it HPWL %.6e %s | learned %d it HPWL %.6e %s%s"
            % (1001 + k, data.num_mov, rule["iters"], rule["hpwl_sites"], rule["stop_reason"], learned["iters"],
               learned["hpwl_sites"], learned["stop_reason"], " (fallback)" if fb else ""))
    it_ratio = [e["learned"]["iters"] / e["rule"]["iters"] for e in evals]
    hp_ratio = [e["learned"]["hpwl_sites"] / e["rule"]["hpwl_sites"] for e in evals]
    if designs:
        summary = {"designs": a.designs, "cells": [int(c) for c in cells], "rollouts": a.rollouts,
                   "rounds": a.rounds, "explore": a.explore, "psigma": a.psigma, "sigma": a.sigma, "hold": a.hold,
                   "beta": a.beta, "hpwl_weight": a.hpwl_weight, "history": history,
                   "train_seconds": time.time() - t_start}
    else:
        summary = {k: v for k, v in prior.items() if not k.startswith("eval")}
    summary.update({"eval": evals, "eval_cells": a.eval_cells,
                    "eval_iter_ratio_mean": float(np.mean(it_ratio)) if it_ratio else None,
                    "eval_hpwl_ratio_mean": float(np.mean(hp_ratio)) if hp_ratio else None,
                    "seconds": time.time() - t_start})
    controller.save(a.out, params, report=json.dumps(summary))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(summary, f, indent=1)
    say("held-out: iterations x%.3f, HPWL x%.4f (means over %d designs); %.0f s"
        % (summary["eval_iter_ratio_mean"] or float("nan"), summary["eval_hpwl_ratio_mean"] or float("nan"),
           len(evals), summary["seconds"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
