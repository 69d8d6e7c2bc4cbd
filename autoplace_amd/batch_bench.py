"""Measurement of batched placement (64 blocks of 2k .. 50k cells, grids 64 .. 256) against 64 sequential plain runs
of the same device loops on the same GPU.

    python -m autoplace_amd.batch_bench [--blocks 64] [--lo 2000] [--hi 50000] [--seed 0] [--json PATH]

Both start from every block's state at the start of its device loop (batch.start_state: host preparation and the
wirelength-only start run first, untimed). Reported: aggregate iterations per second (all blocks' passes / wall
time), per-block completion time p50 / p99 (from the start of the timed region; the batch sees a block's stop at
its next 16-pass check), the autotuned tile layouts, and whether every block's final state and placement is the same
bit for bit in both runs (but for the last pass's accept / is_best flags, which every pass after a stop resets).
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=64)
    ap.add_argument("--lo", type=int, default=2000)
    ap.add_argument("--hi", type=int, default=50000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--json", default="")
    a = ap.parse_args(argv)
    from . import batch, gpu_loop
    say = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    t0 = time.time()
    blocks = batch.make_blocks(a.blocks, a.lo, a.hi, a.seed)
    say("%d blocks generated (%.1f s): cells %d .. %d" % (len(blocks), time.time() - t0,
                                                         min(b.num_mov for b in blocks),
                                                         max(b.num_mov for b in blocks)))
    states = [batch.start_state(b) for b in blocks]
    say("start states ready (%.1f s)" % (time.time() - t0))
    seq = batch.run_sequential(states)
    say("sequential: %.3f s" % seq["seconds"])
    bat = batch.run_batched(states, autotune=True)
    say("batched (autotuned): %.3f s" % bat["seconds"])
    transient = [i for i, f in enumerate(gpu_loop.FIELDS) if f in ("accept", "is_best")]
    same = all([x for i, x in enumerate(ra["state"]) if i not in transient] ==
               [x for i, x in enumerate(rb["state"]) if i not in transient] and np.array_equal(pa, pb)
               for ra, rb, pa, pb in zip(seq["blocks"], bat["blocks"], seq["positions"], bat["positions"]))
    iters = sum(r["iters"] for r in seq["blocks"])

    def summary(res):
        c = np.array([r["completion_s"] for r in res["blocks"]])
        return {"seconds": res["seconds"], "iterations_per_s": iters / res["seconds"],
                "completion_p50_s": float(np.percentile(c, 50)), "completion_p99_s": float(np.percentile(c, 99))}

    out = {"metric": "aggregate iterations/s, %d blocks" % len(blocks), "blocks": len(blocks),
           "cells": [b.num_mov for b in blocks], "grids": [batch.block_grid(b.num_mov) for b in blocks],
           "iterations": iters, "stops": sorted({r["stop"] for r in seq["blocks"]}),
           "sequential": summary(seq), "batched": summary(bat), "tile_layout": bat["tile_layout"],
           "identical_results": bool(same)}
    out["speedup"] = out["batched"]["iterations_per_s"] / out["sequential"]["iterations_per_s"]
    line = json.dumps(out)
    print(line)
    if a.json:
        with open(a.json, "w") as f:
            f.write(line + "\n")
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
