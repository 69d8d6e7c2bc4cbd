"""Batched placement: many independent blocks through the GP iteration together (include/autoplace_gp.h, batch).

Every block is a whole placement problem of its own (die, netlist, grid, control state S / kp / c). Its host
preparation and wirelength-only start run as in the plain flow (run_placement); from the start of the device loop on,
the blocks' passes run together, each stage of a pass one launch over all blocks' workgroups and a run of passes
one HIP graph (ap_batch_*). A stopped block is skipped, so every block ends exactly where its own device loop
(gpu_loop.run_level_on_gpu) would have left it.

    blocks = make_blocks(64, 2000, 50000, seed=0)     # synthetic designs, grids 64 .. 256
    states = [start_state(b) for b in blocks]        # each block's state at the start of its device loop
    res = run_batched(states, autotune=True)          # or run_sequential(states)
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib, gpu_loop, synth
from .database import PlaceData
from .problem import default_args
from .run_placement import place


def block_grid(cells):
    """Power-of-two grid side of a block, 64 .. 256: the power of two at or above sqrt(cells)."""
    return int(min(256, max(64, 2 ** math.ceil(math.log2(math.sqrt(cells))))))


def make_blocks(count=64, lo=2000, hi=50000, seed=0):
    """Synthetic blocks with cells log-uniform in [lo, hi] (nets 1.05 x cells), seeds seed*1000 + 1, 2, ..."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(int)
    out = []
    for k, n in enumerate(cells):
        lef, d = synth.generate(int(n), seed=seed * 1000 + k + 1, name="block%d" % k)
        out.append(PlaceData(lef, d))
    return out


class _Stop(Exception):
    pass


def block_args(data):
    """One level on the block's own grid (no clustering, force grid = overflow grid)."""
    nb = block_grid(data.num_mov)
    return default_args(num_bin=nb, force_num_bin=nb, cluster_ratio=1.0)


def start_state(data, device="cuda:0"):
    """The block's state at the start of its device loop: the plain flow runs until the loop would start.
    -> dict(wl, den, u, v, g, S, kp, c, data)."""
    snap = {}

    def hook(objective, u, v, g, S, kp, c):
        snap.update(wl=objective.wirelength, den=objective.density, u=u.clone(), v=v.clone(), g=g.clone(),
                    S=S.clone(), kp=kp.clone(), c=c.clone())
        raise _Stop()

    try:
        place(data, block_args(data), device, opts={"hook": hook})
    except _Stop:
        pass
    if not snap:
        raise RuntimeError("%s: the flow ended before its device loop" % data.name)
    snap["data"] = data
    return snap


class _Block(object):
    """One block's engine buffers (copies of its start state) and level description."""

    def __init__(self, st):
        self.S, self.kp, self.c = st["S"].clone(), st["kp"].clone(), st["c"].clone()
        self.u, self.v, self.g = st["u"].clone(), st["v"].clone(), st["g"].clone()
        self.best = self.v.clone()
        gpu_loop.setup_level(self, st["wl"], st["den"], self.u, self.v, self.g, self.best, self.S, self.kp, self.c,
                             "fft", True)

    def finish(self):
        if self.perm is not None:
            for mine, theirs in zip((self.u, self.v, self.g, self.best), self.user):
                theirs[self.perm] = mine

    def result(self):
        return _result(self.S)


def _result(S):
    S = S.tolist()
    K = gpu_loop.K
    return {"stop": gpu_loop.REASON.get(int(S[K["stop"]]), "running"), "iters": int(S[K["loop_it"]]),
            "hpwl": S[K["res_hpwl"]], "overflow": S[K["res_ovf"]], "state": S}


def run_sequential(states, check_every=16, max_passes=20000):
    """Every block's device loop alone, one after the other (ap_level, graph replays of check_every passes).
    -> dict(blocks: per-block result + completion time, seconds, positions)."""
    out, pos = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for st in states:
        S, kp, c = st["S"].clone(), st["kp"].clone(), st["c"].clone()
        u, v, g = st["u"].clone(), st["v"].clone(), st["g"].clone()
        eng = gpu_loop.Engine(st["wl"], st["den"], u, v, g, v.clone(), S, kp, c, "fft", True)
        done = 0
        while done < max_passes:
            eng.run(check_every, True)
            done += check_every
            if S[gpu_loop.K["stop"]].item():
                break
        eng.finish()
        eng.close()
        torch.cuda.synchronize()
        r = _result(S)
        r["completion_s"] = time.perf_counter() - t0
        out.append(r)
        pos.append(v.cpu().numpy())
    return {"blocks": out, "seconds": time.perf_counter() - t0, "positions": pos}


def run_batched(states, check_every=16, max_passes=20000, autotune=False, reps=5):
    """All blocks' device loops together (ap_batch). -> dict(blocks, seconds, positions, tile_layout: per block
    [side, tiles per workgroup])."""
    blocks = [_Block(st) for st in states]
    descs = (_lib.LevelDesc * len(blocks))(*[b.desc for b in blocks])
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ap_batch_create(descs, len(blocks), ctypes.byref(h)), "batch_create")
    lib = _lib.lib()
    lay = (ctypes.c_int * (2 * len(blocks)))()
    try:
        if autotune:
            _lib.check(lib.ap_batch_autotune(h, int(reps), lay), "batch_autotune")
        _lib.check(lib.ap_batch_tile_layout(h, lay), "batch_tile_layout")
        _lib.check(lib.ap_batch_prepare(h, check_every), "batch_prepare")
        stop_idx = gpu_loop.K["stop"]
        S_all = [b.S for b in blocks]
        done_at = [None] * len(blocks)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < max_passes:
            _lib.check(lib.ap_batch_run(h, check_every, 1, stream), "batch_run")
            done += check_every
            stops = torch.stack([S[stop_idx] for S in S_all]).tolist()
            now = time.perf_counter() - t0
            for k, v in enumerate(stops):
                if v and done_at[k] is None:
                    done_at[k] = now
            if all(d is not None for d in done_at):
                break
        seconds = time.perf_counter() - t0
    finally:
        torch.cuda.synchronize()
        lib.ap_batch_destroy(h)
    out, pos = [], []
    for b, t in zip(blocks, done_at):
        b.finish()
        r = b.result()
        r["completion_s"] = t
        out.append(r)
        pos.append(b.user[1].cpu().numpy())          # v in the caller's node order
    return {"blocks": out, "seconds": seconds, "positions": pos,
            "tile_layout": [[lay[2 * k], lay[2 * k + 1]] for k in range(len(blocks))]}
