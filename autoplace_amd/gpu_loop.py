"""The rest of a level with the run's control on the device (src/gpu_loop.py:1-81 of the reference).

What NesterovOptimizer.step, ParamScheduler.step and ParamScheduler.need_to_stop decide on the host is decided by
the control kernel, in the same order and in double precision; the kernels read gamma, the density weight and the
step from device memory. Each host check runs `check_every` passes in one native call (ap_level_run, a captured HIP
graph after the first call); once the run has stopped the control freezes, so the result is that of a run checked
after every pass.
"""
import ctypes
import math
import time

import torch

from . import _lib
from .locality import renumber_level

FIELDS = ("lam gam alpha a iter hpwl_prev motion ww mu settled best_hpwl best_ovf is_best improved stop accept "
          "loop_it evals res_hpwl res_ovf has_best").split()
K = {name: i for i, name in enumerate(FIELDS)}
REASON = {1: "converged", 2: "stalled", 3: "diverged", 4: "max_iter", 5: "handover"}


def usable(optimizer):
    return optimizer.kernel


def as_log(log):
    """A line sink from what the caller passes: a logging.Logger (the reference passes one, gpu_loop.py:35,66-68, and
    calls its info), any callable taking one string, or None (silent)."""
    if log is None or callable(log) and not hasattr(log, "info"):
        return log
    if hasattr(log, "info"):
        return log.info
    raise TypeError("log must be a logging.Logger, a callable or None, not %s" % type(log).__name__)


def mark(timeline, name, device):
    """Instrumentation: with a timeline list, synchronise the device and append (name, wall time)."""
    if timeline is not None:
        torch.cuda.synchronize(device)
        timeline.append((name, time.time()))


def next_coef(a):
    a_next = (1 + math.sqrt(4 * a * a + 1)) / 2
    return (a - 1) / a_next


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def setup_level(obj, wirelength, density, u, v, g, best, S, kp, c, poisson="auto", reorder=True):
    """The ap_level_desc of a level (obj.desc), with obj holding the tensors it points to (obj._keep), the spatial
    renumbering (obj.perm; None without reorder) and the engine's copies of u / v / g / best (obj.u ...; the
    caller's in obj.user)."""
    w, d, j = wirelength.opt, density, density.judge
    obj.user = (u, v, g, best)
    n, N = d.num_mov, d.num_nodes
    if reorder:
        perm, net = renumber_level(v, d.nbx, d.nby, float(d.bin_np[0]), float(d.bin_np[1]), n, w["cell32"], w["off"],
                                   w["net_start"])
    else:
        perm = None
        net = w
    pm = (lambda t: t[perm].contiguous()) if reorder else (lambda t: t)           # noqa: E731
    pc = (lambda t: t[perm[:n]].contiguous()) if reorder else (lambda t: t)       # noqa: E731
    obj.perm = perm
    obj.S, obj.kp, obj.c = S, kp, c
    obj.u, obj.v, obj.g, obj.best = pm(u), pm(v), pm(g), pm(best)
    arrays = dict(size=pm(d.smooth["size"]), ratio=pm(d.smooth["ratio"]), pins=pm(d.pins), area=pm(d.area),
                  lo=pm(d.pos_lo), hi=pm(d.pos_hi))
    obj._keep = [net["cell32"], net["off"], net["net_start"], net["cell_start"], net["cell_pin"], obj.u, obj.v,
                 obj.g, obj.best, S, kp, c, d.fixed_map] + list(arrays.values())
    for t in obj._keep:
        if not t.is_contiguous():
            raise ValueError("engine tensors must be contiguous")
    desc = _lib.LevelDesc()
    desc.num_nodes, desc.num_mov = N, n
    desc.num_pins, desc.num_nets = net["cell32"].numel(), net["num_nets"]
    desc.pin_cell, desc.pin_off, desc.net_start = _p(net["cell32"]), _p(net["off"]), _p(net["net_start"])
    desc.cell_start, desc.cell_pin = _p(net["cell_start"]), _p(net["cell_pin"])
    desc.size, desc.ratio = _p(arrays["size"]), _p(arrays["ratio"])
    desc.pins, desc.area, desc.lo, desc.hi = _p(arrays["pins"]), _p(arrays["area"]), _p(arrays["lo"]), \
        _p(arrays["hi"])
    desc.nbx, desc.nby = d.nbx, d.nby
    desc.bx, desc.by = float(d.bin_np[0]), float(d.bin_np[1])
    desc.fixed, desc.target = _p(d.fixed_map), float(d.target)
    desc.bin_area, desc.mov_area = d.bin_area, d.mov_area
    if j is not None:
        js, jr = pc(j.smooth["size"]), pc(j.smooth["ratio"])
        obj._keep += [js, jr, j.fixed_map]
        desc.has_judge, desc.njx, desc.njy = 1, j.nbx, j.nby
        desc.jbx, desc.jby, desc.judge_bin_area = float(j.bin_np[0]), float(j.bin_np[1]), j.bin_area
        desc.judge_size, desc.judge_ratio, desc.judge_fixed = _p(js), _p(jr), _p(j.fixed_map)
    desc.poisson_method = {"auto": 0, "gemm": 1, "fft": 2}[poisson]
    desc.u, desc.v, desc.g, desc.best = _p(obj.u), _p(obj.v), _p(obj.g), _p(obj.best)
    desc.S, desc.kp, desc.c = _p(S), _p(kp), _p(c)
    # the density layer's Poisson plan (built with the layer, before the level's passes): the level uses it instead of
    # building its own; obj keeps the plan alive
    obj._plan = d.plan
    desc.poisson_plan = d.plan._h if d.plan is not None else None
    obj.desc = desc


class Engine(object):
    """A level's problem bound to the native GP iteration engine (ap_level).

    With reorder=True (default) the engine works on a spatially renumbered copy of the level (locality.py): the
    node buffers u / v / g / best are permuted in, and `finish()` writes them back in the caller's order. Keeps
    every tensor it hands to the library alive."""

    def __init__(self, wirelength, density, u, v, g, best, S, kp, c, poisson="auto", reorder=True):
        setup_level(self, wirelength, density, u, v, g, best, S, kp, c, poisson, reorder)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ap_level_create(ctypes.byref(self.desc), ctypes.byref(h)), "level_create")
        self._h = h

    def run(self, passes, use_graph=True):
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().ap_level_run(self._h, passes, int(use_graph), stream), "level_run")

    def autotune(self, reps=5, online=False):
        """Time the density tile layouts around the set-up ones (side, tiles per workgroup) and keep the fastest
        (results unchanged). Offline: from the current point, restored afterwards. online: on the level's own next
        passes, which advance it (no extra passes). -> ((force side, group), (overflow side, group) or (0, 0))."""
        lay = (ctypes.c_int * 4)()
        flags = _lib.AP_AUTOTUNE_ONLINE if online else 0
        _lib.check(_lib.lib().ap_level_autotune(self._h, int(reps), flags, lay), "level_autotune")
        return (lay[0], lay[1]), (lay[2], lay[3])

    def set_tile_layout(self, force, overflow=(0, 0)):
        """force / overflow: (side, group); group 0 = the default for the side."""
        _lib.check(_lib.lib().ap_level_set_tile_layout(self._h, int(force[0]), int(force[1]), int(overflow[0]),
                                                       int(overflow[1])), "level_set_tile_layout")

    def launch_shape(self):
        """(WA plan wavefronts per launched wavefront, WA lanes per workgroup, gather lanes per workgroup)."""
        out = (ctypes.c_int * 3)()
        _lib.check(_lib.lib().ap_level_launch_shape(self._h, ctypes.cast(out, ctypes.c_void_p)), "level_launch_shape")
        return tuple(int(x) for x in out)

    def set_launch_shape(self, wa_per, wa_lanes, gather_lanes):
        """Set the WA and gather launch shapes (the autotuner's choice, replayed; results do not depend on them)."""
        _lib.check(_lib.lib().ap_level_set_launch_shape(self._h, int(wa_per), int(wa_lanes), int(gather_lanes)),
                   "level_set_launch_shape")

    def set_bucket_room(self, percent=150, extra=64):
        """Each density bucket's room per pass: percent % of its last count + extra (smaller: more entries spill)."""
        _lib.check(_lib.lib().ap_level_set_bucket_room(self._h, int(percent), int(extra)), "level_set_bucket_room")

    def bucket_spills(self):
        """(force, overflow): entries that took the density binning's spill path since the engine was created."""
        out = (ctypes.c_ulonglong * 2)()
        _lib.check(_lib.lib().ap_level_bucket_spills(self._h, ctypes.cast(out, ctypes.c_void_p)), "level_bucket_spills")
        return int(out[0]), int(out[1])

    def set_tile_split(self, slice_entries=4096, factor=4):
        """Split a density tile's bucket over up to 16 workgroups when it had 2 x slice_entries or more and factor x
        the grid's mean per tile or more (slice_entries 0: never)."""
        _lib.check(_lib.lib().ap_level_set_tile_split(self._h, int(slice_entries), int(factor)),
                   "level_set_tile_split")

    def tile_splits(self):
        """(force, overflow): split density tiles summed since the engine was created."""
        out = (ctypes.c_ulonglong * 2)()
        _lib.check(_lib.lib().ap_level_tile_splits(self._h, ctypes.cast(out, ctypes.c_void_p)), "level_tile_splits")
        return int(out[0]), int(out[1])

    def finish(self):
        """Write u / v / g / best back into the caller's tensors (caller's node order)."""
        if self.perm is None:
            return
        for mine, theirs in zip((self.u, self.v, self.g, self.best), self.user):
            theirs[self.perm] = mine

    def start_begin(self, tol, window, max_passes):
        """Reset the wirelength-only start (ap_level_start_begin): the rule's tolerance and window, the pass limit."""
        _lib.check(_lib.lib().ap_level_start_begin(self._h, float(tol), int(window), int(max_passes)),
                   "level_start_begin")

    def start_run(self, passes, use_graph=True):
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().ap_level_start_run(self._h, passes, int(use_graph), stream), "level_start_run")

    def start_state(self, history=False):
        """-> (done: 0 running, 1 the start rule, 2 diverged, 3 the pass limit; passes taken; [(WA, HPWL)] of every
        pass taken when history)."""
        st = (ctypes.c_double * 8)()
        _lib.check(_lib.lib().ap_level_start_state(self._h, st, None, 0), "level_start_state")
        done, count = int(st[0]), int(st[1])
        if not history or count == 0:
            return done, count, []
        h = (ctypes.c_double * (2 * count))()
        _lib.check(_lib.lib().ap_level_start_state(self._h, st, h, count), "level_start_state")
        return done, count, [(h[2 * k], h[2 * k + 1]) for k in range(count)]

    def rebind(self, u, v, g, best, S, kp, c):
        """Continue with the level's current point as the start of the next phase: the caller's tensors that
        finish() writes, the control state's new values (copied into the level's S / kp / c), best = the current
        point."""
        self.user = (u, v, g, best)
        self.S.copy_(S)
        self.kp.copy_(kp)
        self.c.copy_(c)
        self.best.copy_(self.v)

    def prepare(self, passes):
        """Capture the graph of `passes` passes now (outside any timed region)."""
        _lib.check(_lib.lib().ap_level_prepare(self._h, passes), "level_prepare")

    def profile(self, passes):
        """Average ms per pass of every kernel class (direct launches with HIP events). -> dict."""
        ms = (ctypes.c_double * len(_lib.PROF_CLASSES))()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().ap_level_profile(self._h, passes, ms, stream), "level_profile")
        return dict(zip(_lib.PROF_CLASSES, list(ms)))

    def time_class(self, name):
        """Graphs captured from now on carry HIP events around the kernels of class `name` (None: no events)."""
        cls = -1 if name is None else _lib.PROF_CLASSES.index(name)
        _lib.check(_lib.lib().ap_level_time_class(self._h, cls), "level_time_class")

    def time_every(self, every):
        """Events around the timed class in passes 0, every, 2 every, ... of a run only (1: every pass)."""
        _lib.check(_lib.lib().ap_level_time_every(self._h, int(every)), "level_time_every")

    def class_ms(self):
        """(average ms of the timed class over the last run's timed passes, number of timed passes)."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _lib.check(_lib.lib().ap_level_class_ms(self._h, ctypes.byref(ms), ctypes.byref(n)), "level_class_ms")
        return ms.value, n.value

    def set_controller(self, ctl, phase):
        """Attach the learned controller (dict: params device f32 tensor or None, sigma, seed, hold, window, log_cap).
        -> the log tensor [log_cap, 16] (None without a log)."""
        params = ctl.get("params")
        cap = int(ctl.get("log_cap", 0))
        log = torch.zeros((cap, 16), dtype=torch.float32, device=self.v.device) if cap > 0 else None
        self._ctl = (params, log)
        seed = (int(ctl.get("seed", 0)) * 2 + (1 if phase else 0)) & 0xFFFFFFFFFFFFFFFF
        _lib.check(_lib.lib().ap_level_set_controller(
            self._h, _p(params) if params is not None else None, float(ctl.get("sigma", 0.0)), seed,
            int(ctl.get("hold", 1)), float(phase),
            _p(log) if log is not None else None, cap, int(ctl.get("window", 50))), "level_set_controller")
        return log

    def controller_state(self):
        st = (ctypes.c_double * 8)()
        _lib.check(_lib.lib().ap_level_controller_state(self._h, st), "level_controller_state")
        return list(st)

    def density(self):
        """Force-grid density [nbx, nby] of the last objective evaluated, in the engine's bin layout."""
        d = self.desc
        out = torch.empty((d.nbx, d.nby), dtype=torch.float32, device=self.v.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().ap_level_density(self._h, _p(out), stream), "level_density")
        return out

    def last_ms(self):
        ms = ctypes.c_float()
        _lib.check(_lib.lib().ap_level_last_ms(self._h, ctypes.byref(ms)), "level_last_ms")
        return ms.value

    def objective(self, pos, grad, stats):
        """One objective evaluation at pos (in the engine's node order) with the parameters in kp."""
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().ap_level_objective(self._h, _p(pos), _p(grad), _p(stats), stream), "level_objective")

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ap_level_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def control_state(ps, optimizer, iteration, args):
    """(S f64[21], kp f32[5], c f64[13]) from the host state (gpu_loop.py:39-53). Before a level's first step (the
    start phase's state) there is no previous HPWL: hpwl_prev is then +inf."""
    S = [0.0] * len(FIELDS)
    S[K["lam"]], S[K["gam"]], S[K["alpha"]], S[K["a"]] = ps.density_weight, ps.wa_coeff, optimizer.alpha, optimizer.a
    S[K["iter"]], S[K["motion"]], S[K["ww"]] = ps.iter, ps.motion, ps.weighted_weight
    S[K["hpwl_prev"]] = ps.hpwl[-1] if ps.hpwl else float("inf")
    S[K["mu"]], S[K["settled"]], S[K["improved"]] = ps.mu, ps.settled, ps.improved_iter
    S[K["best_hpwl"]], S[K["best_ovf"]] = ps.best_hpwl, ps.best_overflow
    S[K["loop_it"]] = iteration
    kp = [ps.wa_coeff, ps.density_weight, ps.density_weight, optimizer.alpha, next_coef(optimizer.a)]
    a = ps.args
    c = [ps.base_gamma, ps.motion_unit, ps.total_pins, ps.total_area, ps.pitch, a.min_iter, a.stop_overflow,
         a.settle_iter, a.stall_iter, args.inner_iter, ps.refine, float(a.wa_rule == "accuracy"),
         ps.handover_gamma or 0.0]
    return S, kp, c


def run_level_on_gpu(optimizer, objective, ps, args, log, iteration, check_every=16, use_graph=True,
                     poisson="auto", reorder=True, hook=None, controller=None, phase=0.0, autotune=None,
                     timeline=None, engine=None):
    """Continue the level from `iteration` (density weight on) until it stops; updates optimizer, ps and
    objective.evals as the host loop would. -> (reason, iterations, hpwl, overflow) of the kept placement.
    hook(objective, u, v, g, S, kp, c), if given, sees the state the device loop starts from (instrumentation).
    controller: None (the plain rule) or a dict for the learned controller (controller.py: `attach`), whose "levels"
    list receives this level's record (passes under the controller, fallback, the logged passes).
    autotune: time the density tile layouts at the start and keep the fastest (None: for the cell level when it has
    100k nodes or more -- the long level, where the timing pays for itself); the results do not depend on it.
    timeline: None, or a list that receives (stage, synchronised wall time) marks (instrumentation).
    engine: the level's Engine left by run_start_on_gpu (its point is the optimizer's), continued instead of a new
    one."""
    device, pos = optimizer.v.device, optimizer.v
    log = as_log(log)
    S, kp, c = control_state(ps, optimizer, iteration, args)
    S = torch.tensor(S, dtype=torch.float64, device=device)
    kp = torch.tensor(kp, dtype=torch.float32, device=device)
    c = torch.tensor(c, dtype=torch.float64, device=device)
    u = optimizer.u.contiguous()
    g = optimizer.g.contiguous()
    if hook is not None:
        hook(objective, u, pos, g, S, kp, c)
    best = pos.clone()
    if engine is not None:
        eng = engine
        eng.rebind(u, pos, g, best, S, kp, c)
        S, kp, c = eng.S, eng.kp, eng.c          # the level reads and writes its own control tensors
    else:
        eng = Engine(objective.wirelength, objective.density, u, pos, g, best, S, kp, c, poisson, reorder)
        mark(timeline, "engine set-up", device)
    ctl_log = None
    if controller is not None:
        ctl_log = eng.set_controller(controller, phase)
    if autotune or (autotune is None and phase == 0.0 and objective.density.num_nodes >= 100000):
        eng.autotune(3, online=True)               # timed on the level's first passes: no extra passes
        mark(timeline, "autotune", device)
    evals_before, logged = objective.evals, iteration // args.log_freq
    watch = torch.tensor([K["stop"], K["loop_it"], K["lam"], K["gam"], K["res_hpwl"], K["res_ovf"]], device=device)
    runs_left = (args.inner_iter - iteration) // check_every + 2     # the control stops at inner_iter
    try:
        while True:
            if runs_left == 0:
                raise RuntimeError("device loop: no stop after %d iterations" % args.inner_iter)
            runs_left -= 1
            eng.run(check_every, use_graph)
            stop, it, lam, gam, hpwl, ovf = S[watch].tolist()
            if log and int(it) // args.log_freq > logged:
                logged = int(it) // args.log_freq
                log("iter: %d | hpwl: %.4E overflow: %.4f density_weight: %.4E wa_coeff: %.4E" % (it, hpwl, ovf, lam, gam))
            if stop:
                break
        mark(timeline, "device loop", device)
        eng.finish()
        objective.bucket_spills = eng.bucket_spills()   # density entries past their bucket's room (instrumentation)
        if controller is not None:
            st = eng.controller_state()
            n = int(st[7])
            controller.setdefault("levels", []).append(
                {"phase": phase, "passes": n, "fallback": bool(st[6]),
                 "log": ctl_log[:min(n, len(ctl_log))].cpu().numpy() if ctl_log is not None else None})
    finally:
        torch.cuda.synchronize(device)
        eng.close()
    S = S.tolist()
    get = lambda name: S[K[name]]   # noqa: E731
    ps.density_weight, ps.wa_coeff, ps.iter, ps.motion = get("lam"), get("gam"), int(get("iter")), get("motion")
    ps.weighted_weight, ps.mu, ps.settled = get("ww"), get("mu"), int(get("settled"))
    ps.best_hpwl, ps.best_overflow, ps.improved_iter = get("best_hpwl"), get("best_ovf"), int(get("improved"))
    ps.hpwl.append(get("hpwl_prev"))
    optimizer.u, optimizer.g, optimizer.alpha, optimizer.a = u, g, get("alpha"), get("a")
    objective.evals = evals_before + int(get("evals"))
    if get("has_best"):
        pos.copy_(best)
    return REASON[int(get("stop"))], int(get("loop_it")), get("res_hpwl"), get("res_ovf")


START_WINDOW = 5      # the start rule's window (param_scheduler.py:38-43)


def run_start_on_gpu(optimizer, objective, ps, args, log, iteration, check_every=16, use_graph=True,
                     poisson="auto", reorder=True, timeline=None):
    """The level's wirelength-only start (run_placement.py:41-76 while the density weight is 0) as engine passes
    (ap_level_start_*): the same steps, the BB step from FP64 norms, the start rule of param_scheduler.py:38-48 on
    the device, `check_every` passes per graph replay and one synchronisation per replay instead of one per step.
    Updates optimizer (u, v, g, alpha, a), ps (iter, the hpwl / wa history) and objective.evals as the host loop
    would, and logs every log_freq-th iteration as it does. -> (engine, done, passes): the level's Engine, holding the
    current point, for run_level_on_gpu to continue; done 1 (the rule fired), 2 (diverged) or 3 (the iteration
    limit)."""
    device, pos = optimizer.v.device, optimizer.v
    log = as_log(log)
    if ps.wa or ps.hpwl:
        raise ValueError("run_start_on_gpu starts a level's wirelength-only phase: the scheduler has history")
    S, kp, c = control_state(ps, optimizer, iteration, args)
    S = torch.tensor(S, dtype=torch.float64, device=device)
    kp = torch.tensor(kp, dtype=torch.float32, device=device)
    c = torch.tensor(c, dtype=torch.float64, device=device)
    u = optimizer.u.contiguous()
    g = optimizer.g.contiguous()
    best = pos.clone()
    eng = Engine(objective.wirelength, objective.density, u, pos, g, best, S, kp, c, poisson, reorder)
    mark(timeline, "start engine set-up", device)
    try:
        eng.start_begin(args.start_tol, START_WINDOW, args.inner_iter - iteration)
        runs_left = (args.inner_iter - iteration) // check_every + 2     # the pass limit ends it before
        while True:
            if runs_left == 0:
                raise RuntimeError("start phase: no end after %d passes" % (args.inner_iter - iteration))
            runs_left -= 1
            eng.start_run(check_every, use_graph)
            done, count, _ = eng.start_state()
            if done:
                break
        done, count, hist = eng.start_state(history=True)
        eng.finish()
    except BaseException:
        torch.cuda.synchronize(device)
        eng.close()
        raise
    mark(timeline, "start passes", device)
    St = S.tolist()
    optimizer.u, optimizer.g = u, g
    optimizer.alpha, optimizer.a = St[K["alpha"]], St[K["a"]]
    objective.evals += count
    for k, (wa, hpwl) in enumerate(hist):
        ps.hpwl.append(hpwl)
        ps.wa.append(wa)
        ps.iter += 1
        if log and (iteration + k) % args.log_freq == 0:
            log("iter: %d | hpwl: %.4E overflow: %.4f density_weight: %.4E wa_coeff: %.4E"
                % (iteration + k, hpwl, float("nan"), ps.density_weight, ps.wa_coeff))
    if hist:
        objective.wa = torch.tensor(hist[-1][0], dtype=torch.float64, device=device)
        objective.hpwl = torch.tensor(hist[-1][1], dtype=torch.float64, device=device)
    if done == 2:
        optimizer.diverged = True
    return eng, done, count
