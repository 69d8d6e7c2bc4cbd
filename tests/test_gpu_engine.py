"""The native GP iteration engine on the GPU: its objective against the oracle, graph replay against direct launches
(bit for bit), determinism, and a short trajectory against the oracle's control loop."""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from conftest import norm_err, rel_err
from oracle import control as octl
from oracle import place as opl

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _setup(data, weight=2e-3, gamma=30.0, seed=1, grids=None):
    from autoplace_amd import gpu_loop
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import _scheduler, build_operators
    args = default_args() if grids is None else default_args(num_bin=grids[0], force_num_bin=grids[1])
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV, "gemm")
    rng = np.random.default_rng(seed)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    ps = _scheduler(den, data.num_mov, data.mov_area, float(data.cell_pins.sum()), float(prob.node_size.prod(1).sum()),
                    len(prob.node_size), args)
    ps.density_weight, ps.wa_coeff, ps.iter = weight, gamma, 40
    ps.hpwl.append(1e30)
    opt = SimpleNamespace(alpha=0.05, a=1.0)
    S, kp, c = gpu_loop.control_state(ps, opt, 40, args)
    return args, prob, wl, den, pos, S, kp, c


def _oracle_level(prob, args, data):
    n = data.num_mov
    judge = None if prob.force_bin == prob.num_bin else {"nb": prob.num_bin, "fraction": prob.judge_fraction,
                                                          "size": prob.node_size[:n]}
    return opl.Level(prob.width, prob.height, prob.force_bin, prob.fixed_fraction, prob.target, prob.node_size, n,
                     prob.node_pins, prob.pin_cell, prob.pin_off, prob.pin_net, prob.num_nets, args.ignore_net_degree,
                     judge)


def _engine(wl, den, pos, S, kp, c, g0, reorder=True, poisson="gemm"):
    from autoplace_amd.gpu_loop import Engine
    t = lambda a, dt=torch.float32: torch.tensor(np.asarray(a), dtype=dt, device=DEV)   # noqa: E731
    u, v, g, best = t(pos), t(pos), t(g0), t(pos)
    dS, dkp, dc = t(S, torch.float64), t(kp), t(c, torch.float64)
    return Engine(wl, den, u, v, g, best, dS, dkp, dc, poisson, reorder), (u, v, g, best, dS, dkp, dc)


def test_engine_objective_vs_oracle(small_design):
    _, _, data = small_design
    args, prob, wl, den, pos, S, kp, c = _setup(data)
    eng, (u, v, g, best, dS, dkp, dc) = _engine(wl, den, pos, S, kp, c, np.zeros_like(pos), reorder=False)
    grad = torch.empty_like(v)
    stats = torch.zeros(8, dtype=torch.float64, device=DEV)
    eng.objective(v, grad, stats)
    torch.cuda.synchronize()
    lvl = _oracle_level(prob, args, data)
    kp32 = np.array(kp, np.float32)
    g_ref, wa, hpwl, ovf = opl._evaluate_kp(lvl, pos, kp32)
    st = stats.cpu().numpy()
    print("engine objective: grad err %.2e, wa %.2e, hpwl %.2e, ovf %.2e" % (
        norm_err(grad.cpu().numpy(), g_ref), rel_err(st[0], wa), rel_err(st[1], hpwl), rel_err(st[2], ovf)))
    assert norm_err(grad.cpu().numpy(), g_ref) < 1e-5
    assert rel_err(st[0], wa) < 1e-6 and rel_err(st[1], hpwl) < 1e-6 and rel_err(st[2], ovf) < 1e-6
    eng.close()


@pytest.mark.parametrize("case", ["small", "grid1024"])
def test_engine_density_bit_exact(small_design, case):
    """The engine's tile-binned density map equals the oracle's fixed-point map bit for bit, and its overflow the
    oracle's: on the small design (16 x 16-bin tiles) and on a 1024 x 1024 force grid (32 x 32-bin tiles) where
    nodes span several tiles and the biggest go through the big-node path; the same for every tile layout."""
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators
    from oracle import density as oden
    if case == "small":
        data = small_design[2]
        args = default_args()
    else:
        lef, d = synth.generate(6000, num_macros=3, macro_frac=0.1, seed=11, io_pins=8, name="g1024")
        data = PlaceData(lef, d)
        args = default_args(force_num_bin=1024)
    prob = build_problem(data, args)
    if case == "grid1024":
        prob.node_size = prob.node_size.copy()
        prob.node_size[:40] *= 6.0                   # 40 cells over 2 tiles wide: the big-node path
    wl, den = build_operators(prob, args, DEV, "fft")
    rng = np.random.default_rng(5)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = (lo + rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    pos[: len(pos) // 3] = (lo[: len(pos) // 3] + 0.5 * (hi - lo)[: len(pos) // 3] * rng.random((len(pos) // 3, 2)))
    S = np.zeros(21)
    kp = np.array([30.0, 1e-3, 1.0, 0.0, 0.0], np.float32)
    c = np.zeros(13)
    eng, _ = _engine(wl, den, pos, S, kp, c, np.zeros_like(pos), reorder=True, poisson="fft")
    grad = torch.empty_like(eng.v)
    stats = torch.zeros(8, dtype=torch.float64, device=DEV)

    def evaluate():
        stats.zero_()
        eng.objective(eng.v, grad, stats)
        return eng.density().cpu().numpy(), float(stats[2])

    rho_dev, ovf_dev = evaluate()
    # every tile layout (side 8 .. 64, 1 .. 8 tiles per workgroup) each grid admits gives the same map and overflow:
    # the force grid's with the overflow grid at (8, 1), then the overflow grid's with the force grid at (32, 1)
    layouts = [(t, g) for t in (8, 16, 32, 64) for g in (1, 2, 4, 8)]
    pairs = [(fl, (8, 1) if den.judge is not None else (0, 0)) for fl in layouts]
    if den.judge is not None:
        pairs += [((32, 1), jl) for jl in layouts]
    tried = []
    for fl, jl in pairs:
        try:
            eng.set_tile_layout(fl, jl)
        except ValueError:                           # a side or group the grid does not admit
            continue
        r, o = evaluate()
        assert np.array_equal(r, rho_dev), (fl, jl)
        assert rel_err(o, ovf_dev) < 1e-12, (fl, jl, o, ovf_dev)   # FP64 sums of per-tile partials
        tried.append((fl, jl))
    print("tile layouts with the same map and overflow: %s" % tried)
    assert len(tried) >= 6
    # rooms cut short (0 %: every entry; 50 %: about half of them) send entries down the spill path: the same map
    spills = []
    for room in ((0, 0), (50, 0), (100, 0), (150, 64)):
        eng.set_bucket_room(*room)
        for _ in range(2):                           # rooms from the counts of the evaluation before
            before = eng.bucket_spills()
            r, o = evaluate()
            assert np.array_equal(r, rho_dev), room
            assert rel_err(o, ovf_dev) < 1e-12, (room, o, ovf_dev)
            spills.append(sum(eng.bucket_spills()) - sum(before))
    print("spilled entries per evaluation at rooms 0 %%, 50 %%, 100 %%, default: %s" % spills)
    assert min(spills[:4]) > 0                       # 0 % and 50 %
    assert max(spills[4:]) == 0                      # rooms of at least the counts at the same positions
    eng.close()
    smooth, ratio = den.smooth["size"].cpu().numpy(), den.smooth["ratio"].cpu().numpy()
    bx, by = float(den.bin_np[0]), float(den.bin_np[1])
    m0, m1 = oden.density_maps(pos, smooth, ratio, data.num_mov, bx, by, den.nbx, den.nby)
    rho, over = oden.density_post(m0, m1, den.fixed_map.cpu().numpy(), den.target, den.nbx, den.nby)
    if den.judge is not None:
        j = den.judge
        js, jr = j.smooth["size"].cpu().numpy(), j.smooth["ratio"].cpu().numpy()
        q0, q1 = oden.density_maps(pos[:data.num_mov], js, jr, data.num_mov, float(j.bin_np[0]), float(j.bin_np[1]),
                                   j.nbx, j.nby)
        _, over = oden.density_post(q0, q1, j.fixed_map.cpu().numpy(), j.target, j.nbx, j.nby)
        # the engine divides by the force layer's movable area (equal to the overflow layer's whenever the cells'
        # sizes are the same in both, as in every run; the enlarged cells above make them differ here)
        ovf = oden.overflow(over, j.bin_area, den.mov_area)
    else:
        ovf = oden.overflow(over, den.bin_area, den.mov_area)
    if case == "grid1024":
        span = smooth / np.array([bx, by], np.float32)
        print("1024 grid: nodes over 33 bins: %d of %d" % (int((span.max(1) > 33).sum()), len(span)))
        assert (span.max(1) > 33).sum() >= 40
    print("engine density (%s, %d x %d): max |diff| %.3g, overflow rel %.2e"
          % (case, den.nbx, den.nby, float(np.abs(rho_dev - rho).max()), rel_err(ovf_dev, ovf)))
    assert np.array_equal(rho_dev, rho)
    assert rel_err(ovf_dev, ovf) < 1e-12


@pytest.mark.parametrize("spread", [1.0, 0.04])
def test_split_tiles_change_nothing(spread):
    """Tiles whose buckets are summed by several workgroups (a slice of the entries each, the sums met in the spill
    map, the last workgroup writing the tile) give the oracle's map bit for bit and its overflow: every tile split
    (slices of 64 and 512 entries without the mean condition), and the level's start case, all nodes in a corner of
    4 % of the die's side, at the defaults."""
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators
    from oracle import density as oden
    lef, d = synth.generate(30000, num_macros=3, macro_frac=0.1, seed=12, io_pins=8, name="split")
    data = PlaceData(lef, d)
    args = default_args(force_num_bin=256)
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV, "fft")
    rng = np.random.default_rng(6)
    lo, hi = den.pos_lo.cpu().numpy(), den.pos_hi.cpu().numpy()
    pos = (lo + spread * rng.random(lo.shape) * (hi - lo)).astype(np.float32)
    S = np.zeros(21)
    kp = np.array([30.0, 1e-3, 1.0, 0.0, 0.0], np.float32)
    c = np.zeros(13)
    eng, _ = _engine(wl, den, pos, S, kp, c, np.zeros_like(pos), reorder=True, poisson="fft")
    grad = torch.empty_like(eng.v)
    stats = torch.zeros(8, dtype=torch.float64, device=DEV)

    def evaluate():
        stats.zero_()
        eng.objective(eng.v, grad, stats)
        return eng.density().cpu().numpy(), float(stats[2])

    smooth, ratio = den.smooth["size"].cpu().numpy(), den.smooth["ratio"].cpu().numpy()
    bx, by = float(den.bin_np[0]), float(den.bin_np[1])
    m0, m1 = oden.density_maps(pos, smooth, ratio, data.num_mov, bx, by, den.nbx, den.nby)
    rho, _ = oden.density_post(m0, m1, den.fixed_map.cpu().numpy(), den.target, den.nbx, den.nby)
    # (slice, mean factor, whether tiles must split): spread out, the tiles hold a few hundred entries each and split
    # only without the mean condition; in the corner, at the defaults
    slices = ((64, 0, True), (512, 0, False), (64, 8, False)) if spread == 1.0 else ((4096, 4, True), (64, 0, True))
    rho0, ovf0 = evaluate()
    assert np.array_equal(rho0, rho)
    for sl, factor, must in slices:
        eng.set_tile_split(sl, factor)
        for _ in range(2):                           # the split plan comes from the evaluation before
            before = eng.tile_splits()
            r, o = evaluate()
            after = eng.tile_splits()
            print("slice %d (spread %.2f): split tiles %s, map equal %s" % (sl, spread, (after[0] - before[0],
                  after[1] - before[1]), np.array_equal(r, rho)), flush=True)
            assert np.array_equal(r, rho), sl
            assert rel_err(o, ovf0) < 1e-12, (sl, o, ovf0)
        assert sum(after) > sum(before) or not must  # tiles were split
    eng.set_tile_split(0, 0)
    r, o = evaluate()
    assert np.array_equal(r, rho) and rel_err(o, ovf0) < 1e-12
    eng.close()


def _run(data, passes, use_graph, calls=1, reorder=True, poisson="gemm", time_cls=None, profile=False, grids=None,
         room=None, spills=None, every=1, shape=None):
    args, prob, wl, den, pos, S, kp, c = _setup(data, grids=grids)
    g0 = np.zeros_like(pos)
    eng, bufs = _engine(wl, den, pos, S, kp, c, g0, reorder=False, poisson=poisson)
    grad = torch.empty_like(bufs[1])
    stats = torch.zeros(8, dtype=torch.float64, device=DEV)
    eng.objective(bufs[1], grad, stats)            # g at the start point (caller's order)
    torch.cuda.synchronize()
    eng.close()
    g0 = grad.cpu().numpy()
    eng, bufs = _engine(wl, den, pos, S, kp, c, g0, reorder=reorder, poisson=poisson)
    if room:
        eng.set_bucket_room(*room)
    if shape:
        eng.set_launch_shape(*shape)
        assert eng.launch_shape() == tuple(shape)
    if time_cls:
        eng.time_every(every)
        eng.time_class(time_cls)
    for _ in range(calls):
        if profile:
            eng.profile(passes)
        else:
            eng.run(passes, use_graph)
    timing = eng.class_ms() if time_cls else None
    eng.finish()
    torch.cuda.synchronize()
    out = [b.cpu().numpy().copy() for b in bufs]
    if spills is not None:
        spills.append(eng.bucket_spills())
    eng.close()
    return out, (args, prob, pos, S, kp, c, grad.cpu().numpy(), timing)


def test_engine_graph_equals_direct_and_is_deterministic(small_design):
    _, _, data = small_design
    a, _ = _run(data, 16, False, calls=2)
    b, _ = _run(data, 16, True, calls=2)
    c, _ = _run(data, 16, True, calls=2)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(y, z)
    # without the spatial renumbering the same passes give the same control decisions (FP64 sums in another order)
    d, _ = _run(data, 16, True, calls=2, reorder=False)
    K = octl.K
    for name in ("iter", "loop_it", "evals", "stop"):
        assert d[4][K[name]] == a[4][K[name]]
    assert norm_err(d[1], a[1]) < 1e-4
    K = octl.K
    assert a[4][K["loop_it"]] == 40 + 32 and a[4][K["evals"]] == 32


def test_cached_device_memory_changes_nothing(small_design):
    """An engine built on the blocks a destroyed one left in the library's device-memory cache (their old contents
    still in them) runs its passes bit for bit as one built on blocks fresh from HIP."""
    from autoplace_amd import _lib
    lib = _lib.lib()
    _, _, data = small_design
    lib.ap_release_cached_memory()
    a, _ = _run(data, 16, True, calls=2)
    kept = lib.ap_cached_device_bytes()
    assert kept > 0
    b, _ = _run(data, 16, True, calls=2)
    assert lib.ap_cached_device_bytes() == kept          # the second engine took the same blocks back
    assert lib.ap_release_cached_memory() == 0 and lib.ap_cached_device_bytes() == 0
    c, _ = _run(data, 16, True, calls=2)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(y, z)


@pytest.mark.parametrize("grids", [None, (32, 64)])
def test_engine_two_streams_equal_one_stream(small_design, grids):
    """Graph and direct passes, which run WA and the overflow grid on a second stream beside the force grid, end bit
    for bit where profiled passes end, which run every kernel in order on one stream (with an overflow grid of its
    own: force grid 64, overflow grid 32)."""
    _, _, data = small_design
    one, (_, prob, _, _, _, _, _, _) = _run(data, 16, False, profile=True, grids=grids)
    print("force grid %d, overflow grid %d" % (prob.force_bin, prob.num_bin))
    if grids:
        assert prob.force_bin != prob.num_bin
    for use_graph in (True, False):
        two, _ = _run(data, 16, use_graph, grids=grids)
        for x, y in zip(one, two):
            assert np.array_equal(x, y)


def test_launch_shapes_change_nothing(small_design):
    """The WA and gather launch shapes the autotuner chooses from (WA: 1, 2 or 4 plan wavefronts per launched
    wavefront, 64 .. 512 lanes; gather: 64 .. 512 lanes) leave 16 passes bit for bit as they are: their partials are
    per wavefront, summed in a fixed order."""
    _, _, data = small_design
    ref, _ = _run(data, 16, True, grids=(32, 64))
    for shape in ((1, 128, 128), (4, 512, 512), (1, 64, 64), (4, 256, 128), (2, 512, 256)):
        for use_graph in (True, False):
            out, _ = _run(data, 16, use_graph, grids=(32, 64), shape=shape)
            for x, y in zip(out, ref):
                assert np.array_equal(x, y), shape


@pytest.mark.parametrize("room", [(0, 0), (50, 0)])
def test_engine_spilled_entries_change_nothing(small_design, room):
    """Passes whose density buckets are given too little room (every entry, or about half of them, past it: the
    spill path) end bit for bit where passes with the default rooms end, graph and direct."""
    _, _, data = small_design
    spills = []
    ref, _ = _run(data, 16, True, grids=(32, 64), spills=spills)
    for use_graph in (True, False):
        out, _ = _run(data, 16, use_graph, grids=(32, 64), room=room, spills=spills)
        for x, y in zip(out, ref):
            assert np.array_equal(x, y)
    print("spilled entries (force, overflow) at the default rooms, then at %s %% + %d: %s" % (room + (spills,)))
    # the default rooms are set from counts at the current point (the set-up count included): small moves do not
    # outgrow them
    assert spills[0] == (0, 0)
    assert min(spills[1]) > 0 and min(spills[2]) > 0


def test_engine_timing_events_change_nothing(small_design):
    """Event-record nodes around a kernel class (captured graph) and event records (direct launches) leave the results
    bit for bit as they are, and time the class."""
    _, _, data = small_design
    ref, _ = _run(data, 16, True)
    for use_graph in (True, False):
        for cls in ("density_tiles", "wa", "gather"):
            out, (_, _, _, _, _, _, _, timing) = _run(data, 16, use_graph, time_cls=cls)
            ms, n = timing
            print("timed %s (%s): %.4f ms per pass over %d passes" % (cls, "graph" if use_graph else "direct", ms, n))
            assert n == 16 and 0 < ms < 50
            for x, y in zip(out, ref):
                assert np.array_equal(x, y)
        # events in every 5th pass only: passes 0, 5, 10, 15 timed
        out, (_, _, _, _, _, _, _, (ms, n)) = _run(data, 16, use_graph, time_cls="density_tiles", every=5)
        assert n == 4 and 0 < ms < 50
        for x, y in zip(out, ref):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("poisson", ["gemm", "fft"])
def test_engine_trajectory_vs_oracle(small_design, poisson):
    """16 passes of the engine against 16 oracle passes from the same state: the control decisions agree and the
    points stay within FP32 drift of each other."""
    _, _, data = small_design
    (u, v, g, best, S_dev, kp_dev, _), (args, prob, pos, S, kp, c, g0, _) = _run(data, 16, True, poisson=poisson)
    lvl = _oracle_level(prob, args, data)
    S = list(S)
    kp = np.array(kp, np.float32)
    uo, vo, go = pos.copy(), pos.copy(), g0.copy()
    K = octl.K
    for _ in range(16):
        un, vn = octl.nesterov_step(uo, vo, go, lvl.pos_lo, lvl.pos_hi, kp[3], kp[4])
        gn, wa, hpwl, ovf = opl._evaluate_kp(lvl, vn, kp)
        octl.control_pass(S, kp, c, octl.dist(vn, vo), octl.dist(gn, go), hpwl, ovf, wa)
        if S[K["accept"]]:
            uo, vo, go = un, vn, gn
    for name in ("iter", "loop_it", "evals", "settled", "stop"):
        assert S_dev[K[name]] == S[K[name]], name
    for name in ("lam", "gam", "alpha", "hpwl_prev"):
        assert rel_err(S_dev[K[name]], S[K[name]]) < 1e-4, (name, S_dev[K[name]], S[K[name]])
    print("engine vs oracle after 16 passes: v err %.2e" % norm_err(v, vo))
    assert norm_err(v, vo) < 1e-4


def test_graph_cache_eviction_without_host_syncs(small_design):
    """More distinct captured runs than the level keeps (4 graphs: pass counts 1..6, each with another timed kernel
    class) replayed back to back with no host synchronisation between them: every eviction destroys a graph whose
    replay may still be running, which the engine waits for first. The state ends bit for bit where direct launches
    of the same passes end."""
    _, _, data = small_design
    args, prob, wl, den, pos, S, kp, c = _setup(data)
    classes = ("wa", "density_tiles", "gather", None, "nesterov", "overflow_tiles")
    outs = []
    for use_graph in (True, False):
        eng, bufs = _engine(wl, den, pos, S, kp, c, np.zeros_like(pos), reorder=True, poisson="fft")
        for passes, cls in zip(range(1, 7), classes):
            eng.time_class(cls)
            eng.run(passes, use_graph)
        eng.run(3, use_graph)                        # a cached graph again (3 passes, timing class of the last run)
        eng.finish()
        torch.cuda.synchronize()
        outs.append([b.cpu().numpy().copy() for b in bufs])
        eng.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
    assert outs[0][4][octl.K["evals"]] == 24
