"""Clustering and the coarse problem on the device against the reference's own outputs (golden_small.npz: cluster ids
and areas, cluster sizes, the clusters' pins, the coarse grid, the expansion) bit for bit; the free-area fraction on
the device against the host form; a density layer built from device tensors against one built from host arrays."""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def small(small_design):
    from autoplace_amd.problem import build_problem, default_args
    from autoplace_amd.run_placement import build_operators
    _, _, data = small_design
    args = default_args()
    prob = build_problem(data, args)
    wl, den = build_operators(prob, args, DEV)
    return data, args, prob, wl, den


def test_cluster_ids_and_areas_match_the_reference(small, golden_small):
    from autoplace_amd.cluster import AREA_SCALE, cluster_cells
    data, args, prob, wl, den = small
    g = golden_small
    size = torch.tensor(data.size, dtype=torch.float32, device=DEV)
    area = (size.prod(1).double() * AREA_SCALE).round().long()
    cid, carea = cluster_cells(wl.all["cell"], wl.all["net_start"], area, 5.0, 0)
    assert np.array_equal(cid.cpu().numpy(), g["cluster_id"])
    assert np.array_equal(carea.cpu().numpy(), g["cluster_area"])


def test_coarse_problem_matches_the_reference(small, golden_small):
    from autoplace_amd.cluster import coarse_problem, expand
    data, args, prob, wl, den = small
    g = golden_small
    n = data.num_mov
    coarse = coarse_problem(data, prob.node_size, wl, den, prob.fixed_fraction, args)
    assert coarse.num == int(g["coarse_num"])
    assert np.array_equal(coarse.cid.cpu().numpy(), g["cluster_id"])
    assert np.array_equal(coarse.size.cpu().numpy(), g["coarse_size"])
    assert coarse.density.nbx == int(g["coarse_nb"])
    assert coarse.total_pins == float(g["coarse_total_pins"])
    pins = coarse.wirelength.pins()
    assert np.array_equal(pins["cell"].cpu().numpy(), g["coarse_pin_cell"])
    assert np.array_equal(pins["net"].cpu().numpy(), g["coarse_pin_net"])
    assert np.array_equal(pins["off"].cpu().numpy(), g["coarse_pin_off"])
    pr = torch.tensor(g["pos_rand"], device=DEV)
    cpos = torch.cat([pr[:coarse.num], pr[n:]])
    assert np.array_equal(expand(cpos, coarse, prob.node_size, 0).cpu().numpy(), g["expand_pr"])
    # the same with the draws made beforehand on a host thread, as the flow makes them
    from autoplace_amd.cluster import ExpansionDraws
    draws = ExpansionDraws(n, 0)
    assert np.array_equal(expand(cpos, coarse, prob.node_size, 0, draws).cpu().numpy(), g["expand_pr"])


def test_device_density_layer_equals_the_host_built_one(small):
    """The coarse layer's arrays when built from device tensors equal those of a layer built from host arrays."""
    from autoplace_amd.density import ElectronicDensityLayer
    data, args, prob, wl, den = small
    rng = np.random.default_rng(2)
    node = np.abs(rng.normal(3.0, 2.0, (700, 2))) + 0.3
    frac = rng.random((32, 32)) * 0.4
    pins = rng.integers(0, 9, 700).astype(np.float64)
    host = ElectronicDensityLayer(data.width, data.height, 32, 32, frac, 0.9, node, 500, DEV, pins)
    dev = ElectronicDensityLayer(data.width, data.height, 32, 32, torch.tensor(frac, device=DEV), 0.9,
                                 torch.tensor(node, device=DEV), 500, DEV, torch.tensor(pins, device=DEV))
    for a, b in ((host.smooth["size"], dev.smooth["size"]), (host.smooth["ratio"], dev.smooth["ratio"]),
                 (host.real["size"], dev.real["size"]), (host.fixed_map, dev.fixed_map), (host.pos_lo, dev.pos_lo),
                 (host.pos_hi, dev.pos_hi), (host.pins, dev.pins), (host.area, dev.area)):
        assert torch.equal(a, b)
    assert rel_err(dev.mov_area, host.mov_area) < 1e-14


@pytest.mark.parametrize("config", ["small", "B"])
def test_free_fraction_on_the_device(small, config):
    """The device's free-area fraction against the host form (its sums run in another order) at several grids, on
    the small design (2 macros, design pins) and on the 200k-cell design (8 macros on 20 %)."""
    if config == "small":
        data = small[0]
    else:
        from autoplace_amd import synth
        from autoplace_amd.database import PlaceData
        lef, d = synth.generate_config("B", seed=0)
        data = PlaceData(lef, d)
    worst = 0.0
    for nbx, nby in ((16, 16), (48, 40), (64, 64), (512, 512), (1024, 1024), (100, 300)):
        host = data.fixed_fraction(nbx, nby)
        dev = data.fixed_fraction_on(DEV, nbx, nby).cpu().numpy()
        worst = max(worst, float(np.abs(dev - host).max()))
        assert np.abs(dev - host).max() < 1e-12, (nbx, nby)
        assert 0.0 < host.mean() < 1.0
    print("free fraction, device vs host (%s): max |diff| %.2e" % (config, worst))
