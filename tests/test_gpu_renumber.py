"""The library's spatial renumbering (ap_level_renumber) against its torch statement (tests/renumber_ref.py) on the
same device tensors: the permutation, the renumbered netlist and the cells' pin lists bit for bit, on a generated
design and on edge cases (nodes off the die and on bin edges, no fillers, a non-square and a one-bin grid, nets
without pins, a net of many pins)."""
import numpy as np
import pytest
import torch

from renumber_ref import hilbert_order, renumber_netlist

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _check(pos, nbx, nby, bx, by, n, cell32, off, net_start):
    from autoplace_amd.locality import renumber_level
    print("renumber: %d nodes, %d cells, %d pins, %d nets, grid %d x %d" % (pos.shape[0], n, cell32.numel(),
                                                                            net_start.numel() - 1, nbx, nby),
          flush=True)
    perm, got = renumber_level(pos, nbx, nby, bx, by, n, cell32, off, net_start)
    ref_perm = hilbert_order(pos, nbx, nby, bx, by, n)
    ref = renumber_netlist(cell32, off, net_start, n, ref_perm)
    assert torch.equal(perm.cpu(), ref_perm.cpu())
    for k in ("cell32", "net_start", "cell_start"):
        assert got[k].dtype == ref[k].dtype and torch.equal(got[k].cpu(), ref[k].cpu()), k
    assert torch.equal(got["off"].cpu(), ref["off"].cpu())
    m = ref["cell_pin"].numel()
    assert torch.equal(got["cell_pin"][:m].cpu(), ref["cell_pin"].cpu())
    assert got["num_nets"] == ref["num_nets"]
    # the entries past the movable cells' are the fixed pins, each once
    fixed = torch.nonzero(got["cell32"] == n).flatten().int()
    assert torch.equal(torch.sort(got["cell_pin"][m:cell32.numel()]).values.cpu(), fixed.cpu())
    return perm, got


def _netlist(rng, n, nets, max_deg, fixed_frac=0.05, empty=0):
    deg = rng.integers(2, max_deg + 1, nets)
    if empty:
        deg[rng.choice(nets, empty, replace=False)] = 0
    start = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    cell = rng.integers(0, n, int(start[-1])).astype(np.int32)
    cell[rng.random(cell.size) < fixed_frac] = n
    off = rng.normal(0, 0.5, (cell.size, 2)).astype(np.float32)
    t = lambda a: torch.tensor(a, device=DEV)   # noqa: E731
    return t(cell), t(off), t(start)


def test_renumber_of_a_generated_design():
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.wirelength import net_csr
    lef, d = synth.generate(20000, seed=3, io_pins=40)
    data = PlaceData(lef, d)
    n = data.num_mov
    csr = net_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, n, 100)
    rng = np.random.default_rng(5)
    pos = torch.tensor(rng.uniform(0, [data.width, data.height], (n + 3000, 2)), dtype=torch.float32, device=DEV)
    t = lambda a: torch.as_tensor(np.asarray(a)).to(DEV).contiguous()   # noqa: E731
    nb = 128
    _check(pos, nb, nb, data.width / nb, data.height / nb, n, t(csr["cell"]), t(csr["off"]), t(csr["net_start"]))


@pytest.mark.parametrize("case", ["off_die_and_edges", "no_fillers", "non_square", "one_bin", "empty_nets",
                                  "big_net"])
def test_renumber_edge_cases(case):
    rng = np.random.default_rng(11)
    n, fillers, nbx, nby, bx, by = 3000, 500, 64, 64, 1.7, 1.3
    if case == "no_fillers":
        fillers = 0
    if case == "non_square":
        nbx, nby = 100, 7
    if case == "one_bin":
        nbx = nby = 1
    N = n + fillers
    pos = rng.uniform(0, [nbx * bx, nby * by], (N, 2)).astype(np.float32)
    if case == "off_die_and_edges":
        pos[:200] = rng.uniform(-50, 50 + max(nbx * bx, nby * by), (200, 2))
        k = rng.integers(0, nbx, 300)
        pos[200:500, 0] = (k * np.float32(bx)).astype(np.float32)       # on (or one rounding off) bin edges
        pos[200:500, 1] = (rng.integers(0, nby, 300) * np.float32(by)).astype(np.float32)
    nets, max_deg, empty = 4000, 8, 0
    if case == "empty_nets":
        empty = 300
    cell, off, start = _netlist(rng, n, nets, max_deg, empty=empty)
    if case == "big_net":
        big = torch.tensor(rng.integers(0, n + 1, 2500).astype(np.int32), device=DEV)
        cell = torch.cat([cell, big])
        off = torch.cat([off, torch.zeros(2500, 2, device=DEV)])
        start = torch.cat([start, start[-1:] + 2500])
    _check(torch.tensor(pos, device=DEV), nbx, nby, bx, by, n, cell.contiguous(), off.contiguous(),
           start.contiguous())
