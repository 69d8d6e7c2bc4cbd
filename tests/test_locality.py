"""The statement of the engine's spatial renumbering in torch (tests/renumber_ref.py, CPU tensors): the Hilbert curve
it follows, a permutation that keeps every net and every cell's pins."""
import numpy as np
import pytest
import torch


def test_morton_order_is_a_permutation_keeping_cells_first():
    from renumber_ref import morton_order
    rng = np.random.default_rng(0)
    pos = torch.tensor(rng.uniform(0, 100, (1000, 2)), dtype=torch.float32)
    perm = morton_order(pos, 64, 64, 100 / 64, 100 / 64, 700)
    assert sorted(perm.tolist()) == list(range(1000))
    assert perm[:700].max() < 700 and perm[700:].min() >= 700
    # index neighbours are spatial neighbours: consecutive cells are close on average
    p = pos[perm[:700]]
    assert (p[1:] - p[:-1]).norm(dim=1).mean() < 10


def test_hilbert_key_walks_the_grid_cell_by_cell():
    """The curve the renumbering follows: on a 2^k x 2^k grid the keys are a permutation of 0 .. 4^k - 1 and cells of
    consecutive keys share an edge; it starts at (0, 0) and ends at (2^k - 1, 0); the 2 x 2 grid is visited
    (0,0) (0,1) (1,1) (1,0). A wrong rotation or quadrant order breaks one of these."""
    from renumber_ref import hilbert_key
    for k in (1, 2, 3, 5):
        m = 1 << k
        ix, iy = torch.meshgrid(torch.arange(m), torch.arange(m), indexing="ij")
        key = hilbert_key(ix.reshape(-1), iy.reshape(-1), k)
        assert sorted(key.tolist()) == list(range(m * m))
        at = torch.empty(m * m, 2, dtype=torch.long)
        at[key] = torch.stack([ix.reshape(-1), iy.reshape(-1)], 1)
        assert torch.all((at[1:] - at[:-1]).abs().sum(1) == 1)
        assert at[0].tolist() == [0, 0] and at[-1].tolist() == [m - 1, 0]
        if k == 1:
            assert at.tolist() == [[0, 0], [0, 1], [1, 1], [1, 0]]


@pytest.mark.parametrize("order", ["morton", "hilbert"])
def test_renumbered_netlist_has_the_same_nets(order):
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from renumber_ref import hilbert_order, morton_order, renumber_netlist
    from autoplace_amd.wirelength import net_csr
    lef, d = synth.generate(800, seed=4, io_pins=5)
    data = PlaceData(lef, d)
    n = data.num_mov
    csr = net_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, n, 100)
    t = lambda a: torch.as_tensor(a)   # noqa: E731
    pos = torch.tensor(np.random.default_rng(1).uniform(0, data.width, (n + 50, 2)), dtype=torch.float32)
    perm = (morton_order if order == "morton" else hilbert_order)(pos, 32, 32, data.width / 32, data.height / 32, n)
    new = renumber_netlist(t(csr["cell"]), t(csr["off"]), t(csr["net_start"]), n, perm)
    old_of = torch.cat([perm[:n], torch.tensor([n])])

    def nets(cell, off, start):
        out = set()
        for i in range(len(start) - 1):
            s, e = int(start[i]), int(start[i + 1])
            out.add(tuple((int(cell[k]), tuple(np.asarray(off[k]).tolist())) for k in range(s, e)))
        return out

    before = nets(csr["cell"], csr["off"], csr["net_start"])
    after = nets(old_of[new["cell32"].long()].numpy(), new["off"].numpy(), new["net_start"].numpy())
    assert before == after
    # cell CSR of the new numbering: every movable pin once, grouped by cell
    cs, cp = new["cell_start"].numpy(), new["cell_pin"].numpy()
    assert sorted(cp.tolist()) == sorted(np.nonzero(new["cell32"].numpy() < n)[0].tolist())
    for c in range(0, n, 37):
        assert np.all(new["cell32"].numpy()[cp[cs[c]:cs[c + 1]]] == c)


def test_device_net_csr_equals_the_host_one():
    """The torch form of the optimized-net CSR (built on the device for the coarse netlist) gives the host form's
    arrays bit for bit, here on CPU tensors; nets of degree 1 and above the limit are dropped by both."""
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    from autoplace_amd.wirelength import net_csr, net_csr_torch
    lef, d = synth.generate(3000, seed=7, io_pins=6)
    data = PlaceData(lef, d)
    n = data.num_mov
    for limit in (100, 6):
        ref = net_csr(data.pin_cell, data.pin_off, data.pin_net, data.num_nets, n, limit)
        got = net_csr_torch(torch.as_tensor(np.asarray(data.pin_cell)), torch.as_tensor(np.asarray(data.pin_off)),
                            torch.as_tensor(np.asarray(data.pin_net)), data.num_nets, n, limit)
        assert got["num_nets"] == ref["num_nets"]
        for k in ("cell", "off", "net_start", "cell_start", "cell_pin"):
            a, b = np.asarray(ref[k]), got[k].numpy()
            assert a.dtype == b.dtype and np.array_equal(a, b), k
