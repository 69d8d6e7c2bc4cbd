"""Weighted-average (WA) wirelength on the device: value, gradient and HPWL in one pass
(src/core/wa_wirelength_hpwl.py:16-94 of the reference, on the HIP kernels).

For a net with pin coordinates x_i and smoothing gamma:
    WA = sum(x_i e^{x_i/gamma}) / sum(e^{x_i/gamma}) - sum(x_i e^{-x_i/gamma}) / sum(e^{-x_i/gamma})
shifted by the net's max / min (which give the HPWL for free). Both axes in one pass.
"""
import numpy as np
import torch

from . import gp_cuda


def net_csr(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree):
    """Optimized nets (2 <= degree <= ignore_net_degree) renumbered, grouped by net, and the cell CSR of their movable
    pins (wa_wirelength_hpwl.py:26-32,38-44). Host numpy -> dict of arrays (int32 indices, float32 offsets)."""
    pin_cell = np.asarray(pin_cell, dtype=np.int64)
    pin_net = np.asarray(pin_net, dtype=np.int64)
    degree = np.bincount(pin_net, minlength=num_nets)
    use = (degree >= 2) & (degree <= ignore_net_degree)
    new_id = np.cumsum(use.astype(np.int64)) - 1
    sel = use[pin_net]
    cell, net = pin_cell[sel], new_id[pin_net[sel]]
    if len(net) > 1 and np.any(np.diff(net) < 0):
        raise ValueError("the pins must be grouped by net")
    n_opt = int(use.sum())
    order = np.argsort(cell, kind="stable")
    cell_start = np.searchsorted(cell[order], np.arange(num_mov + 1))
    return {"cell": cell.astype(np.int32), "off": np.ascontiguousarray(np.asarray(pin_off, np.float32)[sel]),
            "net_start": np.searchsorted(net, np.arange(n_opt + 1)).astype(np.int32),
            "cell_start": cell_start.astype(np.int32), "cell_pin": order[:int(cell_start[-1])].astype(np.int32),
            "num_nets": n_opt}


def net_csr_torch(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree):
    """net_csr with torch ops on the tensors' device (the coarse netlist, built inside the GP clock): the same integer
    steps (degree count, renumbering, a stable sort by cell), so the same arrays bit for bit. -> dict of tensors."""
    pin_cell, pin_net = pin_cell.long(), pin_net.long()
    dev = pin_cell.device
    degree = torch.bincount(pin_net, minlength=num_nets)
    use = (degree >= 2) & (degree <= ignore_net_degree)
    new_id = torch.cumsum(use.long(), 0) - 1
    sel = use[pin_net]
    cell, net = pin_cell[sel], new_id[pin_net[sel]]
    if len(net) > 1 and bool((net[1:] < net[:-1]).any()):
        raise ValueError("the pins must be grouped by net")
    n_opt = int(use.sum())
    order = torch.argsort(cell, stable=True)
    cell_start = torch.searchsorted(cell[order], torch.arange(num_mov + 1, device=dev))
    return {"cell": cell.int(), "off": pin_off.float()[sel].contiguous(),
            "net_start": torch.searchsorted(net, torch.arange(n_opt + 1, device=dev)).int(),
            "cell_start": cell_start.int(), "cell_pin": order[:int(cell_start[-1])].int(), "num_nets": n_opt}


class Wirelength(object):
    def __init__(self, pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree, device):
        """pin_cell / pin_net int [P] grouped by net (num_mov = fixed pin, its absolute position in pin_off),
        pin_off [P, 2]: host arrays, or tensors on the device (then the CSR is built there). Two pin sets: `all`
        (every net: exact HPWL, clustering) and `opt` (the optimized nets) as the kernels' CSR view."""
        self.num_mov, self.device = num_mov, torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the wirelength kernels run on a HIP device (got %s)" % self.device)
        if torch.is_tensor(pin_cell):
            self._from_tensors(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree)
            return
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)   # noqa: E731
        csr = net_csr(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree)
        self.num_opt_nets = csr["num_nets"]
        self.opt = {"cell32": t(csr["cell"], torch.int32), "off": t(csr["off"], torch.float32),
                    "net_start": t(csr["net_start"], torch.int32), "cell_start": t(csr["cell_start"], torch.int32),
                    "cell_pin": t(csr["cell_pin"], torch.int32), "num_nets": csr["num_nets"]}
        pin_net = np.asarray(pin_net, dtype=np.int64)
        self.all = {"cell": t(pin_cell, torch.int32), "off": t(np.asarray(pin_off, np.float32), torch.float32),
                    "net": t(pin_net, torch.int32), "num_nets": num_nets,
                    "net_start": t(np.searchsorted(pin_net, np.arange(num_nets + 1)), torch.int32)}
        self._gamma = torch.empty(1, device=self.device)

    def _from_tensors(self, pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree):
        dev = self.device
        pin_cell, pin_off, pin_net = pin_cell.to(dev), pin_off.to(dev), pin_net.to(dev)
        csr = net_csr_torch(pin_cell, pin_off, pin_net, num_nets, num_mov, ignore_net_degree)
        self.num_opt_nets = csr["num_nets"]
        self.opt = {"cell32": csr["cell"], "off": csr["off"], "net_start": csr["net_start"],
                    "cell_start": csr["cell_start"], "cell_pin": csr["cell_pin"], "num_nets": csr["num_nets"]}
        net = pin_net.long()
        self.all = {"cell": pin_cell.int().contiguous(), "off": pin_off.float().contiguous(), "net": net.int(),
                    "num_nets": num_nets,
                    "net_start": torch.searchsorted(net, torch.arange(num_nets + 1, device=dev)).int()}
        self._gamma = torch.empty(1, device=dev)

    def pins(self):
        """Every pin of every net: dict(cell, off, net, num_nets), long indices."""
        a = self.all
        return {"cell": a["cell"].long(), "off": a["off"], "net": a["net"].long(), "num_nets": a["num_nets"]}

    def hpwl(self, pos):
        """Exact HPWL over every net, summed in FP64 on the device -> 0-d float64 tensor."""
        a = self.all
        return gp_cuda.net_hpwl(pos, a["cell"], a["off"], a["net_start"], self.num_mov)

    def forward(self, pos, gamma):
        """gamma: float or 1-element device tensor -> (WA, gradient [num_mov, 2], HPWL of the optimized nets); the
        totals are 0-d float64 device tensors."""
        s = self.opt
        if not torch.is_tensor(gamma):
            gamma = self._gamma.fill_(gamma)
        # a level evaluates this host-side twin a few times only (the engine has its own plan): the plan-free
        # kernel, which needs no set-up
        grad, out = gp_cuda.wa(pos, s["cell32"], s["off"], s["net_start"], s["cell_start"], s["cell_pin"],
                               self.num_mov, gamma, planned=False)
        total = out.double().sum(0)
        return total[0], grad, total[1]
