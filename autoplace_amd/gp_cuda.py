"""Drop-in for the reference's compiled module `gp_cuda` (cpp_to_py/gp_cuda.cpp:31-85): the same six functions with
the same tensor arguments and results, on the MI355X HIP kernels of libautoplace_gp.so.

Like the reference's binding, every function checks that its tensors are contiguous device tensors of the expected
dtype (RuntimeError otherwise), allocates its outputs on the inputs' device, and launches on the current stream
(torch.cuda.current_stream(), so the calls compose with torch and with stream capture).

`poisson_field` replaces the torch.fft call of src/core/electronic_density_layer.py:133.
"""
import collections
import ctypes
import weakref

import torch

from . import _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t, dtype, what):
    if not (t.is_cuda and t.is_contiguous() and t.dtype == dtype):
        raise RuntimeError("want a contiguous %s device tensor (%s)" % (str(dtype).replace("torch.", ""), what))
    return ctypes.c_void_p(t.data_ptr())


F = lambda t, w="": _ptr(t, torch.float32, w)      # noqa: E731
D = lambda t, w="": _ptr(t, torch.float64, w)      # noqa: E731
I = lambda t, w="": _ptr(t, torch.int32, w)        # noqa: E731


class _WaPlan(object):
    """A netlist's WA plan (ap_wa_plan), destroyed with this object."""

    def __init__(self, pin_cell, pin_off, net_start, num_nets):
        torch.cuda.current_stream().synchronize()      # the plan reads the netlist now, outside the stream
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ap_wa_plan_create(I(pin_cell, "pin_cell"), F(pin_off, "pin_off"),
                                                I(net_start, "net_start"), num_nets, pin_cell.numel(),
                                                ctypes.byref(h)), "wa_plan_create")
        self.handle = h
        self.versions = (pin_cell._version, pin_off._version, net_start._version)

    def close(self):
        if self.handle is not None:
            _lib.lib().ap_wa_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:                              # interpreter shutdown
            pass


# the plans of the netlists wa() has seen, by the netlist tensors' storage; an entry goes when one of its tensors is
# freed, when one of them was changed in place (torch's version counter), or as the oldest of more than _WA_PLANS
_WA_PLANS = 8
_wa_plans = collections.OrderedDict()


def _wa_plan(pin_cell, pin_off, net_start, num_nets):
    key = (pin_cell.data_ptr(), pin_off.data_ptr(), net_start.data_ptr(), pin_cell.numel(), num_nets)
    plan = _wa_plans.get(key)
    if plan is not None and plan.versions == (pin_cell._version, pin_off._version, net_start._version):
        _wa_plans.move_to_end(key)
        return plan
    if plan is not None:
        _wa_plans.pop(key).close()
    plan = _WaPlan(pin_cell, pin_off, net_start, num_nets)
    _wa_plans[key] = plan
    for t in (pin_cell, pin_off, net_start):
        weakref.finalize(t, _drop_wa_plan, key, plan)
    while len(_wa_plans) > _WA_PLANS:
        _wa_plans.popitem(last=False)[1].close()
    return plan


def _drop_wa_plan(key, plan):
    if _wa_plans.get(key) is plan:
        del _wa_plans[key]
    plan.close()


def wa(pos, pin_cell, pin_off, net_start, cell_start, cell_pin, num_mov, gamma, planned=True):
    """-> (gradient [num_mov, 2], per-net (WA, HPWL) [num_nets, 2]) (gp_cuda.cpp:31-38).

    planned (not in the reference's signature; default on): the netlist's WA plan, made on the first call with
    these netlist tensors and kept (ap_wa_wirelength_planned); False: the plan-free kernel (ap_wa_wirelength). The
    two agree to float rounding."""
    num_nets = net_start.size(0) - 1
    pin_grad = torch.empty_like(pin_off)
    out = pos.new_empty((num_nets, 2))
    grad = pos.new_empty((num_mov, 2))
    args = (F(pos, "pos"), I(pin_cell, "pin_cell"), F(pin_off, "pin_off"), I(net_start, "net_start"),
            I(cell_start, "cell_start"), I(cell_pin, "cell_pin"), num_nets, num_mov, F(gamma, "gamma"), F(pin_grad),
            F(out), F(grad), pin_cell.numel(), _stream())
    if planned and num_nets > 0:
        plan = _wa_plan(pin_cell, pin_off, net_start, num_nets)
        _lib.check(_lib.lib().ap_wa_wirelength_planned(plan.handle, *args), "wa")
    else:
        _lib.check(_lib.lib().ap_wa_wirelength(*args), "wa")
    return grad, out


def density_map(pos, size, ratio, num_mov, bx, by, nbx, nby, fixed, target):
    """-> (total density mirrored [2 nbx, 2 nby], density of cells + blocked area above the target [nbx, nby])
    (gp_cuda.cpp:41-48)."""
    L = _lib.lib()
    nbytes = L.ap_density_map_work_bytes(pos.size(0), num_mov, nbx, nby)
    if nbytes < 0:
        raise ValueError("density_map: bad sizes (%d nodes, %d cells, grid %d x %d)"
                         % (pos.size(0), num_mov, nbx, nby))
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=pos.device)   # torch's caching allocator
    rho = pos.new_empty((2 * nbx, 2 * nby))
    over = pos.new_empty((nbx, nby))
    _lib.check(L.ap_density_map_ws(F(pos, "pos"), F(size, "size"), F(ratio, "ratio"), pos.size(0), num_mov, float(bx),
                                   float(by), nbx, nby, F(fixed, "fixed"), float(target), F(rho), F(over),
                                   ctypes.c_void_p(work.data_ptr()), nbytes, _stream()), "density_map")
    return rho, over


def nesterov_step(u, v, g, lo, hi, alpha, coef):
    """-> (u_next, v_next), v_next clamped to [lo, hi] (gp_cuda.cpp:51-56)."""
    un = torch.empty_like(v)
    vn = torch.empty_like(v)
    _lib.check(_lib.lib().ap_nesterov_step(F(u, "u"), F(v, "v"), F(g, "g"), F(lo, "lo"), F(hi, "hi"),
                                           F(alpha, "alpha"), F(coef, "coef"), v.numel(), F(un), F(vn), _stream()),
               "nesterov_step")
    return un, vn


def density_grad(pos, size, ratio, field, wl_grad, pins, area, weight, precond, bx, by, nbx, nby):
    """field [2, 2 nbx, 2 nby]; wl_grad [num_wl, 2] -> preconditioned gradient [num_nodes, 2] (gp_cuda.cpp:59-65)."""
    grad = torch.empty_like(pos)
    _lib.check(_lib.lib().ap_density_grad(F(pos, "pos"), F(size, "size"), F(ratio, "ratio"), F(field, "field"),
                                          F(wl_grad, "wl_grad"), F(pins, "pins"), F(area, "area"), pos.size(0),
                                          wl_grad.size(0), F(weight, "weight"), F(precond, "precond"), float(bx),
                                          float(by), nbx, nby, F(grad), _stream()), "density_grad")
    return grad


def control(S, kp, c, moved, turned, hpwl, ovf, wa_total):
    """One pass of the run's control; writes S and kp in place (gp_cuda.cpp:69-71)."""
    _lib.check(_lib.lib().ap_control(D(S, "S"), F(kp, "kp"), D(c, "c"), F(moved, "moved"), F(turned, "turned"),
                                     F(hpwl, "hpwl"), F(ovf, "ovf"), F(wa_total, "wa"), _stream()), "control")


def select(S, u, v, g, u_next, v_next, g_next, best):
    """Take an accepted step (u, v, g in place) and keep the best placement (gp_cuda.cpp:74-76)."""
    _lib.check(_lib.lib().ap_select_step(D(S, "S"), F(u, "u"), F(v, "v"), F(g, "g"), F(u_next, "u_next"),
                                         F(v_next, "v_next"), F(g_next, "g_next"), F(best, "best"), v.numel(),
                                         _stream()), "select")


def net_hpwl(pos, pin_cell, pin_off, net_start, num_mov):
    """Exact HPWL over every net (pins grouped by net_start), summed in FP64 on the device -> 0-d float64 tensor."""
    num_nets = net_start.size(0) - 1
    work = torch.empty(max((num_nets + 255) // 256, 1), dtype=torch.float64, device=pos.device)
    out = torch.empty((), dtype=torch.float64, device=pos.device)
    _lib.check(_lib.lib().ap_net_hpwl(F(pos, "pos"), I(pin_cell, "pin_cell"), F(pin_off, "pin_off"),
                                      I(net_start, "net_start"), num_nets, num_mov, D(work), D(out), _stream()),
               "net_hpwl")
    return out


class PoissonPlan(object):
    """Device plan of the Poisson solve on an nbx x nby grid (method: 'auto', 'gemm' or 'fft')."""

    METHODS = {"auto": _lib.AP_POISSON_AUTO, "gemm": _lib.AP_POISSON_GEMM, "fft": _lib.AP_POISSON_FFT}

    def __init__(self, nbx, nby, bx, by, method="auto"):
        self.nbx, self.nby = nbx, nby
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ap_poisson_create(nbx, nby, float(bx), float(by), self.METHODS[method],
                                                ctypes.byref(h)), "poisson_create")
        self._h = h
        self.method = {1: "gemm", 2: "fft"}[_lib.lib().ap_poisson_method(h)]

    def solve(self, rho, rho_ld, field_x, field_y, field_ld):
        """Raw form: device tensors / views with the given row strides."""
        _lib.check(_lib.lib().ap_poisson_solve(self._h, ctypes.c_void_p(rho.data_ptr()), rho_ld,
                                               ctypes.c_void_p(field_x.data_ptr()), ctypes.c_void_p(field_y.data_ptr()),
                                               field_ld, _stream()), "poisson_solve")

    def energy(self, rho, rho_ld=None):
        """Electrostatic energy 1/2 sum rho phi of the map (GEMM plans) -> 0-d float64 device tensor."""
        out = torch.empty((), dtype=torch.float64, device=rho.device)
        _lib.check(_lib.lib().ap_poisson_energy(self._h, ctypes.c_void_p(rho.data_ptr()), rho_ld or self.nby,
                                                ctypes.c_void_p(out.data_ptr()), _stream()), "poisson_energy")
        return out

    def field(self, rho):
        """rho [nbx, nby] (unmirrored) -> field [2, nbx, nby]."""
        if not (rho.is_cuda and rho.is_contiguous() and rho.dtype == torch.float32):
            raise RuntimeError("want a contiguous float32 device tensor (rho)")
        out = rho.new_empty((2, self.nbx, self.nby))
        self.solve(rho, self.nby, out[0], out[1], self.nby)
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.lib().ap_poisson_destroy(h)
            except Exception:
                pass
            self._h = None


def poisson_field(mirrored, plan):
    """Field of a mirrored map [2 nbx, 2 nby] in the reference's layout [2, 2 nbx, 2 nby]; the [:, :nbx, :nby]
    quadrant is the one density_grad reads and the only one written."""
    nbx, nby = plan.nbx, plan.nby
    if not (mirrored.is_cuda and mirrored.is_contiguous() and tuple(mirrored.shape) == (2 * nbx, 2 * nby)):
        raise RuntimeError("want the contiguous mirrored map [2 nbx, 2 nby] on the device")
    field = mirrored.new_empty((2, 2 * nbx, 2 * nby))
    plan.solve(mirrored, 2 * nby, field[0], field[1], 2 * nby)
    return field
