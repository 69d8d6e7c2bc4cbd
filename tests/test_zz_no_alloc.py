"""ap_net_hpwl_noalloc (no work buffer, no allocation) against ap_net_hpwl (the caller's work buffer) through the C ABI:
two calls in a row, the second on a smaller netlist, each bit for bit equal to the existing entry point (same
kernels, same fixed-order FP64 sums) and to the FP64 oracle within float32 coordinate rounding; sizes above the fixed
scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import wa as owa

DEV = "cuda:0"


def _netlist(cells, nets, seed):
    from autoplace_amd import synth
    from autoplace_amd.database import PlaceData
    lef, d = synth.generate(cells, num_nets=nets, seed=seed, io_pins=6, name="hpwl%d" % seed)
    data = PlaceData(lef, d)
    rng = np.random.default_rng(seed)
    pos = (rng.random((data.num_mov, 2)) * [data.width, data.height]).astype(np.float32)
    return data, pos


def _t(a, dt):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)


def _both(lib, data, pos):
    p = _t(pos, torch.float32)
    cell = _t(data.pin_cell, torch.int32)
    off = _t(data.pin_off, torch.float32)
    start = _t(np.searchsorted(data.pin_net, np.arange(data.num_nets + 1)), torch.int32)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())   # noqa: E731
    work = torch.empty((data.num_nets + 255) // 256, dtype=torch.float64, device=DEV)
    a = torch.empty((), dtype=torch.float64, device=DEV)
    b = torch.full((), float("nan"), dtype=torch.float64, device=DEV)
    assert lib.ap_net_hpwl(ptr(p), ptr(cell), ptr(off), ptr(start), data.num_nets, data.num_mov, ptr(work), ptr(a),
                           stream) == 0
    assert lib.ap_net_hpwl_noalloc(ptr(p), ptr(cell), ptr(off), ptr(start), data.num_nets, data.num_mov, ptr(b),
                                   stream) == 0
    torch.cuda.synchronize()
    return float(a), float(b)


@pytest.mark.gpu
def test_noalloc_hpwl_equals_the_work_buffer_entry_point_twice():
    from autoplace_amd import _lib
    lib = _lib.lib()
    big, pos_big = _netlist(60000, 63000, 1)
    small, pos_small = _netlist(3000, 3150, 2)
    for data, pos in ((big, pos_big), (small, pos_small)):
        with_work, without = _both(lib, data, pos)
        ref = owa.hpwl_exact(pos, data.pin_cell, data.pin_off, data.pin_net, data.num_nets, data.num_mov)
        print("%d nets: work buffer %.17g, no allocation %.17g, oracle %.17g" % (data.num_nets, with_work, without,
                                                                                 ref))
        assert without == with_work
        assert abs(without - ref) / ref < 1e-6


def test_noalloc_hpwl_refuses_more_nets_than_its_scratch():
    """Checked on the host before anything touches a device (runs without a GPU)."""
    from autoplace_amd import _lib
    lib = _lib.lib()
    too_many = 65536 * 256 + 1
    dummy = ctypes.c_void_p(8)
    st = lib.ap_net_hpwl_noalloc(dummy, dummy, dummy, dummy, too_many, 10, dummy, None)
    assert st == _lib.AP_ERR_ARG
    assert b"above its fixed scratch" in lib.ap_last_error()
