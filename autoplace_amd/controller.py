"""The learned schedule controller (include/autoplace_gp.h part 4): its parameters, its training step on the device,
and the options that attach it to the placement flow.

A 12 -> 64 -> 64 -> 2 ReLU MLP in FP32. In every pass of the engine's device loop the control kernel evaluates it
on the pass's statistics; its outputs, clipped to +-0.2, are the logs of factors on the rule's density-weight growth
and on its new smoothing (the reference's rule, gp_cuda_kernel.cu:198-209, is what they scale; it has no learned
counterpart). After `window` passes without a 0.1 % overflow improvement the plain rule takes over for the level.
The output layer starts at zero, so an untrained network is the plain rule exactly.

Training (train_controller.py) is reward-weighted regression on trajectories this placer generates; the forward,
gradient and Adam kernels are the library's own (gp_mlp.hip).
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "controller_weights.npz")
IN, HID, OUT = 12, 64, 2
CLIP = 0.2
FEATURES = ("overflow", "100 x overflow change", "log HPWL / first HPWL", "10 x relative HPWL change",
            "WA model error", "log gamma / base gamma", "density share", "log step / first step",
            "0.2 log |wire grad| / |density grad|", "iteration / 1000", "log(1 + motion)", "coarse level")


def param_count():
    return _lib.lib().ap_mlp_param_count()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed=0, device="cuda"):
    """He-initialised hidden layers, a zero output layer (the plain rule until trained)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), np.float32)
    w1 = rng.normal(0, np.sqrt(2.0 / IN), (HID, IN))
    w2 = rng.normal(0, np.sqrt(2.0 / HID), (HID, HID))
    p[:HID * IN] = w1.ravel()
    o = HID * IN + HID
    p[o:o + HID * HID] = w2.ravel()
    return torch.tensor(p, device=device)


def load(path=WEIGHTS, device="cuda"):
    """-> (parameters on the device, metadata dict)."""
    z = np.load(path, allow_pickle=False)
    p = z["params"].astype(np.float32)
    if p.size != param_count():
        raise ValueError("%s: %d parameters, the network has %d" % (path, p.size, param_count()))
    meta = {k: z[k].tolist() for k in z.files if k != "params"}
    return torch.tensor(p, device=device), meta


def save(path, params, **meta):
    np.savez(path, params=params.detach().cpu().numpy().astype(np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def forward(params, x):
    """x [n, 12] device f32 -> [n, 2] unclipped outputs."""
    x = x.contiguous()
    out = torch.empty((x.shape[0], OUT), dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().ap_mlp_forward(_p(params), _p(x), x.shape[0], _p(out), _stream()), "mlp_forward")
    return out


def loss_grad(params, x, t, w):
    """(sum_s w_s |mlp(x_s) - t_s|^2 as a 0-d f64 tensor, its gradient [params])."""
    n = x.shape[0]
    grad = torch.empty_like(params)
    loss = torch.empty((), dtype=torch.float64, device=x.device)
    work = torch.empty(_lib.lib().ap_mlp_work_floats(n), dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().ap_mlp_grad(_p(params), _p(x.contiguous()), _p(t.contiguous()), _p(w.contiguous()), n,
                                      _p(grad), _p(loss), _p(work), _stream()), "mlp_grad")
    return loss, grad


class Adam(object):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.params, self.lr, self.betas, self.eps = params, lr, betas, eps
        self.m, self.v, self.step_count = torch.zeros_like(params), torch.zeros_like(params), 0

    def step(self, grad):
        self.step_count += 1
        _lib.check(_lib.lib().ap_mlp_adam(_p(self.params), _p(self.m), _p(self.v), _p(grad), self.step_count,
                                          float(self.lr), float(self.betas[0]), float(self.betas[1]),
                                          float(self.eps), _stream()), "mlp_adam")


def fit(params, x, t, w, epochs=200, lr=2e-3, batch=8192, seed=0):
    """Minimise sum_s w_s |mlp(x_s) - t_s|^2 (w normalised to sum 1 over each batch) with Adam on the device.
    params is updated in place. -> loss of the last full-data evaluation."""
    n = x.shape[0]
    opt = Adam(params, lr)
    g = torch.Generator(device="cpu").manual_seed(seed)
    for _ in range(epochs):
        order = torch.randperm(n, generator=g).to(x.device)
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            wb = w[idx]
            wb = wb / wb.sum().clamp_min(1e-30)
            _, grad = loss_grad(params, x[idx], t[idx], wb)
            opt.step(grad)
    loss, _ = loss_grad(params, x, t, w / w.sum().clamp_min(1e-30))
    return float(loss)


def options(params, sigma=0.0, seed=0, hold=1, window=50, log_cap=0):
    """The `controller` entry of the placement options (run_placement.place(opts=...)). params None: the plain rule,
    with no controller bookkeeping and no log; a network with a zero output layer is the plain rule with both."""
    return {"params": params, "sigma": float(sigma), "seed": int(seed), "hold": int(hold), "window": int(window),
            "log_cap": int(log_cap)}
