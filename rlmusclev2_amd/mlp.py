"""Warm-start MLP driver: parameters in torch device tensors, forward and
backward through the HIP kernels of the C ABI (xpbd_mlp_forward /
xpbd_mlp_backward), Adam in torch.

Per-tet network 24 -> 64 -> 64 -> 4 (22 features, 3 multipliers used).
Layers 1-2 use BF16 weights on MFMA with FP32 accumulation; FP32 master
weights are kept for the optimizer and cast to BF16 before each call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi

NIN, NH, NOUT = 24, 64, 4
KEYS = ("W1", "b1", "W2", "b2", "W3", "b3")
SHAPES = {"W1": (NH, NIN), "b1": (NH,), "W2": (NH, NH), "b2": (NH,), "W3": (NOUT, NH), "b3": (NOUT,)}


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Mlp:
    def __init__(self, params: dict | None = None, device="cuda", seed: int = 0):
        import torch
        self.torch = torch
        self.lib = capi.load_library()
        self.device = device
        if params is None:
            rng = np.random.default_rng(seed)
            params = {
                "W1": rng.normal(0, np.sqrt(2.0 / 22), (NH, NIN)), "b1": np.zeros(NH),
                "W2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
                "W3": rng.normal(0, 0.1 * np.sqrt(1.0 / NH), (NOUT, NH)), "b3": np.zeros(NOUT)}
        self.p = {k: torch.tensor(np.asarray(params[k], np.float32).reshape(SHAPES[k]), device=device)
                  for k in KEYS}
        self.ng = int(self.lib.xpbd_mlp_grad_size())
        self._refresh()
        self._adam = None

    def _refresh(self):
        t = self.torch
        self.w1b = self.p["W1"].to(t.bfloat16).contiguous()
        self.w2b = self.p["W2"].to(t.bfloat16).contiguous()
        self.dev = capi.XpbdMlpDev(_ptr(self.w1b), _ptr(self.p["b1"]), _ptr(self.w2b), _ptr(self.p["b2"]),
                                   _ptr(self.p["W3"]), _ptr(self.p["b3"]))

    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def forward(self, X, keep=False):
        t = self.torch
        n = X.shape[0]
        Y = t.empty((n, NOUT), dtype=t.float32, device=self.device)
        H1 = t.empty((n, NH), dtype=t.float32, device=self.device) if keep else None
        H2 = t.empty((n, NH), dtype=t.float32, device=self.device) if keep else None
        rc = self.lib.xpbd_mlp_forward(C.byref(self.dev), _ptr(X), n, _ptr(Y),
                                       _ptr(H1) if keep else None, _ptr(H2) if keep else None, self.stream())
        if rc != 0:
            raise capi.XpbdError(rc, "xpbd_mlp_forward")
        return (Y, H1, H2) if keep else Y

    def backward(self, X, H1, H2, dY):
        t = self.torch
        n = X.shape[0]
        nblk = (n + 31) // 32
        scratch = t.empty(nblk * self.ng, dtype=t.float32, device=self.device)
        grad = t.empty(self.ng, dtype=t.float32, device=self.device)
        rc = self.lib.xpbd_mlp_backward(C.byref(self.dev), _ptr(X), _ptr(H1), _ptr(H2), _ptr(dY), n, _ptr(grad),
                                        _ptr(scratch), scratch.numel(), self.stream())
        if rc != 0:
            raise capi.XpbdError(rc, "xpbd_mlp_backward")
        out, o = {}, 0
        for k in KEYS:
            m = int(np.prod(SHAPES[k]))
            out[k] = grad[o:o + m].view(SHAPES[k])
            o += m
        return out

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        t = self.torch
        if self._adam is None:
            self._adam = {"t": 0, "m": {k: t.zeros_like(v) for k, v in self.p.items()},
                          "v": {k: t.zeros_like(v) for k, v in self.p.items()}}
        st = self._adam
        st["t"] += 1
        for k in KEYS:
            g = grads[k] + wd * self.p[k]
            st["m"][k].mul_(b1).add_(g, alpha=1 - b1)
            st["v"][k].mul_(b2).addcmul_(g, g, value=1 - b2)
            mh = st["m"][k] / (1 - b1 ** st["t"])
            vh = st["v"][k] / (1 - b2 ** st["t"])
            self.p[k].sub_(lr * mh / (vh.sqrt() + eps))
        self._refresh()

    def numpy_params(self):
        return {k: v.detach().cpu().numpy().astype(np.float64) for k, v in self.p.items()}
