"""xpbd_mlp_backward_noalloc (static device workspace, allocates nothing)
against xpbd_mlp_backward (caller's workspace), through the C ABI: the same
kernels in the same order, so the gradients must be bit-identical, also on a
second call with fewer rows."""
import ctypes as C

import numpy as np
import pytest

from rlmusclev2_amd import capi

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_backward_equals_workspace_backward():
    import torch
    from rlmusclev2_amd.mlp import NIN, NOUT, Mlp
    lib = capi.load_library()
    net = Mlp(seed=3)
    ng = int(lib.xpbd_mlp_grad_size())
    rng = np.random.default_rng(5)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in (9000, 777):                     # the second call has fewer rows
        print("rows", n, flush=True)
        X = rng.normal(size=(n, NIN)).astype(np.float32)
        X[:, 22:] = 0.0
        Xt = torch.tensor(X, device="cuda")
        _, H1, H2 = net.forward(Xt, keep=True)
        dY = torch.tensor(rng.normal(size=(n, NOUT)).astype(np.float32), device="cuda")
        nblk = (n + 31) // 32
        work = torch.empty(nblk * ng, dtype=torch.float32, device="cuda")
        g_ws = torch.full((ng,), np.nan, dtype=torch.float32, device="cuda")
        rc = lib.xpbd_mlp_backward(C.byref(net.dev), _p(Xt), _p(H1), _p(H2), _p(dY), n, _p(g_ws), _p(work),
                                   work.numel(), stream)
        assert rc == capi.XPBD_OK
        g_na = torch.full((ng,), np.nan, dtype=torch.float32, device="cuda")
        rc = lib.xpbd_mlp_backward_noalloc(C.byref(net.dev), _p(Xt), _p(H1), _p(H2), _p(dY), n, _p(g_na), stream)
        assert rc == capi.XPBD_OK
        torch.cuda.synchronize()
        a, b = g_ws.cpu().numpy(), g_na.cpu().numpy()
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        np.testing.assert_array_equal(b, a)


def test_noalloc_backward_rejects_rows_beyond_the_workspace():
    """checked on the host before any launch: nothing is read"""
    import torch
    from rlmusclev2_amd.mlp import Mlp
    lib = capi.load_library()
    net = Mlp(seed=3)
    t = torch.zeros(8, dtype=torch.float32, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.xpbd_mlp_backward_noalloc(C.byref(net.dev), _p(t), _p(t), _p(t), _p(t), 16385, _p(t), stream)
    assert rc == capi.XPBD_EINVAL
