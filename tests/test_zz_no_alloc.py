"""akua_mlp_adam_step_device, the Adam step that allocates nothing (gradient partials in a
fixed-size device array of the library), against akua_mlp_adam_step on the same inputs: two
calls in a row on one GPU, the second with fewer samples, both bit-identical."""
import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _bind(L):
    return L.akua_mlp_adam_step_device   # argument types set by _lib.load()


def test_no_alloc_adam_step_equals_the_allocating_one():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd import _lib, learned
    L = _lib.load()
    step = _bind(L)
    rng = np.random.default_rng(21)
    F = (rng.normal(size=(5000, 16)) * rng.uniform(0.1, 5, 16)).astype(np.float32)
    T = (rng.normal(size=(5000, 4)) * np.array([0.3, 1.0, 0.1, 2.0])).astype(np.float32)
    p = M.random_params(rng, F, T).astype(np.float32)
    m = np.zeros(M.TRAINABLE, np.float32)
    v = np.zeros(M.TRAINABLE, np.float32)
    w = np.array([1.0, 0.5, 2.0, 0.25], np.float32)
    dp, dm, dv = (torch.from_numpy(a.copy()).cuda() for a in (p, m, v))
    dloss = torch.zeros(1, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for t, n in ((1, 5000), (2, 300)):
        Fn, Tn = F[:n], T[:n]
        p, m, v, loss = learned.mlp_adam_step(p, m, v, t, 1e-3, Fn, Tn, w)
        dF, dT = torch.from_numpy(Fn.copy()).cuda(), torch.from_numpy(Tn.copy()).cuda()
        st = step(ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(dm.data_ptr()),
                  ctypes.c_void_p(dv.data_ptr()), t, 1e-3, ctypes.c_void_p(dF.data_ptr()),
                  ctypes.c_void_p(dT.data_ptr()), n, w.ctypes.data_as(ctypes.c_void_p),
                  ctypes.c_void_p(dloss.data_ptr()), ctypes.c_void_p(stream) if stream else None)
        assert st == 0, L.akua_last_error()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dp.cpu().numpy(), p, err_msg=f"params after call {t}")
        np.testing.assert_array_equal(dm.cpu().numpy(), m, err_msg=f"m after call {t}")
        np.testing.assert_array_equal(dv.cpu().numpy(), v, err_msg=f"v after call {t}")
        assert dloss.cpu().numpy()[0] == np.float32(loss)


def test_no_alloc_adam_step_rejects_bad_arguments():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd import _lib
    L = _lib.load()
    step = _bind(L)
    buf = torch.zeros(8192, dtype=torch.float32, device="cuda")
    a = ctypes.c_void_p(buf.data_ptr())
    w = np.ones(4, np.float32)
    wp = w.ctypes.data_as(ctypes.c_void_p)
    assert step(a, a, a, 1, 1e-3, a, a, 0, wp, a, None) == 1            # n < 1
    assert step(a, a, a, 0, 1e-3, a, a, 16, wp, a, None) == 1           # t < 1
    assert step(None, a, a, 1, 1e-3, a, a, 16, wp, a, None) == 1        # NULL
    w[2] = -1.0
    assert step(a, a, a, 1, 1e-3, a, a, 16, wp, a, None) == 1           # negative loss weight
