"""Helpers for the GPU parity tests: device buffers are torch tensors and all
compute goes through the C ABI (libbi_amd.capi)."""
import ctypes

import numpy as np

from libbi_amd import capi


def dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a.copy()).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empty(shape, dtype):
    import torch
    return torch.empty(shape, dtype=dtype, device="cuda")


def f64p(x):
    return ctypes.byref(x)


def oracle_model(orc, gm):
    """The oracle's model matching a capi model (float kernels draw 24-bit normals)."""
    m = orc.Model()
    m.model, m.d, m.integrator = gm.model, gm.d, gm.integrator
    m.nsteps, m.host_quirk = gm.nsteps, gm.host_quirk
    m.noise_bits = 32 if gm.precision == capi.F64 else 24
    m.h0, m.atol, m.rtol = gm.h0, gm.atol, gm.rtol
    for i in range(16):
        m.par[i] = gm.par[i]
    return m


def l96_spun_up(orc, d=40, seed=0):
    m = orc.make_model("l96", d=d)
    x = np.full(d, 8.0)
    x[0] += 0.01
    x, _ = orc.flow(m, x, 0.0, 10.0)
    return x
