"""Synthetic observation records generated with the library's own kernels.

A truth trajectory is one particle run through Model::transitionSamples
(bihip_transition) and its observations are drawn with
Model::observationSamples (bihip_obs_sample), as LibBi's `sample --target
prior` would produce them.  No dataset is downloaded.
"""
import ctypes

import numpy as np

from . import capi


def _c(t):
    return ctypes.c_void_p(t.data_ptr())


def simulate(ctx, model, x0, T, dt, seed, t0=0.0, batch=0):
    """Truth (T+1 x d) and observations (T x nobs) of one particle, in float64.

    Step k (k = 1..T) uses Philox step k for both the transition noise and
    the observation noise, as the filter does.
    """
    import torch
    m = capi.Model.from_buffer_copy(model)
    m.precision = capi.F64
    L = capi.lib()
    x = torch.tensor(np.asarray(x0, np.float64).reshape(m.d, 1), device="cuda")
    nobs = capi.nobs(m)
    y = torch.empty((T, nobs), dtype=torch.float64, device="cuda")
    xs = torch.empty((T + 1, m.d), dtype=torch.float64, device="cuda")
    xs[0] = x[:, 0]
    for k in range(T):
        capi.check(L.bihip_transition(ctx.h, ctypes.byref(m), seed, batch, k + 1,
                                      t0 + k * dt, t0 + (k + 1) * dt, 1, _c(x), _c(x), 1, None,
                                      None))
        capi.check(L.bihip_obs_sample(ctx.h, ctypes.byref(m), seed, batch, k + 1, 1, _c(x), 1,
                                      _c(y[k])))
        xs[k + 1] = x[:, 0]
    torch.cuda.synchronize()
    return xs.cpu().numpy(), y.cpu().numpy()


def l96_spinup(ctx, d=40, F=8.0, duration=10.0):
    """A point on the Lorenz-96 attractor: x = F + 0.01 e_0 run for `duration`
    time units without noise (DOPRI5, FP64, rtol = atol = 1e-5)."""
    import torch
    m = capi.make_model("l96", d=d, precision=capi.F64, F=F, noise_sd=0.0)
    x0 = np.full(d, F)
    x0[0] += 0.01
    x = torch.tensor(x0.reshape(d, 1), device="cuda")
    capi.check(capi.lib().bihip_transition(ctx.h, ctypes.byref(m), 0, 0, 0, 0.0, duration, 1,
                                           _c(x), _c(x), 1, None, None))
    torch.cuda.synchronize()
    return x[:, 0].cpu().numpy()
