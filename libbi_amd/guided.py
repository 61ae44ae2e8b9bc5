"""Learned guided proposal for the Lorenz-96 d=200 filter.

Training triples come from the library's own model kernels: states x of a
stochastic ensemble on the attractor, the exact flow Phi(x) (FP32 DOPRI5, the
flow the filter computes), x' = Phi(x) + q xi (the transition), and
observations y of x'.  The network (301-256-256-256-400, BF16 MFMA, FP32
accumulation) is trained on the GPU by bihip_mlp_train with the Gaussian
negative log-likelihood of the target (x' - Phi) / q given
u = ((Phi - x_shift) / x_scale, (y - H Phi) / obs_sd, (F - F0) / F_scale):
it learns the one-step optimal proposal p(x' | x, y).  The filter then draws
from it and corrects the weights exactly (bihip_pf_set_guided).

Data preparation (assembling the inputs) uses torch elementwise operations on
the device tensors; training and inference run in the library's kernels.
"""
import ctypes

import numpy as np

from . import capi, data

D, NOBS = 200, 100
DIMS = [D + NOBS + 1, 256, 256, 256, 2 * D]


def _c(t):
    return ctypes.c_void_p(t.data_ptr())


def model(F=8.0, precision=capi.F32):
    return capi.make_model("l96", d=D, precision=precision, F=F, noise_sd=0.1, obs_sd=1.0,
                           init_sd=1.0, atol=1e-5, rtol=1e-5, h0=0.05)


def triples(ctx, n, F=8.0, dt=0.05, seed=1, burn=20, chunk=1 << 18):
    """(Phi, Xn, Y): n x 200, n x 200, n x 100 device tensors (float32, float32,
    float64) with Xn = Phi + q xi and Y = H Xn + r eta."""
    import torch
    mean0 = data.l96_spinup(ctx, d=D, F=F)
    m = model(F)
    L = capi.lib()
    P = min(chunk, n)
    X = torch.empty((D, P), dtype=torch.float32, device="cuda")
    m0 = np.ascontiguousarray(mean0, np.float64)
    capi.check(L.bihip_init_samples(ctx.h, ctypes.byref(m), seed, 0, P, _c(X), P,
                                    m0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    Phi = torch.empty((n, D), dtype=torch.float32, device="cuda")
    Xn = torch.empty((n, D), dtype=torch.float32, device="cuda")
    Y = torch.empty((n, NOBS), dtype=torch.float64, device="cuda")
    yb = torch.empty((P, NOBS), dtype=torch.float64, device="cuda")
    filled, step = 0, 0
    while filled < n:
        step += 1
        if step > burn:
            Z = X.clone()
            capi.check(L.bihip_flow(ctx.h, ctypes.byref(m), (step - 1) * dt, step * dt, P, _c(Z),
                                    P, None, None))
        capi.check(L.bihip_transition(ctx.h, ctypes.byref(m), seed, 0, step, (step - 1) * dt,
                                      step * dt, P, _c(X), _c(X), P, None, None))
        if step <= burn:
            continue
        capi.check(L.bihip_obs_sample(ctx.h, ctypes.byref(m), seed + 1, 0, step, P, _c(X), P,
                                      _c(yb)))
        take = min(P, n - filled)
        Phi[filled:filled + take] = Z[:, :take].T
        Xn[filled:filled + take] = X[:, :take].T
        Y[filled:filled + take] = yb[:take]
        filled += take
    torch.cuda.synchronize()
    return Phi, Xn, Y


def normalisation(Phi, F=8.0, obs_sd=1.0):
    return capi.GuidedNorm(x_shift=float(Phi.mean()), x_scale=float(Phi.std()),
                           innov_scale=obs_sd, par_shift=F, par_scale=1.0)


def training_tensors(Phi, Xn, Y, norm, F=8.0, q=0.1):
    """Network inputs (n x 301, float32, the kernel's formulas) and targets (n x 200)."""
    import torch
    n = Phi.shape[0]
    U = torch.zeros((n, DIMS[0]), dtype=torch.float32, device=Phi.device)
    U[:, :D] = (Phi - norm.x_shift) * (1.0 / norm.x_scale)
    U[:, D:D + NOBS] = ((Y - Phi[:, 0::2].double()) * (1.0 / norm.innov_scale)).float()
    U[:, -1] = (F - norm.par_shift) / norm.par_scale
    T = ((Xn - Phi) / q).contiguous()
    return U, T


def train(ctx, n=1 << 21, steps=(3000, 2000, 1000), lrs=(1e-3, 3e-4, 1e-4), batch=4096, seed=1,
          F=8.0, data_=None, burn=20, chunk=1 << 18):
    """Train the proposal on n triples (an ensemble of `chunk` particles,
    recorded every interval after `burn` intervals); returns (mlp, norm, report)."""
    Phi, Xn, Y = data_ if data_ is not None else triples(ctx, n, F=F, seed=seed, burn=burn,
                                                         chunk=chunk)
    norm = normalisation(Phi, F)
    U, T = training_tensors(Phi, Xn, Y, norm, F)
    net = capi.MLP(ctx, DIMS, seed=seed)
    net.set_loss(capi.LOSS_GAUSS)
    losses = [net.train(U, T, batch=batch, steps=s, lr=lr, seed=seed + 5 * k)
              for k, (s, lr) in enumerate(zip(steps, lrs))]
    losses = np.concatenate(losses) if losses else np.zeros(0)
    return net, norm, {"triples": int(Phi.shape[0]), "steps": int(sum(steps)),
                       "final_minibatch_loss": float(losses[-100:].mean()) if len(losses) else None}
