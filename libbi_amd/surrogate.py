"""Learned surrogate of the Lorenz-96 flow over one observation interval.

The training pairs come from the library's own exact integrator: states x_t
of a stochastic ensemble on the attractor (the distribution a filter's
particles occupy), and targets Phi_dt(x_t) from the FP64 DOPRI5 flow at
rtol = atol = 1e-8 (bihip_flow).  The network (41-256-256-256-40, GELU,
BF16 MFMA with FP32 accumulation) is trained on the GPU by bihip_mlp_train
to predict the normalised increment (Phi_dt(x) - x) / dx_scale from
((x - x_shift) / x_scale, (F - F0) / F_scale).  The trained network then
runs inside the filter step (bihip_pf_set_surrogate), and the exact
integrator stays available to measure its log-likelihood error.

Data preparation (normalising the pairs) uses torch elementwise operations on
the device tensors; every training and inference computation runs in the
library's kernels.
"""
import ctypes

import numpy as np

from . import capi, data

DIMS = [41, 256, 256, 256, 40]


def _c(t):
    return ctypes.c_void_p(t.data_ptr())


def flow_pairs(ctx, n, d=40, F=8.0, dt=0.05, seed=1, burn=20, chunk=1 << 20, tol=1e-8,
               spread=1.0):
    """(X, Y): n x d float32 device tensors with Y = Phi_dt(X).

    X: an ensemble of `chunk` particles drawn around a point on the attractor
    (sd `spread`), run through the stochastic model (noise sd 0.1 per
    interval) for `burn` intervals, then recorded every interval."""
    import torch
    mean0 = data.l96_spinup(ctx, d=d, F=F)
    msto = capi.make_model("l96", d=d, precision=capi.F64, F=F, noise_sd=0.1, init_sd=spread,
                           atol=1e-5, rtol=1e-5, h0=dt)
    mflow = capi.make_model("l96", d=d, precision=capi.F64, F=F, noise_sd=0.0, atol=tol,
                            rtol=tol, h0=dt)
    L = capi.lib()
    P = min(chunk, n)
    Xs = torch.empty((d, P), dtype=torch.float64, device="cuda")
    m0 = np.ascontiguousarray(mean0, np.float64)
    capi.check(L.bihip_init_samples(ctx.h, ctypes.byref(msto), seed, 0, P, _c(Xs), P,
                                    m0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    X = torch.empty((n, d), dtype=torch.float32, device="cuda")
    Y = torch.empty((n, d), dtype=torch.float32, device="cuda")
    filled, step = 0, 0
    while filled < n:
        step += 1
        capi.check(L.bihip_transition(ctx.h, ctypes.byref(msto), seed, 0, step, (step - 1) * dt,
                                      step * dt, P, _c(Xs), _c(Xs), P, None, None))
        if step <= burn:
            continue
        Z = Xs.clone()
        capi.check(L.bihip_flow(ctx.h, ctypes.byref(mflow), 0.0, dt, P, _c(Z), P, None, None))
        take = min(P, n - filled)
        X[filled:filled + take] = Xs[:, :take].T.float()
        Y[filled:filled + take] = Z[:, :take].T.float()
        filled += take
    torch.cuda.synchronize()
    return X, Y


def normalisation(X, Y, F=8.0):
    """Scalar normalisation (the model is symmetric under cyclic shifts)."""
    dx = (Y - X)
    return capi.SurrogateNorm(x_shift=float(X.mean()), x_scale=float(X.std()), par_shift=F,
                              par_scale=1.0, dx_scale=float(dx.std()))


def training_tensors(X, Y, norm, F=8.0):
    """Network inputs (n x 41) and normalised targets (n x 40)."""
    import torch
    n = X.shape[0]
    U = torch.empty((n, X.shape[1] + 1), dtype=torch.float32, device=X.device)
    U[:, :-1] = (X - norm.x_shift) / norm.x_scale
    U[:, -1] = (F - norm.par_shift) / norm.par_scale
    T = ((Y - X) / norm.dx_scale).contiguous()
    return U, T


def train(ctx, n_pairs=1 << 22, steps=(4000, 3000, 2000), lrs=(2e-3, 5e-4, 1e-4), batch=4096,
          seed=1, F=8.0, pairs=None):
    """Train the surrogate; returns (mlp, norm, report)."""
    X, Y = pairs if pairs is not None else flow_pairs(ctx, n_pairs, F=F, seed=seed)
    norm = normalisation(X, Y, F)
    U, T = training_tensors(X, Y, norm, F)
    net = capi.MLP(ctx, DIMS, seed=seed)
    losses = []
    for k, (s, lr) in enumerate(zip(steps, lrs)):
        losses.append(net.train(U, T, batch=batch, steps=s, lr=lr, seed=seed + 7 * k))
    losses = np.concatenate(losses) if losses else np.zeros(0)
    return net, norm, {"pairs": int(X.shape[0]), "steps": int(sum(steps)),
                       "final_minibatch_loss": float(losses[-100:].mean()) if len(losses) else None}


def train_over_forcings(ctx, Fs, pairs_per_F, steps=(6000, 6000, 6000, 6000),
                        lrs=(2e-3, 1e-3, 3e-4, 1e-4), batch=4096, seed=1, chunk=1 << 15):
    """Train one surrogate for a range of forcings (the network's parameter
    input), from each forcing's prior predictive; returns (mlp, norm, report).
    This is the network SMC^2 over F needs: trained at one F it ignores its
    parameter input."""
    import torch
    Fs = [float(F) for F in Fs]
    pairs = [flow_pairs(ctx, pairs_per_F, F=F, seed=seed + 11 * i, burn=0, chunk=chunk)
             for i, F in enumerate(Fs)]
    X = torch.cat([x for x, _ in pairs])
    Y = torch.cat([y for _, y in pairs])
    F0 = 0.5 * (min(Fs) + max(Fs))
    norm = normalisation(X, Y, F0)
    norm.par_scale = max(0.5 * (max(Fs) - min(Fs)), 1e-3)
    del X, Y
    Us, Ts = [], []
    for (x, y), F in zip(pairs, Fs):
        U, T = training_tensors(x, y, norm, F)
        Us.append(U)
        Ts.append(T)
    del pairs
    U, T = torch.cat(Us), torch.cat(Ts)
    del Us, Ts
    net = capi.MLP(ctx, DIMS, seed=seed)
    losses = []
    for k, (s, lr) in enumerate(zip(steps, lrs)):
        losses.append(net.train(U, T, batch=batch, steps=s, lr=lr, seed=seed + 7 * k))
    losses = np.concatenate(losses) if losses else np.zeros(0)
    return net, norm, {"forcings": Fs, "pairs": int(U.shape[0]), "steps": int(sum(steps)),
                       "final_minibatch_loss": float(losses[-100:].mean()) if len(losses) else None}


def flow_error(ctx, net, norm, X, Y, F=8.0):
    """RMS error of the surrogate flow against the exact flow on pairs (X, Y) (n x d)."""
    Xc = X.T.contiguous()
    out = net.surrogate_flow(norm, F, Xc)
    err = (out.T - Y).float()
    return float(err.pow(2).mean().sqrt()), float((Y - X).pow(2).mean().sqrt())
