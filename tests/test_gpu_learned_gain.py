"""What the trained networks gain over the method without them.

Both comparisons start every step from the same particle cloud: the exact
bootstrap filter's weighted particles, resampled to equal weights, so that the
methods differ only in the step under test and use the same transition noise
(common random numbers).

Guided proposal (Lorenz-96 d=200, q = 0.1, observations of the even
dimensions with r = 1): with additive Gaussian noise and a linear Gaussian
observation the one-step optimal proposal is known in closed form, and its
weight is p(y | x) = N(y; H Phi(x), (q^2 + r^2) I).  The network, trained
briefly on the GPU by the library's own BF16 MFMA kernels, reaches most of
the effective sample size of that optimum (two thirds on MI355X), about three
times the bootstrap step's, where the untrained network falls below the
bootstrap step; and its one-step evidence estimate is closer to the exact
conditional value (the mean of the optimal weights) than the bootstrap
step's (0.10 against 0.32 nats per step).  A whole guided filter estimates the same evidence as the bootstrap
filter (the weights are corrected exactly).

Surrogate flow (Lorenz-96 d=40): trained on pairs from the exact integrator,
its flow error falls far below the process noise of one interval, from an
untrained network's error far above it; and on the exact filter's clouds its
log-likelihood increment per interval stays close to the exact flow's (about
0.1 nats on MI355X), where an untrained network's is off by tens of nats.

Training data for both come from the model's own prior predictive: an
ensemble drawn around the filter's initial mean and recorded at every
interval, so the networks see the states the filter visits."""
import ctypes

import numpy as np
import pytest

from libbi_amd import capi, data, guided, surrogate

pytestmark = pytest.mark.gpu

DT = 0.05
Q, R = 0.1, 1.0


def _ess(lw):
    import torch
    w = torch.exp(lw - lw.max())
    return float(w.sum() ** 2 / (w * w).sum()) / lw.numel()


def _logmeanexp(lw):
    import torch
    return float(torch.logsumexp(lw, 0)) - np.log(lw.numel())


def _clouds(ctx, gm, y, mean0, P, T, seed, torch):
    """The exact filter's particles before step k+1, k = 0..T-1, resampled to
    equal weights (multinomial, torch generator): (k, cloud d x P)."""
    pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=seed, ess_rel=0.5)
    pf.set_observations(y)
    pf.init(mean0)
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 + seed)
    try:
        for k in range(T):
            if k > 0:
                pf.step(1)
            X, _, lws = pf.state()
            w = torch.exp(lws - lws.max())
            idx = torch.multinomial(w, P, replacement=True, generator=g)
            yield k, X[:, :P][:, idx].contiguous()
    finally:
        pf.close()


def _flow(ctx, m, X, t1, t2):
    Z = X.clone()
    capi.check(capi.lib().bihip_flow(ctx.h, ctypes.byref(m), t1, t2, Z.shape[1],
                                     ctypes.c_void_p(Z.data_ptr()), Z.shape[1], None, None))
    return Z


def _zero_output_net(ctx, dims, seed=3):
    net = capi.MLP(ctx, dims, seed=seed)
    p = net.get_params()
    last = dims[-1] * dims[-2] + dims[-1]
    p[-last:] = 0.0  # m = 0, s = 0: the proposal is the transition density
    net.set_params(p)
    return net


def test_trained_guided_proposal_approaches_the_optimal_proposal(ctx, torch_gpu):
    torch = torch_gpu
    # triples from the model's own prior predictive: an ensemble drawn around
    # the filter's initial mean, recorded at every interval from the first
    net, norm, info = guided.train(ctx, n=1 << 20, steps=(2000, 1500, 500), seed=3, burn=0,
                                   chunk=1 << 15)
    zero = _zero_output_net(ctx, guided.DIMS)
    untrained = capi.MLP(ctx, guided.DIMS, seed=11)
    gm = guided.model(8.0)
    P, T = 1 << 15, 10
    mean0 = data.l96_spinup(ctx, d=guided.D)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=31)
    ess = {k: [] for k in ("bootstrap", "guided", "untrained", "optimal")}
    err = {k: [] for k in ("bootstrap", "guided")}
    for k, C in _clouds(ctx, gm, y, mean0, P, T, 4, torch):
        Phi = _flow(ctx, gm, C, k * DT, (k + 1) * DT)
        yk = torch.as_tensor(y[k], device="cuda")
        opt = (-0.5 * ((yk[:, None] - Phi[0::2].double()) ** 2).sum(0) / (Q * Q + R * R)
               - 0.5 * guided.NOBS * np.log(2 * np.pi * (Q * Q + R * R)))
        ess["optimal"].append(_ess(opt))
        exact = _logmeanexp(opt)  # E[p(y | x') | cloud], the exact conditional value
        for name, n in (("bootstrap", zero), ("guided", net), ("untrained", untrained)):
            X = Phi.clone()
            inc = n.guided_propose(norm, 8.0, X, y[k], seed=9, step=k + 1)
            ess[name].append(_ess(inc))
            if name in err:
                err[name].append(abs(_logmeanexp(inc) - exact))
    m = {k: float(np.mean(v)) for k, v in ess.items()}
    e = {k: float(np.mean(v)) for k, v in err.items()}
    print(f"training {info}; mean one-step ESS fraction over {T} clouds of {P}: "
          + ", ".join(f"{k} {v:.4f}" for k, v in m.items())
          + f"; per-step |evidence error| bootstrap {e['bootstrap']:.4f}, guided "
          f"{e['guided']:.4f} nats", flush=True)
    print("per cloud:", {k: np.round(v, 4).tolist() for k, v in ess.items()}, flush=True)
    # the optimal proposal is the ceiling, and the trained network gets close
    # to it; the untrained network is worse than not having one
    assert m["guided"] > 0.5 * m["optimal"]
    assert m["guided"] > 1.5 * m["bootstrap"]
    assert m["untrained"] < m["bootstrap"]
    # its evidence increment is nearer the exact conditional value
    assert e["guided"] < e["bootstrap"]
    for n in (net, zero, untrained):
        n.close()


def test_guided_filter_estimates_the_bootstrap_evidence(ctx, torch_gpu):
    """The exact weight correction keeps the evidence: guided and bootstrap
    filters agree within their Monte Carlo error over six seeds."""
    net, norm, _ = guided.train(ctx, n=1 << 19, steps=(1500, 1000, 500), seed=3, burn=0,
                                chunk=1 << 15)
    T, P, seeds = 8, 1 << 15, (1, 2, 3, 4, 5, 6)
    mean0 = data.l96_spinup(ctx, d=guided.D)
    _, y = data.simulate(ctx, guided.model(8.0), mean0, T, DT, seed=31)
    z = {}
    for kind in ("bootstrap", "guided"):
        zs = []
        for s in seeds:
            pf = capi.ParticleFilter(ctx, guided.model(8.0), P, T, DT, seed=s, ess_rel=0.5)
            pf.set_observations(y)
            pf.init(mean0)
            if kind == "guided":
                pf.set_guided(net, norm)
            pf.step(T)
            zs.append(pf.term())
            pf.close()
        z[kind] = np.array(zs)
    zb, zg = z["bootstrap"], z["guided"]
    se = np.sqrt(zb.var(ddof=1) / len(seeds) + zg.var(ddof=1) / len(seeds))
    print(f"bootstrap logZ {zb.mean():.2f} sd {zb.std(ddof=1):.3f}; guided logZ "
          f"{zg.mean():.2f} sd {zg.std(ddof=1):.3f}; difference {zg.mean() - zb.mean():.2f} "
          f"(se {se:.2f})", flush=True)
    assert abs(zg.mean() - zb.mean()) < 4 * se + 0.1
    net.close()


def test_trained_surrogate_flow_and_local_loglik_error(ctx, torch_gpu):
    torch = torch_gpu
    # pairs from the model's own prior predictive: an ensemble drawn around the
    # filter's initial mean, recorded at every interval from the first
    X, Y = surrogate.flow_pairs(ctx, 1 << 20, seed=5, burn=0, chunk=1 << 15)
    Xt, Yt = surrogate.flow_pairs(ctx, 1 << 16, seed=99, burn=0, chunk=1 << 12)
    norm0 = surrogate.normalisation(X, Y)
    untrained = capi.MLP(ctx, surrogate.DIMS, seed=7)
    err0, scale = surrogate.flow_error(ctx, untrained, norm0, Xt, Yt)
    net, norm, info = surrogate.train(ctx, steps=(5000, 3000), lrs=(2e-3, 5e-4), seed=7,
                                      pairs=(X, Y))
    err, _ = surrogate.flow_error(ctx, net, norm, Xt, Yt)
    print(f"surrogate flow RMSE on held-out pairs: untrained {err0:.4f}, trained {err:.4f} "
          f"(increments RMS {scale:.3f}, process noise sd {Q}); {info}", flush=True)
    assert err0 > 0.3       # an untrained network is no flow at all
    assert err < 0.06       # well below the noise the filter adds every interval
    # per-interval log-likelihood increments on the exact filter's clouds: the
    # exact FP32 flow, the trained and the untrained surrogate, then the same
    # noise and the observation density
    P, T = 1 << 16, 20
    mean0 = data.l96_spinup(ctx, d=40)
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=Q, obs_sd=R,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=41)
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    diff = {"trained": [], "untrained": []}
    rmse = []
    for k, C in _clouds(ctx, gm, y, mean0, P, T, 8, torch):
        eps = Q * torch.randn(C.shape, generator=g, device="cuda", dtype=torch.float32)
        yk = torch.as_tensor(y[k], device="cuda")

        def ll(Z):
            Xn = (Z + eps).double()
            inc = (-0.5 * ((yk[:, None] - Xn[0::2]) / R) ** 2).sum(0)
            return _logmeanexp(inc)
        Zx = _flow(ctx, gm, C, k * DT, (k + 1) * DT)
        Zs = net.surrogate_flow(norm, 8.0, C)
        rmse.append(float((Zs - Zx).pow(2).mean().sqrt()))
        exact = ll(Zx)
        diff["trained"].append(ll(Zs) - exact)
        diff["untrained"].append(ll(untrained.surrogate_flow(norm0, 8.0, C)) - exact)
    dt_ = np.array(diff["trained"])
    du = np.array(diff["untrained"])
    print("per interval: rmse", np.round(rmse, 4).tolist(), "trained d", np.round(dt_, 3).tolist(),
          "untrained d", np.round(du, 1).tolist(), flush=True)
    print(f"on the filter's clouds: flow RMSE {np.mean(rmse):.4f}; log-likelihood increment "
          f"error per interval, trained: mean |d| {np.abs(dt_).mean():.4f}, sum {dt_.sum():.4f}; "
          f"untrained: mean |d| {np.abs(du).mean():.1f}", flush=True)
    assert np.mean(rmse) < 0.06
    assert np.abs(du).mean() > 50.0 * np.abs(dt_).mean()
    assert np.abs(dt_).mean() < 0.5
    untrained.close()
    net.close()
