"""Learned guided proposal (bihip_guided_propose, bihip_pf_set_guided) on the
Lorenz-96 d=200 filter: a network that outputs zero is the bootstrap step;
the fused proposal kernel against the GEMM path of the same network and the
exact weight formula; a random proposal still estimates the same evidence;
and the Gaussian-loss gradient against the oracle."""
import numpy as np
import pytest

from libbi_amd import capi, guided

pytestmark = pytest.mark.gpu

D, NOBS = 200, 100


def _zero_output_net(ctx, seed=3):
    net = capi.MLP(ctx, guided.DIMS, seed=seed)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] = 0.0  # m = 0, s = 0: the proposal is the transition density
    net.set_params(p)
    return net


def _norm():
    return capi.GuidedNorm(x_shift=2.3, x_scale=3.6, innov_scale=1.0, par_shift=8.0,
                           par_scale=1.0)


def _filter(ctx, y, mean0, P, seed, record=False):
    pf = capi.ParticleFilter(ctx, guided.model(), P, len(y), 0.05, seed, ess_rel=0.5,
                             record=record)
    pf.set_observations(y)
    pf.init(mean0)
    return pf


def _data(ctx, T):
    from libbi_amd import data
    mean0 = data.l96_spinup(ctx, d=D)
    _, y = data.simulate(ctx, guided.model(), mean0, T, 0.05, seed=21)
    return mean0, y


def test_zero_proposal_is_bootstrap(ctx, torch_gpu):
    T, P = 6, 2048
    mean0, y = _data(ctx, T)
    net = _zero_output_net(ctx)
    res = []
    for use in (False, True):
        pf = _filter(ctx, y, mean0, P, 5, record=True)
        if use:
            pf.set_guided(net, _norm())
        pf.step(T)
        res.append((pf.term(), pf.outputs(), [pf.ancestors(k) for k in range(T)]))
        pf.close()
    (z0, o0, a0), (z1, o1, a1) = res
    same = (np.array(a0) == np.array(a1)).mean()
    print("bootstrap", z0, "zero proposal", z1, "ancestors equal", same, flush=True)
    # the proposal kernel forms x' = Phi + q (0 + 1 eps) and the log-weight in its
    # own summation order, so the two agree to FP32 rounding of the state
    # (one ulp of x' moves log p(y | x') by ~1e-4 over 100 observations)
    assert z1 == pytest.approx(z0, rel=1e-5)
    np.testing.assert_array_equal(o0["resampled"], o1["resampled"])
    np.testing.assert_allclose(o1["logIncrements"], o0["logIncrements"], rtol=1e-5)
    assert same > 0.99
    net.close()


def test_proposal_kernel_matches_gemm_path(ctx, torch_gpu, orc):
    torch = torch_gpu
    P, seed, step = 1000, 7, 3
    net = capi.MLP(ctx, guided.DIMS, seed=11)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.3  # moderate m and s
    net.set_params(p)
    norm = _norm()
    rng = np.random.default_rng(2)
    Phi = (2.3 + 3.6 * rng.normal(size=(D, P))).astype(np.float32)
    y = rng.normal(size=NOBS) + 2.0
    X = torch.from_numpy(Phi.copy()).cuda()
    inc = net.guided_propose(norm, 8.0, X, y, seed, step).cpu().numpy()
    xn = X.cpu().numpy().astype(np.float64)
    # the same network through the GEMM path on the kernel's inputs
    U = np.zeros((P, guided.DIMS[0]), np.float32)
    U[:, :D] = (Phi.T - np.float32(2.3)) * np.float32(1 / np.float32(3.6))
    U[:, D:D + NOBS] = (y[None, :] - Phi.T[:, 0::2].astype(np.float64)).astype(np.float32)
    out = net.forward(torch.from_numpy(U).cuda()).cpu().numpy().astype(np.float64)
    m, s = out[:, :D].T, out[:, D:].T
    z = np.stack([orc.normals(seed, i, step, orc.TAG_NOISE, D, bits=24) for i in range(P)], axis=1)
    d = m + np.exp(0.5 * s) * z
    x_ref = Phi.astype(np.float64) + 0.1 * d
    np.testing.assert_allclose(xn, x_ref, rtol=0, atol=5e-5)
    obs = (-0.5 * (y[:, None] - xn[0::2]) ** 2 - 0.5 * np.log(2 * np.pi)).sum(axis=0)
    inc_ref = obs + (-0.5 * d * d + 0.5 * z * z + 0.5 * s).sum(axis=0)
    print("max |inc - ref|", np.abs(inc - inc_ref).max(), flush=True)
    np.testing.assert_allclose(inc, inc_ref, rtol=0, atol=2e-2)
    net.close()


def test_random_proposal_estimates_the_same_evidence(ctx, torch_gpu):
    """Any proposal with the exact weight correction estimates p(y): the
    guided and bootstrap filters agree on log Z within Monte Carlo error."""
    T, P, R = 3, 1 << 14, 8
    mean0, y = _data(ctx, T)
    net = capi.MLP(ctx, guided.DIMS, seed=13)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.2
    net.set_params(p)
    zs = {False: [], True: []}
    for use in (False, True):
        for r in range(R):
            pf = _filter(ctx, y, mean0, P, 100 + r)
            if use:
                pf.set_guided(net, _norm())
            pf.step(T)
            zs[use].append(pf.term())
            pf.close()
    a, b = np.array(zs[False]), np.array(zs[True])
    se = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
    print("bootstrap", a.mean(), a.std(), "guided", b.mean(), b.std(), "se", se, flush=True)
    assert abs(a.mean() - b.mean()) < 4 * se + 0.05
    net.close()


def test_gaussian_loss_gradient_matches_oracle(ctx, torch_gpu, orc):
    torch = torch_gpu
    dims = [41, 64, 64, 16]
    net = capi.MLP(ctx, dims, seed=4)
    net.set_loss(capi.LOSS_GAUSS)
    prm = net.get_params().astype(np.float64)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(500, 41)).astype(np.float32)
    T = rng.normal(size=(500, 8)).astype(np.float32)
    loss, g = net.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    rl, rg = orc.mlp_loss_grad(dims, prm, X.astype(np.float64), T.astype(np.float64), bf16=True,
                               loss=orc.LOSS_GAUSS)
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print("gauss loss", loss, rl, "grad rel", rel, flush=True)
    assert loss == pytest.approx(rl, rel=1e-4)
    assert rel < 5e-3
    net.close()


def test_proposal_kernel_matches_oracle(ctx, torch_gpu, orc):
    """The fused proposal kernel (bihip_guided_propose) against the FP64
    oracle's guided step (orc_guided_step, bf16 rounding points of the GPU
    network) on the same 301-256-256-256-400 network, flowed states,
    observations and noise: the proposed states and the log-weight increments
    (a sum of 300 terms) agree to the network's bf16 precision."""
    torch = torch_gpu
    P, seed, step = 2000, 9, 4
    net = capi.MLP(ctx, guided.DIMS, seed=17)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.3
    net.set_params(p)
    prm = net.get_params().astype(np.float64)
    norm = _norm()
    rng = np.random.default_rng(4)
    Phi = (2.3 + 3.6 * rng.normal(size=(D, P))).astype(np.float32)
    y = rng.normal(size=NOBS) + 2.0
    X = torch.from_numpy(Phi.copy()).cuda()
    inc = net.guided_propose(norm, 8.0, X, y, seed, step).cpu().numpy()
    xg = X.cpu().numpy().astype(np.float64)
    nv = (norm.x_shift, norm.x_scale, norm.innov_scale, norm.par_shift, norm.par_scale)
    xo, io = orc.guided_step(guided.DIMS, prm, nv, 8.0, Phi.astype(np.float64), y, seed, step,
                             q=0.1, obs_sd=1.0)
    dx = np.abs(xg - xo)
    di = np.abs(inc - io)
    scale = np.abs(io).mean()
    print(f"guided vs oracle: max |dx| {dx.max():.3e} (mean {dx.mean():.3e}); "
          f"max |dinc| {di.max():.3e} (mean {di.mean():.3e}), mean |inc| {scale:.1f}", flush=True)
    # x' = Phi + 0.1 (m + e^{s/2} eps): bf16 network outputs, FP32 state
    np.testing.assert_allclose(xg, xo, rtol=0, atol=2e-4)
    assert dx.mean() < 2e-5
    # increments: relative to their size, the bf16 network's precision
    assert di.max() < 1e-3 * scale
    assert di.mean() < 2e-4 * scale
    net.close()
