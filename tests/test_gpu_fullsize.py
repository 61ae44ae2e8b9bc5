"""Parity at the configured sizes, not only at small test sizes.

- Batched SMC^2 inner filters at the full population: 1024 parameter
  particles (mu, sigma, c, e from log-normal priors) with N_j log-uniform in
  [256, 16384] (about 4.1 M particles), one persistent launch per step;
  sampled inner filters (the smallest, the largest and random ones) equal the
  oracle's filter with their theta_j and batch id j: ancestors bit-exact at
  every step, log-likelihood increments to 1e-9.
- The guided-proposal configuration's filter, Lorenz-96 d = 200 at N = 2^18 in
  FP32: against the FP64 oracle on the device run's own resampling history,
  and the fused proposal kernel against the oracle's guided step at the full
  N = 2^18.
- That configuration as configured: 64 forcing values x 2^18 particles in one
  guided filter bank, its first and last filters equal to the lone guided
  filters with their forcing and batch id.
- The surrogate's batched form as configured: 256 forcing values x 4096
  particles of Lorenz-96 d = 40 in one surrogate filter bank, its first and
  last filters equal to the lone surrogate filters."""
import numpy as np
import pytest

from libbi_amd import capi, data, guided
from gpu_helpers import oracle_model

pytestmark = pytest.mark.gpu

DT = 0.05


def smc2_population(B=1024, seed=2024):
    rng = np.random.default_rng(seed)
    N = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), B))).astype(np.int32)
    mu = np.exp(rng.normal(np.log(0.5), 0.2, B))
    sigma = np.exp(rng.normal(np.log(0.2), 0.2, B))
    c = np.exp(rng.normal(np.log(0.25), 0.1, B))
    e = np.exp(rng.normal(np.log(0.3), 0.1, B))
    return N, np.stack([mu, sigma, c, e], axis=1)


def test_full_smc2_population_matches_oracle(ctx, orc):
    T = 12
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, 100, 1)[:, 0]
    _, y = orc.simulate(m, 100, T, 1.0, x0)
    N, theta = smc2_population()
    gm = capi.make_model("pz", precision=capi.F64)
    bf = capi.BatchFilter(ctx, gm, N, theta, T, 1.0, seed=77, ess_rel=0.5, record=True)
    bf.set_observations(y)
    bf.init()
    bf.step(T)
    out = bf.outputs()
    blk, grid, _ = bf.tuning()
    print(f"1024 filters, {int(N.sum())} particles; autotuned block {blk}, grid {grid}; "
          f"resampled fraction {out['resampled'].mean():.3f}", flush=True)
    assert (out["status"] == 0).all()
    assert out["resampled"].sum() > 1000
    rng = np.random.default_rng(7)
    picks = sorted({int(np.argmin(N)), int(np.argmax(N))} | set(rng.choice(1024, 6, replace=False).tolist()))
    for j in picks:
        om = oracle_model(orc, gm)
        for i, v in enumerate(theta[j]):
            om.par[i] = v
        ref = orc.pf_run(om, y, int(N[j]), 1.0, 77, ess_rel=0.5, batch=j, want_ancestors=True)
        np.testing.assert_allclose(out["logIncrements"][j], ref["logIncrements"], rtol=1e-9,
                                   atol=1e-9)
        np.testing.assert_array_equal(out["resampled"][j], ref["resampled"])
        for k in range(T):
            np.testing.assert_array_equal(bf.ancestors(j, k), ref["ancestors"][k], f"{j} {k}")
    bf.close()


def test_guided_configuration_filter_vs_fp64_oracle(ctx, orc):
    """L96 d=200 FP32 at N = 2^18 (one of the 64 forcing values) against the
    FP64 oracle on the device run's resampling history, T = 10."""
    P, T, F = 1 << 18, 10, 7.9
    gm = guided.model(F)
    mean0 = data.l96_spinup(ctx, d=gm.d, F=F)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=31)
    pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=5, ess_rel=0.5, record=True)
    pf.set_observations(y)
    pf.init(mean0)
    pf.step(T)
    out = pf.outputs()
    logZ = pf.term()
    anc = np.array([pf.ancestors(k) for k in range(T)])
    pf.close()
    ref = orc.pf_run(oracle_model(orc, gm), y, P, DT, 5, ess_rel=0.5, mean0=mean0,
                     imposed_resampled=out["resampled"], imposed_ancestors=anc)
    rel = abs(logZ - ref["logZ"]) / abs(ref["logZ"])
    print(f"d=200 N=2^18 FP32: logZ {logZ:.6f} fp64 {ref['logZ']:.6f} rel {rel:.3e}; "
          f"resampled {int(out['resampled'].sum())}/{T}", flush=True)
    assert out["resampled"].sum() >= 3
    assert rel < 1e-5


def test_guided_proposal_kernel_at_full_size_matches_oracle(ctx, torch_gpu, orc):
    """The fused proposal over N = 2^18 particles against the oracle's guided
    step on the same network, states, observation and noise."""
    torch = torch_gpu
    P, seed, step = 1 << 18, 3, 2
    D, NOBS = guided.D, guided.NOBS
    net = capi.MLP(ctx, guided.DIMS, seed=23)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.3
    net.set_params(p)
    prm = net.get_params().astype(np.float64)
    norm = capi.GuidedNorm(x_shift=2.3, x_scale=3.6, innov_scale=1.0, par_shift=8.0,
                           par_scale=1.0)
    rng = np.random.default_rng(11)
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
    print(f"N=2^18 guided vs oracle: max |dx| {dx.max():.3e} (mean {dx.mean():.3e}); "
          f"max |dinc| {di.max():.3e} (mean {di.mean():.3e}), mean |inc| {scale:.1f}", flush=True)
    np.testing.assert_allclose(xg, xo, rtol=0, atol=2e-4)
    assert dx.mean() < 2e-5
    assert di.max() < 1e-3 * scale
    assert di.mean() < 2e-4 * scale
    net.close()


def test_guided_bank_at_full_size_equals_lone_filters(ctx, torch_gpu):
    """64 filters of 2^18 Lorenz-96 d=200 particles (FP32, forcings in
    [7.5, 8.5]) stepped as one bank with the learned proposal: filters 0 and
    63 equal the lone guided filters with their forcing and batch id, ancestors
    and states bit for bit, log-likelihood increments to rounding."""
    B, P, T = 64, 1 << 18, 4
    Fs = np.linspace(7.5, 8.5, B)
    gm = guided.model(8.0)
    mean0 = data.l96_spinup(ctx, d=gm.d, F=8.0)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=41)
    net = capi.MLP(ctx, guided.DIMS, seed=23)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.3
    net.set_params(p)
    norm = capi.GuidedNorm(x_shift=2.3, x_scale=3.6, innov_scale=1.0, par_shift=8.0,
                           par_scale=1.0)
    bank = capi.FilterBank(ctx, gm, P, T, DT, 6, Fs, ess_rel=0.5, record=True)
    bank.set_guided(net, norm)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    print(f"guided bank 64 x 2^18: resampled {int(out['resampled'].sum())}/{B * T}",
          flush=True)
    for j in (0, B - 1):
        m = capi.Model.from_buffer_copy(gm)
        m.par[0] = Fs[j]
        pf = capi.ParticleFilter(ctx, m, P, T, DT, 6, ess_rel=0.5, batch=j, record=True)
        pf.set_guided(net, norm)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-12)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k))
        Xb, _, _ = bank.state(j)
        Xp, _, _ = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy()), j
        pf.close()
    bank.close()
    net.close()


def test_surrogate_bank_at_full_size_equals_lone_filters(ctx, torch_gpu):
    """256 filters of 4096 Lorenz-96 d=40 particles (FP32, forcings in
    [7.5, 8.5]) stepped as one bank with the learned flow: filters 0 and 255
    equal the lone surrogate filters with their forcing and batch id."""
    B, P, T = 256, 4096, 6
    Fs = np.linspace(7.5, 8.5, B)
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=43)
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=5)
    p = net.get_params()
    p *= 0.1  # a near-zero flow correction
    net.set_params(p)
    norm = capi.SurrogateNorm(x_shift=2.3, x_scale=3.6, dx_scale=0.05, par_shift=8.0,
                              par_scale=1.0)
    bank = capi.FilterBank(ctx, gm, P, T, DT, 8, Fs, ess_rel=0.5, record=True)
    bank.set_surrogate(net, norm)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    for j in (0, B - 1):
        m = capi.Model.from_buffer_copy(gm)
        m.par[0] = Fs[j]
        pf = capi.ParticleFilter(ctx, m, P, T, DT, 8, ess_rel=0.5, batch=j, record=True)
        pf.set_surrogate(net, norm)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-12)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k))
        Xb, _, _ = bank.state(j)
        Xp, _, _ = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy()), j
        pf.close()
    bank.close()
    net.close()
