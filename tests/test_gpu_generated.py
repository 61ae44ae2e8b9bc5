"""Models generated from LibBi model files (libbi_amd/bi/*.bi, compiled into
the library by libbi_amd.dsl) run through the same filter kernels as the
hand-written models.

Pinned three ways: the PZ model of the LibBi user guide
(docs/src/users.tex:52-84) and the linear-Gaussian model, written in the
model language, against the oracle's filter with the same equations and
constants (FP64: log-likelihood to 1e-10, ancestors bit-exact); the same
files against the hand-written device models (FP32); and the discrete-time
AR1 model, which no hand-written model covers, against its exact Kalman
likelihood, also with observations between its delta times."""
import ctypes

import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import empty, host, oracle_model

pytestmark = pytest.mark.gpu


def _run(ctx, m, y, P, dt, seed, resampler=capi.SYSTEMATIC, ess_rel=0.5, record=True,
         batch=0):
    T = len(y)
    pf = capi.ParticleFilter(ctx, m, P, T, dt, seed, resampler=resampler, ess_rel=ess_rel,
                             record=record, batch=batch)
    pf.set_observations(y)
    pf.init()
    pf.step(T)
    z = pf.term()
    out = pf.outputs()
    anc = np.array([pf.ancestors(k) for k in range(T)]) if record else None
    pf.close()
    return z, out, anc


def _pz_pair(prec, integrator=None, mu=0.5, sigma=0.2):
    """The generated PZ model and the hand-written one with the file's constants
    (and the same integrator settings: the file's ode block gives h = 0.1 and
    leaves atoler and rtoler at their defaults of 1e-3)."""
    h = capi.make_model("pz", precision=prec, integrator=integrator, mu=mu, sigma=sigma,
                        c=0.25, e=0.3, m_l=0.1, m_q=0.1, obs_sd=0.2,
                        P0_logmean=np.log(2.0), P0_logsd=0.2, Z0_logmean=np.log(2.0),
                        Z0_logsd=0.1)
    g = capi.generated_model("PZ", precision=prec, integrator=integrator, mu=mu, sigma=sigma,
                             h0=h.h0, atol=h.atol, rtol=h.rtol, nsteps=h.nsteps)
    return g, h


def _pz_data(orc, T, seed=100):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return y


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.STRATIFIED, capi.MULTINOMIAL])
def test_generated_pz_matches_oracle(ctx, orc, resampler):
    """PZ.bi, FP64, RK4 h = 0.1, N = 4096, T = 60, resampling at ESS < 0.7 N."""
    y = _pz_data(orc, 60)
    g, h = _pz_pair(capi.F64)
    om = oracle_model(orc, h)
    P = 4096
    z, out, anc = _run(ctx, g, y, P, 1.0, 42, resampler, 0.7)
    ref = orc.pf_run(om, y, P, 1.0, 42, resampler=resampler, ess_rel=0.7, want_ancestors=True)
    assert z == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_generated_pz_dopri5_matches_oracle(ctx, orc):
    """The ode block's algorithm replaced at run time by RK5(4) (DOPRI5)."""
    y = _pz_data(orc, 30)
    g, h = _pz_pair(capi.F64, integrator=capi.DOPRI5)
    om = oracle_model(orc, h)
    z, out, anc = _run(ctx, g, y, 2048, 1.0, 7, ess_rel=0.5)
    ref = orc.pf_run(om, y, 2048, 1.0, 7, ess_rel=0.5, want_ancestors=True)
    assert z == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_array_equal(anc, ref["ancestors"])


@pytest.mark.parametrize("integrator", [capi.RK4, capi.DOPRI5])
def test_generated_pz_fp32_equals_handwritten(ctx, orc, integrator):
    """FP32 state: the generated functor against the hand-written PZ kernel, no
    resampling (the particles stay paired), logZ to the FP32 bar of 1e-5."""
    y = _pz_data(orc, 20)
    g, h = _pz_pair(capi.F32, integrator=integrator)
    zg, og, _ = _run(ctx, g, y, 4096, 1.0, 3, ess_rel=0.0, record=False)
    zh, oh, _ = _run(ctx, h, y, 4096, 1.0, 3, ess_rel=0.0, record=False)
    assert not og["resampled"].any()
    assert zg == pytest.approx(zh, rel=1e-5)
    np.testing.assert_allclose(og["logIncrements"], oh["logIncrements"], rtol=1e-4, atol=1e-4)


def test_generated_init_and_obs_samples_equal_handwritten(ctx, orc, torch_gpu):
    """bihip_init_samples and bihip_obs_sample of PZ.bi equal the hand-written model's."""
    g, h = _pz_pair(capi.F64)
    P = 3000
    L = capi.lib()
    out = []
    for m in (g, h):
        X = empty((2, P), torch_gpu.float64)
        capi.check(L.bihip_init_samples(ctx.h, ctypes.byref(m), 21, 3, P,
                                        ctypes.c_void_p(X.data_ptr()), P, None))
        Y = empty((P, 1), torch_gpu.float64)
        capi.check(L.bihip_obs_sample(ctx.h, ctypes.byref(m), 21, 3, 5, P,
                                      ctypes.c_void_p(X.data_ptr()), P,
                                      ctypes.c_void_p(Y.data_ptr())))
        out.append((host(X), host(Y)))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-15, atol=0)


def test_generated_lg_matches_oracle_and_kalman(ctx, orc):
    """LinearGaussian.bi (DOPRI5, then additive noise) against the oracle's LG
    filter, and its likelihood against the Kalman filter's."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, atol=1e-10, rtol=1e-10,
                             h0=0.25, **par)
    h = capi.make_model("lg", precision=capi.F64, atol=1e-10, rtol=1e-10, h0=0.25, **par)
    om = oracle_model(orc, h)
    x0 = orc.init_samples(om, 11, 1)[:, 0]
    _, y = orc.simulate(om, 11, 40, 1.0, x0)
    z, out, anc = _run(ctx, g, y, 4096, 1.0, 1, ess_rel=0.5)
    ref = orc.pf_run(om, y, 4096, 1.0, 1, ess_rel=0.5, want_ancestors=True)
    assert z == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_array_equal(anc, ref["ancestors"])
    exact = orc.kalman_loglik_lg(om, 1.0, y)
    zs = [_run(ctx, g, y, 1 << 16, 1.0, s, record=False)[0] for s in (1, 2)]
    assert abs(np.mean(zs) - exact) < 0.05, (zs, exact)


def _ar1_kalman(y, theta, sigma2, obs_sd, fires):
    """Exact log-likelihood of x <- theta x + w (w ~ N(0, sigma2)) applied at
    the intervals flagged in fires, y_k ~ N(x, obs_sd^2) after interval k,
    x_0 ~ N(0, 1)."""
    m, v, ll = 0.0, 1.0, 0.0
    for k, yk in enumerate(y):
        if fires[k]:
            m, v = theta * m, theta * theta * v + sigma2
        s = v + obs_sd * obs_sd
        ll += -0.5 * np.log(2 * np.pi * s) - 0.5 * (yk - m) ** 2 / s
        gain = v / s
        m, v = m + gain * (yk - m), (1 - gain) * v
    return ll


def _ar1_data(T, theta, sigma2, obs_sd, fires, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal()
    y = np.empty(T)
    for k in range(T):
        if fires[k]:
            x = theta * x + np.sqrt(sigma2) * rng.normal()
        y[k] = x + obs_sd * rng.normal()
    return y


@pytest.mark.parametrize("dt", [1.0, 0.5])
def test_generated_ar1_matches_kalman(ctx, dt):
    """AR1.bi: a discrete-time model (no ode block) with delta = 1.  At dt = 0.5
    its actions fire at every other interval only (times that are multiples of
    delta, share/src/bi/state/Schedule.hpp:230-292); the Kalman filter then
    predicts with the identity between them."""
    theta, sigma2, obs_sd = 0.8, 0.5, 0.5
    T = 40 if dt == 1.0 else 80
    fires = [((k * dt) % 1.0) == 0.0 for k in range(T)]
    y = _ar1_data(T, theta, sigma2, obs_sd, fires, seed=5)
    exact = _ar1_kalman(y, theta, sigma2, obs_sd, fires)
    m = capi.generated_model("AR1", precision=capi.F64, theta=theta, sigma2=sigma2,
                             obs_sd=obs_sd)
    zs = [_run(ctx, m, y, 1 << 16, dt, s, record=False)[0] for s in (1, 2)]
    assert abs(np.mean(zs) - exact) < 0.05, (zs, exact)


def test_generated_batch_equals_single_filter(ctx, orc):
    """SMC^2 inner filters on PZ.bi: filter j of a batch (theta_j = (mu, sigma))
    equals bihip_pf run alone with batch id j, and the oracle's filter."""
    T = 15
    y = _pz_data(orc, T)
    g, _ = _pz_pair(capi.F64)
    Ns = [300, 2048, 7000, 1]
    rng = np.random.default_rng(4)
    theta = np.stack([np.exp(rng.normal(np.log(0.5), 0.2, len(Ns))),
                      np.exp(rng.normal(np.log(0.2), 0.2, len(Ns)))], axis=1)
    bf = capi.BatchFilter(ctx, g, Ns, theta, T, 1.0, seed=5, ess_rel=0.6, record=True)
    bf.set_observations(y)
    bf.init()
    bf.step(T)
    out = bf.outputs()
    assert (out["status"] == 0).all()
    for j, N in enumerate(Ns):
        m, h = _pz_pair(capi.F64, mu=theta[j, 0], sigma=theta[j, 1])
        z, o, _ = _run(ctx, m, y, N, 1.0, 5, ess_rel=0.6, record=False, batch=j)
        np.testing.assert_allclose(out["logIncrements"][j], o["logIncrements"], rtol=1e-10,
                                   atol=1e-10)
        assert out["logLikelihood"][j] == pytest.approx(z, rel=1e-10)
        ref = orc.pf_run(oracle_model(orc, h), y, N, 1.0, 5, ess_rel=0.6, batch=j,
                         want_ancestors=True)
        for k in range(T):
            np.testing.assert_array_equal(bf.ancestors(j, k), ref["ancestors"][k], f"{j} {k}")
    bf.close()


def test_generated_model_rejects_flow(ctx, torch_gpu):
    g, _ = _pz_pair(capi.F64)
    X = empty((2, 4), torch_gpu.float64)
    rc = capi.lib().bihip_flow(ctx.h, ctypes.byref(g), 0.0, 1.0, 4,
                               ctypes.c_void_p(X.data_ptr()), 4, None, None)
    assert rc == capi.EINVAL


def _l96_pair(prec):
    par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
    g = capi.generated_model("Lorenz96", precision=prec, **par)
    h = capi.make_model("l96", d=8, precision=prec, atol=g.atol, rtol=g.rtol, h0=g.h0, **par)
    return g, h


def test_generated_lorenz96_matches_oracle(ctx, orc):
    """Lorenz96.bi (eight components over a cyclic dimension, DOPRI5, vector
    noise, observations of the even components) against the oracle's
    Lorenz-96 filter: FP64, N = 4096, T = 40, ESS < 0.5 N."""
    g, h = _l96_pair(capi.F64)
    om = oracle_model(orc, h)
    x0 = orc.init_samples(om, 3, 1, mean0=np.full(8, 1.0))[:, 0]
    _, y = orc.simulate(om, 3, 40, 0.05, x0)
    z, out, anc = _run(ctx, g, y, 4096, 0.05, 9, ess_rel=0.5)
    ref = orc.pf_run(om, y, 4096, 0.05, 9, ess_rel=0.5, mean0=np.zeros(8), want_ancestors=True)
    assert out["resampled"].any()
    assert z == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_generated_lorenz96_fp32_equals_handwritten(ctx, orc):
    """FP32: the generated one-lane functor against the hand-written
    two-lanes-per-particle Lorenz-96 kernel, no resampling, logZ to 1e-5."""
    g, h = _l96_pair(capi.F32)
    om = oracle_model(orc, h)
    x0 = orc.init_samples(om, 3, 1, mean0=np.full(8, 1.0))[:, 0]
    _, y = orc.simulate(om, 3, 10, 0.05, x0)
    T, P = len(y), 4096
    pf = capi.ParticleFilter(ctx, h, P, T, 0.05, 4, ess_rel=0.0)
    pf.set_observations(y)
    pf.init(np.zeros(8))
    pf.step(T)
    zh = pf.term()
    pf.close()
    zg, og, _ = _run(ctx, g, y, P, 0.05, 4, ess_rel=0.0, record=False)
    assert not og["resampled"].any()
    assert zg == pytest.approx(zh, rel=1e-5)


def _lg_kalman(y, lam, noise_sd, obs_sd, m0, s0, dt, fires):
    """Exact log-likelihood of dx/dt = -lam x over every interval of length dt,
    plus N(0, noise_sd^2) at the intervals flagged in fires, y_k ~ N(x, obs_sd^2)."""
    a = np.exp(-lam * dt)
    m, v, ll = m0, s0 * s0, 0.0
    for k, yk in enumerate(y):
        m, v = a * m, a * a * v + (noise_sd ** 2 if fires[k] else 0.0)
        s = v + obs_sd * obs_sd
        ll += -0.5 * np.log(2 * np.pi * s) - 0.5 * (yk - m) ** 2 / s
        gain = v / s
        m, v = m + gain * (yk - m), (1 - gain) * v
    return ll


def test_generated_ode_with_delta_between_observations(ctx):
    """LinearGaussian.bi (delta = 1) observed every 0.5: the ode runs over every
    interval, the noise fires only on those starting at whole times."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, atol=1e-10, rtol=1e-10,
                             h0=0.25, **par)
    T, dt = 60, 0.5
    fires = [(k % 2) == 0 for k in range(T)]
    rng = np.random.default_rng(8)
    x, y = rng.normal(), np.empty(T)
    for k in range(T):
        x = np.exp(-0.5 * dt) * x + (0.5 * rng.normal() if fires[k] else 0.0)
        y[k] = x + 0.5 * rng.normal()
    exact = _lg_kalman(y, 0.5, 0.5, 0.5, 0.0, 1.0, dt, fires)
    zs = [_run(ctx, g, y, 1 << 16, dt, s, record=False)[0] for s in (1, 2)]
    assert abs(np.mean(zs) - exact) < 0.05, (zs, exact)


def _sir_counts(T, beta, gamma, k, sigma, seed):
    """Incidence counts simulated from SIR.bi's equations in numpy (RK4 h = 0.1)."""
    rng = np.random.default_rng(seed)
    N, x = 1000.0, np.array([990.0, 10.0, 0.0, 0.0])

    def f(x, w):
        inf = beta * np.exp(w) * x[0] * x[1] / N
        return np.array([-inf, inf - gamma * x[1], gamma * x[1], inf])

    y = np.empty(T)
    for t in range(T):
        w = sigma * rng.normal()
        x[3] = 0.0
        for _ in range(10):
            k1 = f(x, w)
            k2 = f(x + 0.05 * k1, w)
            k3 = f(x + 0.05 * k2, w)
            k4 = f(x + 0.1 * k3, w)
            x = x + 0.1 / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        y[t] = rng.negative_binomial(k, k / (k + x[3]))
    return y


def test_generated_sir_negbin_counts(ctx, torch_gpu):
    """SIR.bi: the negative-binomial count density on the device against
    scipy.stats.nbinom (mean Z, shape k), then the filter on counts simulated
    from the model: two independent estimates of the log-likelihood agree."""
    from scipy import stats
    par = dict(beta=0.6, gamma=0.25, k=8.0, sigma=0.1)
    m = capi.generated_model("SIR", precision=capi.F64, **par)
    rng = np.random.default_rng(2)
    P = 2000
    X0 = np.stack([rng.uniform(100, 900, P), rng.uniform(1, 300, P), rng.uniform(0, 500, P),
                   rng.uniform(0.5, 60.0, P)])
    X = torch_gpu.from_numpy(X0.copy()).cuda()
    for yv in (0.0, 3.0, 25.0, 80.0):
        lws = torch_gpu.zeros(P, dtype=torch_gpu.float64, device="cuda")
        yd = torch_gpu.tensor([yv], dtype=torch_gpu.float64, device="cuda")
        capi.check(capi.lib().bihip_obs_logdensity(ctx.h, ctypes.byref(m),
                                                   ctypes.c_void_p(yd.data_ptr()), P,
                                                   ctypes.c_void_p(X.data_ptr()), P,
                                                   ctypes.c_void_p(lws.data_ptr())))
        ref = stats.nbinom.logpmf(yv, 8.0, 8.0 / (8.0 + X0[3]))
        np.testing.assert_allclose(host(lws), ref, rtol=1e-12, atol=1e-12)
    y = _sir_counts(30, seed=11, **par)
    zs = [_run(ctx, m, y, 1 << 16, 1.0, s, record=False) for s in (1, 2)]
    assert all(np.isfinite(z) for z, _, _ in zs)
    assert abs(zs[0][0] - zs[1][0]) < 0.5, [z for z, _, _ in zs]
    assert (zs[0][1]["ess"] > 100).all()


def test_generated_model_with_schedule_and_inputs(ctx):
    """LinearGaussian.bi over an irregular schedule with unobserved steps and
    per-step inputs: the ode decays over each interval, the noise fires at the
    whole times inside it (delta = 1), the decay rate of step k is lam_k; the
    likelihood is the Kalman filter's over the same pieces."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, atol=1e-10, rtol=1e-10,
                             h0=0.25, **par)
    rng = np.random.default_rng(13)
    T = 50
    times = np.concatenate([[0.0], np.cumsum(rng.uniform(0.3, 1.6, T))])
    observed = (rng.uniform(size=T) > 0.25).astype(np.int32)
    lams = rng.uniform(0.2, 0.9, T)

    def pieces(t1, t2):  # [a, b) parts of [t1, t2) split at whole times; fires at whole starts
        out, a = [], t1
        while a < t2:
            b = min(np.floor(a) + 1.0, t2)
            out.append((a, b, a == np.floor(a)))
            a = b
        return out

    # truth and observations, then the exact likelihood
    x, y = rng.normal(), np.zeros(T)
    for k in range(T):
        for a, b, on in pieces(times[k], times[k + 1]):
            x = np.exp(-lams[k] * (b - a)) * x + (0.5 * rng.normal() if on else 0.0)
        y[k] = x + 0.5 * rng.normal()
    m, v, exact = 0.0, 1.0, 0.0
    for k in range(T):
        for a, b, on in pieces(times[k], times[k + 1]):
            e = np.exp(-lams[k] * (b - a))
            m, v = e * m, e * e * v + (0.25 if on else 0.0)
        if observed[k]:
            s = v + 0.25
            exact += -0.5 * np.log(2 * np.pi * s) - 0.5 * (y[k] - m) ** 2 / s
            gain = v / s
            m, v = m + gain * (y[k] - m), (1 - gain) * v
    zs = []
    for seed in (1, 2):
        pf = capi.ParticleFilter(ctx, g, 1 << 16, T, 1.0, seed, ess_rel=0.5)
        pf.set_schedule(times, observed)
        pf.set_inputs(pf.inputs_from(lam=lams))
        pf.set_observations(y)
        pf.init()
        pf.step(T)
        zs.append(pf.term())
        pf.close()
    assert abs(np.mean(zs) - exact) < 0.05, (zs, exact)


def test_smc2_with_the_model_files_prior_matches_quadrature(ctx, orc):
    """SMC^2 over lam of LinearGaussian.bi with its parameter block's prior
    (lam ~ gamma(2, 0.25)): the evidence and the posterior mean of lam against
    quadrature of the exact Kalman likelihood times that prior."""
    from scipy import stats
    from scipy.special import logsumexp
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, atol=1e-10, rtol=1e-10,
                             h0=0.25, **par)
    h = capi.make_model("lg", precision=capi.F64, atol=1e-10, rtol=1e-10, h0=0.25, **par)
    om = oracle_model(orc, h)
    T = 40
    x0 = orc.init_samples(om, 21, 1)[:, 0]
    _, y = orc.simulate(om, 21, T, 1.0, x0)
    y = np.asarray(y).reshape(T, 1)
    # quadrature over u = log lam
    u = np.linspace(-7.0, 2.0, 1500)
    lam = np.exp(u)
    ll = []
    for v in lam:
        om.par[0] = v
        ll.append(orc.kalman_loglik_lg(om, 1.0, y[:, 0]))
    lf = np.array(ll) + stats.gamma(2.0, scale=0.25).logpdf(lam) + u
    du = u[1] - u[0]
    lZq = logsumexp(lf) + np.log(du)
    post = np.sum(np.exp(lf - lZq) * lam) * du
    zs, ps = [], []
    for seed in (1, 2, 3, 4):
        s = capi.SMC2(ctx, g, [512] * 512, T, None, None, [0.3], seed=seed, nmoves=1,
                      ess_rel_outer=0.8, block=256)
        s.set_observations(y)
        s.run()
        out = s.outputs()
        s.close()
        w = np.exp(out["lw"] - out["lw"].max())
        zs.append(out["logZ"])
        ps.append((w * out["theta"][:, 0]).sum() / w.sum())
        assert out["accepts"].sum() > 100
    assert abs(np.mean(zs) - lZq) < 0.15, (zs, lZq)
    assert abs(np.mean(ps) - post) < 0.03, (ps, post)


def test_pmmh_with_the_model_files_prior(ctx, orc):
    """PMMH chains over lam of LinearGaussian.bi start from the keyed draws of
    its parameter block's prior, stay in its support and move."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, **par)
    h = capi.make_model("lg", precision=capi.F64, **par)
    om = oracle_model(orc, h)
    T, B, C = 20, 64, 60
    x0 = orc.init_samples(om, 5, 1)[:, 0]
    _, y = orc.simulate(om, 5, T, 1.0, x0)
    p = capi.PMMH(ctx, g, [512] * B, T, C, None, None, [0.3], seed=3)
    p.set_observations(np.asarray(y).reshape(T, 1))
    p.run()
    out = p.outputs()
    p.close()
    th = out["theta"][:, :, 0]
    first = np.array([capi.prior_sample(g, 1, 3, j)[0] for j in range(B)])
    np.testing.assert_allclose(th[0], first, rtol=1e-15)
    assert (th > 0).all() and np.isfinite(out["ll"]).all()
    assert out["accepted"].sum() > B


def _chi2_counts(x, d):
    """p-value of a chi-square test of integer draws x against scipy distribution d."""
    from scipy import stats
    n = len(x)
    lo, hi = int(d.ppf(1e-4)), int(d.ppf(1 - 1e-4))
    ks = np.arange(lo, hi + 1)
    obs = np.array([np.sum(x == k) for k in ks], float)
    obs[0] += np.sum(x < lo)
    obs[-1] += np.sum(x > hi)
    exp = d.pmf(ks) * n
    exp[0] += d.cdf(lo - 1) * n
    exp[-1] += d.sf(hi) * n
    keep = exp > 5
    return stats.chi2.sf(np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep]), keep.sum() - 1)


def test_generated_count_draws_on_the_device(ctx, torch_gpu):
    """Rejection samplers in device code: SIR.bi's negative-binomial and
    SIRStochastic.bi's binomial observation samples, and SIRStochastic.bi's
    binomial infections and recoveries over one step, against scipy."""
    from scipy import stats
    L, P = capi.lib(), 20000
    c = lambda t: ctypes.c_void_p(t.data_ptr())

    def obs_samples(m, X0):
        X = torch_gpu.from_numpy(np.ascontiguousarray(X0)).cuda()
        Y = torch_gpu.empty((P, 1), dtype=torch_gpu.float64, device="cuda")
        capi.check(L.bihip_obs_sample(ctx.h, ctypes.byref(m), 9, 0, 3, P, c(X), P, c(Y)))
        return host(Y)[:, 0]

    sir = capi.generated_model("SIR", precision=capi.F64, beta=0.6, gamma=0.25, k=8.0, sigma=0.1)
    X0 = np.tile(np.array([[700.0], [100.0], [200.0], [30.0]]), (1, P))
    y = obs_samples(sir, X0)
    assert _chi2_counts(y, stats.nbinom(8.0, 8.0 / 38.0)) > 1e-4
    st = capi.generated_model("SIRStochastic", precision=capi.F64, beta=0.5, gamma=0.3, rho=0.4)
    X0 = np.tile(np.array([[900.0], [60.0], [40.0], [80.0]]), (1, P))
    y = obs_samples(st, X0)
    assert _chi2_counts(y, stats.binom(80, 0.4)) > 1e-4
    Xin = torch_gpu.from_numpy(np.ascontiguousarray(X0)).cuda()
    Xout = torch_gpu.empty_like(Xin)
    capi.check(L.bihip_transition(ctx.h, ctypes.byref(st), 5, 0, 4, 3.0, 4.0, P, c(Xin),
                                  c(Xout), P, None, None))
    X1 = host(Xout)
    inf, rec = 900.0 - X1[0], X1[2] - 40.0
    np.testing.assert_array_equal(X1[3], inf)
    np.testing.assert_array_equal(X1[0] + X1[1] + X1[2], np.full(P, 1000.0))
    assert _chi2_counts(inf, stats.binom(900, 1 - np.exp(-0.5 * 60 / 1000))) > 1e-4
    assert _chi2_counts(rec, stats.binom(60, 1 - np.exp(-0.3))) > 1e-4


def test_generated_stochastic_sir_filter(ctx):
    """SIRStochastic.bi filtered on cases simulated from its own law in numpy:
    two independent estimates of the log-likelihood agree."""
    rng = np.random.default_rng(17)
    beta, gamma, rho, T = 0.5, 0.3, 0.4, 30
    S, I, R = 990.0, 10.0, 0.0
    y = np.empty(T)
    for t in range(T):
        inf = rng.binomial(int(S), 1 - np.exp(-beta * I / 1000.0))
        rec = rng.binomial(int(I), 1 - np.exp(-gamma))
        S, I, R = S - inf, I + inf - rec, R + rec
        y[t] = rng.binomial(inf, rho)
    m = capi.generated_model("SIRStochastic", precision=capi.F64, beta=beta, gamma=gamma, rho=rho)
    zs = [_run(ctx, m, y, 1 << 16, 1.0, s, record=False)[0] for s in (1, 2)]
    assert all(np.isfinite(zs)) and abs(zs[0] - zs[1]) < 0.5, zs


def test_generated_lorenz96_d40_lane_groups_match_oracle(ctx, orc):
    """Lorenz96d40.bi runs four lanes per particle with the stencil's
    neighbours exchanged between lanes: FP64 against the oracle's Lorenz-96
    filter (N = 4096, T = 20, ESS < 0.5 N), ancestors bit-exact."""
    from gpu_helpers import l96_spun_up
    par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
    g = capi.generated_model("Lorenz96d40", precision=capi.F64, **par)
    h = capi.make_model("l96", d=40, precision=capi.F64, atol=g.atol, rtol=g.rtol, h0=g.h0, **par)
    om = oracle_model(orc, h)
    _, y = orc.simulate(om, 7, 20, 0.05, l96_spun_up(orc))
    z, out, anc = _run(ctx, g, y, 4096, 0.05, 9, ess_rel=0.5)
    ref = orc.pf_run(om, y, 4096, 0.05, 9, ess_rel=0.5, mean0=np.zeros(40), want_ancestors=True)
    assert out["resampled"].any()
    assert z == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-9, atol=1e-8)
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_generated_lorenz96_d40_fp32_and_samples(ctx, orc, torch_gpu):
    """FP32 against the hand-written four-lane kernel (no resampling, logZ to
    1e-5); initial and observation samples of the lane layout equal the
    hand-written model's."""
    from gpu_helpers import l96_spun_up
    par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
    g = capi.generated_model("Lorenz96d40", precision=capi.F32, **par)
    h = capi.make_model("l96", d=40, precision=capi.F32, atol=g.atol, rtol=g.rtol, h0=g.h0, **par)
    om = oracle_model(orc, h)
    _, y = orc.simulate(om, 7, 10, 0.05, l96_spun_up(orc))
    P, T = 4096, 10
    pf = capi.ParticleFilter(ctx, h, P, T, 0.05, 4, ess_rel=0.0)
    pf.set_observations(y)
    pf.init(np.zeros(40))
    pf.step(T)
    zh = pf.term()
    pf.close()
    zg, og, _ = _run(ctx, g, y, P, 0.05, 4, ess_rel=0.0, record=False)
    assert not og["resampled"].any()
    assert zg == pytest.approx(zh, rel=1e-5)
    L, c = capi.lib(), (lambda t: ctypes.c_void_p(t.data_ptr()))
    m0 = np.zeros(40)
    outs = []
    for m in (g, h):
        X = empty((40, P), torch_gpu.float32)
        capi.check(L.bihip_init_samples(ctx.h, ctypes.byref(m), 21, 3, P, c(X), P,
                                        m0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        Y = empty((P, 20), torch_gpu.float64)
        capi.check(L.bihip_obs_sample(ctx.h, ctypes.byref(m), 21, 3, 5, P, c(X), P, c(Y)))
        outs.append((host(X), host(Y)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("which", ["PZ", "Lorenz96d40-f32", "Lorenz96d40-f64"])
def test_generated_noise_recorded_as_handwritten(ctx, orc, tmp_path, which):
    """With record="paths" a generated model's noise variables reach the NetCDF
    file under its names: PZ.bi's alpha and Lorenz96d40.bi's w[n] (four lanes
    per particle, each storing its ten elements) equal the hand-written
    functors' recorded noise bit for bit (the same keyed Philox draws), and
    recording changes no result of the generated filter."""
    from scipy.io import netcdf_file
    from gpu_helpers import l96_spun_up
    if which == "PZ":
        g, h = _pz_pair(capi.F64)
        y = _pz_data(orc, 6)
        dt, P, name, dims, mean0 = 1.0, 500, "alpha", ("nr", "np"), None
    else:
        prec = capi.F32 if which.endswith("f32") else capi.F64
        par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
        g = capi.generated_model("Lorenz96d40", precision=prec, **par)
        h = capi.make_model("l96", d=40, precision=prec, atol=g.atol, rtol=g.rtol, h0=g.h0,
                            **par)
        _, y = orc.simulate(oracle_model(orc, h), 7, 6, 0.05, l96_spun_up(orc))
        dt, P, name, dims, mean0 = 0.05, 1000, "w", ("nr", "n", "np"), np.zeros(40)
    T = len(y)
    runs, noise = {}, {}
    for key, m, rec in (("g0", g, False), ("g", g, "paths"), ("h", h, "paths")):
        pf = capi.ParticleFilter(ctx, m, P, T, dt, 11, ess_rel=0.5, record=rec)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(T)
        X, _, lws = pf.state()
        runs[key] = (pf.outputs()["logIncrements"], X[:, :P].cpu().numpy(), lws.cpu().numpy())
        if rec:
            pf.write_netcdf(tmp_path / f"{key}.nc")
            f = netcdf_file(str(tmp_path / f"{key}.nc"), "r", mmap=False)
            assert f.variables[name].dimensions == dims
            noise[key] = np.array(f.variables[name].data)
            f.close()
        pf.close()
    for a, b in zip(runs["g0"], runs["g"]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(noise["g"][1:]).max() > 0
    np.testing.assert_array_equal(noise["g"], noise["h"])


@pytest.mark.gpu
def test_generated_lane_group_model_rejected_by_batch(ctx):
    """The batched inner filters take one-lane models only: a lane-group model
    is refused at create time with an error, not run."""
    g = capi.generated_model("Lorenz96d40", precision=capi.F64, F=8.0, noise_sd=0.1,
                             obs_sd=1.0, init_sd=1.0)
    with pytest.raises(capi.BihipError):
        capi.BatchFilter(ctx, g, [256, 256], np.zeros((2, 0)), 5, 0.05, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [capi.F32, capi.F64])
def test_generated_one_delta_path_equals_sub_interval_loop(ctx, orc, prec):
    """Steps of one delta run the generated transition without its
    sub-interval loop; the same times given as a schedule take the loop.
    Both give the same log-likelihood, states and ancestors, bit for bit
    (Lorenz96d40.bi over lane groups, FP32 and FP64)."""
    from gpu_helpers import l96_spun_up
    par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
    g = capi.generated_model("Lorenz96d40", precision=prec, **par)
    h = capi.make_model("l96", d=40, precision=capi.F64, atol=g.atol, rtol=g.rtol, h0=g.h0, **par)
    _, y = orc.simulate(oracle_model(orc, h), 7, 12, 0.05, l96_spun_up(orc))
    P, T = 2048, 12
    res = []
    for sched in (False, True):
        pf = capi.ParticleFilter(ctx, g, P, T, 0.05, 5, ess_rel=0.5, record=True)
        pf.set_observations(y)
        if sched:
            pf.set_schedule(0.05 * np.arange(T + 1))
        pf.init()
        pf.step(T)
        X, _, lw = pf.state()
        res.append((pf.term(), host(X), host(lw),
                    np.array([pf.ancestors(k) for k in range(T)])))
        pf.close()
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        np.testing.assert_array_equal(a, b)
