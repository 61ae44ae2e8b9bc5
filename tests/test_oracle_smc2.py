"""Pins the oracle's SMC^2 and PMMH samplers (MarginalSIR over log theta with inner
bootstrap filters, share/src/bi/sampler/MarginalSIR.hpp:281-488), CPU only:

- without moves or outer resampling every slot is the plain filter run with
  its own theta and random-stream batch id, and logZ is the outer
  logsumexp(lw) - log B;
- one outer resampling followed by the final move, rebuilt from the pinned
  primitives (filter runs, systematic offspring, host permutation) and the
  keyed draws;
- on the linear-Gaussian model, the evidence and the posterior mean of lambda
  agree with Gauss-Hermite quadrature of the exact Kalman likelihood.
"""
import numpy as np
import pytest
from scipy.special import logsumexp


def _lg(orc):
    return orc.make_model("lg", lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)


def _data(orc, T, seed=7):
    m = _lg(orc)
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return m, np.asarray(y).reshape(T, -1)


def _with_theta(orc, lam):
    m = _lg(orc)
    m.par[0] = lam
    return m


def _prior_theta(orc, seed, B, mu, sd):
    return np.array([mu + sd * orc.normals(seed, j, 0, orc.TAG_THETA, 1)[0] for j in range(B)])


def test_smc2_without_moves_is_independent_filters(orc):
    T, B = 12, 9
    m, y = _data(orc, T)
    mu, sd = np.log(0.5), 0.7
    N = [33, 64, 100, 7, 256, 1, 50, 128, 90]
    c = orc.smc2_config(B, T, [mu], [sd], [0.3], seed=21, nmoves=1, ess_rel_outer=0.0)
    r = orc.smc2_run(m, c, N, y)
    assert r["rc"] == 0 and not r["resampled"].any() and not r["accepts"].any()
    lth = _prior_theta(orc, 21, B, mu, sd)
    np.testing.assert_allclose(r["theta"][:, 0], np.exp(lth), rtol=1e-15)
    for j in range(B):
        ref = orc.pf_run(_with_theta(orc, np.exp(lth[j])), y, N[j], 1.0, 21, ess_rel=0.5,
                         batch=j)
        assert r["lw"][j] == pytest.approx(ref["logIncrements"].sum(), rel=1e-12, abs=1e-12)
    assert r["logZ"] == pytest.approx(logsumexp(r["lw"]) - np.log(B), rel=1e-12)
    w = np.exp(r["lw"] - r["lw"].max())
    np.testing.assert_allclose(r["ess"][-1], w.sum() ** 2 / (w * w).sum(), rtol=1e-12)


def test_smc2_first_resample_and_final_move(orc):
    """T = 1 with essRel_outer = 1: the outer step resamples, then the final
    move runs one proposal filter per slot (MarginalSIR.hpp:301-303)."""
    B, N, seed = 8, 200, 5
    m, y = _data(orc, 1)
    mu, sd, psd = np.log(0.5), 1.0, 0.5
    c = orc.smc2_config(B, 1, [mu], [sd], [psd], seed=seed, nmoves=1, ess_rel_outer=1.0)
    r = orc.smc2_run(m, c, [N] * B, y)
    assert r["resampled"][0] == 1
    lth = _prior_theta(orc, seed, B, mu, sd)
    li = np.array([orc.pf_run(_with_theta(orc, np.exp(lth[j])), y, N, 1.0, seed, batch=j)
                   ["logIncrements"][0] for j in range(B)])
    lW = logsumexp(li) - np.log(B)
    assert r["logZ"] == pytest.approx(lW, rel=1e-12)
    # outer systematic resampling with the keyed offset, host permutation
    a = orc.uniforms(seed, 0, orc.TAG_OUTER, 1)[0]
    Ws, _ = orc.sumexpu_inclusive_scan(li)
    O, rc = orc.systematic_cumulative_offspring(Ws, a)
    assert rc == 0
    anc = orc.cumulative_offspring_to_ancestors_permute(O)
    lp = lambda v: -0.5 * ((v - mu) / sd) ** 2
    expect, nacc = np.empty(B), 0
    for j in range(B):
        cur, ll1 = lth[anc[j]], li[anc[j]]
        prop = cur + psd * orc.normals(seed, j, 1, orc.TAG_PROPOSAL, 1)[0]
        ll2 = orc.pf_run(_with_theta(orc, np.exp(prop)), y, N, 1.0, seed,
                         batch=B * (1 + 0 + 1 * 1) + j)["logIncrements"][0]
        u = orc.uniform_at(seed, j, 1, orc.TAG_ACCEPT, 0)
        acc = np.log(u) < ll2 - ll1 + lp(prop) - lp(cur)
        expect[j] = prop if acc else cur
        nacc += int(acc)
    assert r["accepts"][1] == nacc and 0 < nacc < B
    np.testing.assert_allclose(r["theta"][:, 0], np.exp(expect), rtol=1e-15)
    np.testing.assert_array_equal(r["lw"], np.full(B, lW))


def test_smc2_matches_quadrature_on_lg(orc):
    T, B, N = 60, 64, 256
    m, y = _data(orc, T)
    mu, sd = np.log(0.5), 1.0
    xs, ws = np.polynomial.hermite.hermgauss(120)
    ls = mu + np.sqrt(2.0) * sd * xs
    lik = np.array([orc.kalman_loglik_lg(_with_theta(orc, np.exp(l)), 1.0, y[:, 0]) for l in ls])
    lZq = logsumexp(lik, b=ws / np.sqrt(np.pi))
    post = np.sum(ws * np.exp(lik - lZq) * np.exp(ls)) / np.sqrt(np.pi)
    zs, ps, moved = [], [], 0
    for seed in range(1, 9):
        c = orc.smc2_config(B, T, [mu], [sd], [0.3], seed=seed, nmoves=1, ess_rel_outer=0.8)
        r = orc.smc2_run(m, c, [N] * B, y)
        assert r["rc"] == 0
        w = np.exp(r["lw"] - r["lw"].max())
        zs.append(r["logZ"])
        ps.append((w * r["theta"][:, 0]).sum() / w.sum())
        moved += r["accepts"].sum()
    assert moved > 100
    # over 40 seeds the spread of log Z-hat is 0.20 and of the posterior mean
    # 0.04 at these sizes; the means of 8 runs are allowed ~4 standard errors
    assert abs(np.mean(zs) - lZq) < 0.3, (zs, lZq)
    assert abs(np.mean(ps) - post) < 0.06, (ps, post)


def test_pmmh_first_iterations_rebuilt(orc):
    """MarginalMH: filter the prior draw, then propose, filter with batch id
    B c + j, accept on the keyed uniform (MarginalMH.hpp:217-266)."""
    B, T, N, seed, C = 4, 8, 300, 9, 3
    m, y = _data(orc, T)
    mu, sd, psd = np.log(0.5), 0.7, 0.4
    c = orc.smc2_config(B, T, [mu], [sd], [psd], seed=seed)
    r = orc.pmmh_run(m, c, C, [N] * B, y)
    lth = _prior_theta(orc, seed, B, mu, sd)
    ll = np.array([orc.pf_run(_with_theta(orc, np.exp(lth[j])), y, N, 1.0, seed, batch=j)
                   ["logIncrements"].sum() for j in range(B)])
    np.testing.assert_allclose(r["ll"][0], ll, rtol=1e-12)
    lp = lambda v: -0.5 * ((v - mu) / sd) ** 2
    acc = np.ones(B, int)
    for it in range(1, C):
        for j in range(B):
            prop = lth[j] + psd * orc.normals(seed, j, it, orc.TAG_PROPOSAL, 1)[0]
            l2 = orc.pf_run(_with_theta(orc, np.exp(prop)), y, N, 1.0, seed,
                            batch=B * it + j)["logIncrements"].sum()
            u = orc.uniform_at(seed, j, it, orc.TAG_ACCEPT, 0)
            if np.log(u) < l2 - ll[j] + lp(prop) - lp(lth[j]):
                lth[j], ll[j] = prop, l2
                acc[j] += 1
        np.testing.assert_allclose(r["theta"][it, :, 0], np.exp(lth), rtol=1e-14)
        np.testing.assert_allclose(r["ll"][it], ll, rtol=1e-12)
    np.testing.assert_array_equal(r["accepted"], acc)


def test_pmmh_matches_quadrature_on_lg(orc):
    T, B, N, C = 60, 32, 256, 120
    m, y = _data(orc, T)
    mu, sd = np.log(0.5), 1.0
    xs, ws = np.polynomial.hermite.hermgauss(120)
    ls = mu + np.sqrt(2.0) * sd * xs
    lik = np.array([orc.kalman_loglik_lg(_with_theta(orc, np.exp(l)), 1.0, y[:, 0]) for l in ls])
    lZq = logsumexp(lik, b=ws / np.sqrt(np.pi))
    post = np.sum(ws * np.exp(lik - lZq) * np.exp(ls)) / np.sqrt(np.pi)
    c = orc.smc2_config(B, T, [mu], [sd], [0.25], seed=3)
    r = orc.pmmh_run(m, c, C, [N] * B, y)
    draws = r["theta"][C // 3:, :, 0]
    est = draws.mean()
    # chain means are independent across chains: their spread gives the error
    se = draws.mean(axis=0).std(ddof=1) / np.sqrt(B)
    assert 0.1 < r["accepted"].mean() / C < 0.9
    assert abs(est - post) < 4 * se + 0.01, (est, post, se)
