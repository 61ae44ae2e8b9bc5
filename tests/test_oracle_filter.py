"""Pins the oracle's bootstrap particle filter: its log-likelihood
conventions (running lws, lws[:] = logLikelihood after resampling, -log P,
logIncrements) against the exact Kalman log-likelihood of a linear-Gaussian
model (CPU only)."""
import numpy as np
import pytest


def _lg(orc):
    return orc.make_model("lg", lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0,
                          atol=1e-10, rtol=1e-10, h0=0.25)


def _data(orc, T, seed=11):
    m = _lg(orc)
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return m, y


@pytest.mark.parametrize("resampler", [0, 1, 2])
@pytest.mark.parametrize("ess_rel", [1.0, 0.5])
def test_pf_matches_kalman(orc, resampler, ess_rel):
    m, y = _data(orc, 40)
    exact = orc.kalman_loglik_lg(m, 1.0, y)
    zs = [orc.pf_run(m, y, 1 << 14, 1.0, seed, resampler=resampler, ess_rel=ess_rel)["logZ"]
          for seed in (1, 2, 3)]
    # sd of log Z-hat at N=2^14 is ~0.02 here; allow a generous 8 sd on the mean
    assert abs(np.mean(zs) - exact) < 0.1, (zs, exact)


def test_pf_without_resampling_short_horizon(orc):
    m, y = _data(orc, 4)
    exact = orc.kalman_loglik_lg(m, 1.0, y)
    r = orc.pf_run(m, y, 1 << 16, 1.0, 5, ess_rel=0.0)
    assert not r["resampled"].any()
    assert abs(r["logZ"] - exact) < 0.05


def test_pf_bookkeeping(orc):
    m, y = _data(orc, 25)
    r = orc.pf_run(m, y, 2048, 1.0, 7, ess_rel=0.5, want_ancestors=True)
    assert r["rc"] == orc.OK
    # term's logZ equals the running logLikelihood = sum of logIncrements
    assert r["logZ"] == pytest.approx(r["logIncrements"].sum(), rel=1e-12)
    # no resampling at t0 (unobserved); afterwards exactly when ess < essRel * P
    assert r["resampled"][0] == 0
    for k in range(1, 25):
        assert r["resampled"][k] == (r["ess"][k - 1] < 0.5 * 2048)
    for k in range(25):
        a = r["ancestors"][k]
        if not r["resampled"][k]:
            np.testing.assert_array_equal(a, np.arange(2048))
        else:
            assert np.all(a[a] == a)


def test_pf_degenerate_weights(orc):
    # An infinite observation gives every particle log-weight -inf: the ESS is
    # 0/0 = NaN and `ess < essRel * P` is false (Resampler.hpp:168), so the
    # filter never reaches the resampler's degenerate branch; the estimate is
    # -inf from then on.  Calling a resampler on such weights reports
    # ParticleFilterDegeneratedException (test_all_minus_inf_is_degenerate).
    m, y = _data(orc, 5)
    y = y.copy()
    y[1] = np.inf
    r = orc.pf_run(m, y, 256, 1.0, 3, ess_rel=1.0)
    assert r["rc"] == orc.OK
    assert np.isnan(r["ess"][1]) and not r["resampled"][2:].any()
    assert r["logZ"] == -np.inf


def test_pz_filter_runs(orc):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, 100, 1)[:, 0]
    _, y = orc.simulate(m, 100, 30, 1.0, x0)
    r = orc.pf_run(m, y, 1024, 1.0, 1, ess_rel=1.0)
    assert r["rc"] == orc.OK and np.isfinite(r["logZ"])
    assert r["resampled"][1:].all()


# ---------------------------------------------------------------- schedules
def _schedule(T, seed=3, p_unobserved=0.3):
    rng = np.random.default_rng(seed)
    times = np.concatenate([[0.0], np.cumsum(rng.uniform(0.3, 1.7, size=T))])
    observed = (rng.uniform(size=T) > p_unobserved).astype(np.int32)
    observed[0] = 1
    return times, observed


def test_uniform_schedule_equals_plain_run(orc):
    m, y = _data(orc, 20)
    times = np.arange(21, dtype=float)
    a = orc.pf_run(m, y, 1024, 1.0, 5, ess_rel=0.5, want_ancestors=True)
    b = orc.pf_run(m, y, 1024, 1.0, 5, ess_rel=0.5, want_ancestors=True, times=times,
                   observed=np.ones(20, np.int32))
    assert a["logZ"] == b["logZ"]
    np.testing.assert_array_equal(a["logIncrements"], b["logIncrements"])
    np.testing.assert_array_equal(a["ancestors"], b["ancestors"])


def test_scheduled_kalman_is_the_joint_gaussian(orc):
    """The scheduled Kalman log-likelihood equals the log-density of the observed
    y under their joint Gaussian (x_{k+1} = A_k x_k + q e, y_k = x_{k+1} + r e,
    A_k = exp(-lambda (t_{k+1} - t_k))), built directly from the covariances."""
    from scipy.stats import multivariate_normal
    m = _lg(orc)
    lam, q, r, m0, s0 = 0.5, 0.5, 0.5, 0.0, 1.0
    T = 6
    times, observed = _schedule(T, seed=8)
    y = np.random.default_rng(9).normal(size=T)
    A = np.exp(-lam * np.diff(times))
    # x_{k+1} = (prod A) x0 + sum_j (prod_{i>j} A_i) q e_j
    C = np.zeros((T, T + 1))  # x_{k+1} in terms of (x0, e_0 .. e_{T-1})
    for k in range(T):
        C[k, 0] = np.prod(A[:k + 1])
        for j in range(k + 1):
            C[k, 1 + j] = np.prod(A[j + 1:k + 1]) * q
    S = np.diag([s0 ** 2] + [1.0] * T)
    cov_x = C @ S @ C.T
    mean_x = C[:, 0] * m0
    idx = np.flatnonzero(observed)
    cov_y = cov_x[np.ix_(idx, idx)] + r ** 2 * np.eye(len(idx))
    exact = multivariate_normal(mean_x[idx], cov_y).logpdf(y[idx])
    got = orc.kalman_loglik_lg_sched(m, times, observed, y)
    assert got == pytest.approx(exact, rel=1e-10)
    assert orc.kalman_loglik_lg_sched(m, np.arange(T + 1.0), np.ones(T), y) == pytest.approx(
        orc.kalman_loglik_lg(m, 1.0, y), rel=1e-12)


@pytest.mark.parametrize("ess_rel", [1.0, 0.5])
def test_scheduled_pf_matches_kalman(orc, ess_rel):
    """Irregular observation times and unobserved steps: the particle filter's
    logZ against the scheduled Kalman log-likelihood."""
    m = _lg(orc)
    T = 30
    times, observed = _schedule(T)
    x0 = orc.init_samples(m, 11, 1)[:, 0]
    _, y = orc.simulate(m, 11, T, 1.0, x0)
    exact = orc.kalman_loglik_lg_sched(m, times, observed, y)
    zs = [orc.pf_run(m, y, 1 << 14, 1.0, seed, ess_rel=ess_rel, times=times,
                     observed=observed)["logZ"] for seed in (1, 2, 3)]
    assert abs(np.mean(zs) - exact) < 0.1, (zs, exact)


def test_scheduled_pf_bookkeeping(orc):
    """Unobserved steps add nothing to the log-likelihood and are never followed
    by a resampling; the ESS carries over."""
    m, y = _data(orc, 25)
    times, observed = _schedule(25, seed=4, p_unobserved=0.4)
    r = orc.pf_run(m, y, 2048, 1.0, 7, ess_rel=0.9, want_ancestors=True, times=times,
                   observed=observed)
    assert r["rc"] == orc.OK
    assert r["logZ"] == pytest.approx(r["logIncrements"].sum(), rel=1e-12)
    for k in range(25):
        if not observed[k]:
            assert r["logIncrements"][k] == 0.0
            if k > 0:
                assert r["ess"][k] == r["ess"][k - 1]
        if k >= 1:
            assert r["resampled"][k] == (observed[k - 1] and r["ess"][k - 1] < 0.9 * 2048)
    assert r["resampled"].any()


# ---------------------------------------------------------------- inputs
def test_inputs_equal_to_the_parameters_change_nothing(orc):
    m, y = _data(orc, 15)
    inputs = np.tile(np.array(m.par[:16]), (15, 1))
    a = orc.pf_run(m, y, 1024, 1.0, 5, ess_rel=0.5, want_ancestors=True)
    b = orc.pf_run(m, y, 1024, 1.0, 5, ess_rel=0.5, want_ancestors=True, inputs=inputs)
    assert a["logZ"] == b["logZ"]
    np.testing.assert_array_equal(a["ancestors"], b["ancestors"])


def test_time_varying_decay_matches_kalman(orc):
    """A per-step decay rate lambda_k (an input) on unit intervals is the
    constant-lambda model on intervals lambda_k / lambda: the filter's logZ
    against the scheduled Kalman likelihood on those intervals."""
    m = _lg(orc)
    T = 30
    lam = 0.5
    lam_k = lam * np.random.default_rng(6).uniform(0.2, 3.0, size=T)
    inputs = np.tile(np.array(m.par[:16]), (T, 1))
    inputs[:, 0] = lam_k
    x0 = orc.init_samples(m, 11, 1)[:, 0]
    _, y = orc.simulate(m, 11, T, 1.0, x0)
    times = np.concatenate([[0.0], np.cumsum(lam_k / lam)])
    exact = orc.kalman_loglik_lg_sched(m, times, np.ones(T, np.int32), y)
    zs = [orc.pf_run(m, y, 1 << 14, 1.0, seed, ess_rel=0.5, inputs=inputs)["logZ"]
          for seed in (1, 2, 3)]
    assert abs(np.mean(zs) - exact) < 0.1, (zs, exact)


# ---------------------------------------------------------------- imposed resampling
@pytest.mark.parametrize("resampler", [0, 1, 2])
def test_imposed_own_history_reproduces_the_run(orc, resampler):
    """The filter re-run on its own resampling history (flags and ancestors)
    reproduces its log-likelihood, increments and ESS bit for bit."""
    om = orc.make_model("l96", d=8)
    mean0 = np.full(8, 8.0)
    _, y = orc.simulate(om, 3, 15, 0.05, mean0)
    P = 512
    r = orc.pf_run(om, y, P, 0.05, 9, resampler=resampler, ess_rel=0.7, mean0=mean0,
                   want_ancestors=True)
    assert 3 <= r["resampled"].sum() < 15
    s = orc.pf_run(om, y, P, 0.05, 9, resampler=resampler, ess_rel=0.7, mean0=mean0,
                   imposed_resampled=r["resampled"], imposed_ancestors=r["ancestors"])
    assert s["logZ"] == r["logZ"]
    np.testing.assert_array_equal(s["logIncrements"], r["logIncrements"])
    np.testing.assert_array_equal(s["ess"], r["ess"])
    np.testing.assert_array_equal(s["resampled"], r["resampled"])


def test_imposed_history_is_followed(orc):
    """Imposed flags and ancestors replace the filter's own: no resampling at
    all equals the essRel = 0 run, and another valid ancestor vector at one
    step changes the estimate (the imposed rows are read)."""
    om = orc.make_model("l96", d=8)
    mean0 = np.full(8, 8.0)
    _, y = orc.simulate(om, 3, 12, 0.05, mean0)
    P, T = 256, 12
    none = orc.pf_run(om, y, P, 0.05, 4, ess_rel=0.7, mean0=mean0,
                      imposed_resampled=np.zeros(T, np.int32),
                      imposed_ancestors=np.zeros((T, P), np.int32))
    plain = orc.pf_run(om, y, P, 0.05, 4, ess_rel=0.0, mean0=mean0)
    assert none["logZ"] == plain["logZ"]
    r = orc.pf_run(om, y, P, 0.05, 4, ess_rel=0.7, mean0=mean0, want_ancestors=True)
    k = int(np.nonzero(r["resampled"])[0][0])
    anc = r["ancestors"].copy()
    anc[k] = np.roll(anc[k], 1)  # the same multiset, another layout
    moved = orc.pf_run(om, y, P, 0.05, 4, ess_rel=0.7, mean0=mean0,
                       imposed_resampled=r["resampled"], imposed_ancestors=anc)
    # the layout changes which noise each copy draws, hence the estimate
    assert moved["logZ"] != r["logZ"]
    np.testing.assert_array_equal(moved["logIncrements"][:k], r["logIncrements"][:k])


def test_step_filter_restarts_from_a_given_state(orc):
    """StepFilter.set_state with a run's own state after k steps, then steps
    with the run's resampling imposed, continues the run bit for bit."""
    om = orc.make_model("l96", d=8)
    mean0 = np.full(8, 8.0)
    _, y = orc.simulate(om, 3, 10, 0.05, mean0)
    P, k0 = 300, 4
    full = orc.pf_run(om, y, P, 0.05, 6, ess_rel=0.8, mean0=mean0, want_ancestors=True)
    head = orc.StepFilter(om, P, 0.05, 6, ess_rel=0.8)
    head.init(mean0)
    for k in range(k0):
        head.step(y[k])
    X, lws = head.state()
    ll, ess = head.log_likelihood(), head.ess()
    head.close()
    f = orc.StepFilter(om, P, 0.05, 6, ess_rel=0.8)
    f.set_state(k0, X, lws, ll, ess)
    for k in range(k0, 10):
        rc, li = f.step(y[k], full["resampled"][k], full["ancestors"][k])
        assert rc == orc.OK and li == full["logIncrements"][k]
    assert f.term() == full["logZ"]
    f.close()
