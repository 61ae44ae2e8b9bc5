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
