"""Pins the oracle's model right-hand sides, integrators and observation
densities against closed forms, convergence orders and scipy (CPU only)."""
import math

import numpy as np
import pytest
from scipy import integrate, stats


def test_l96_rhs_fixed_point_and_hand_values(orc):
    m = orc.make_model("l96", d=40)
    np.testing.assert_array_equal(orc.rhs(m, np.full(40, 8.0)), np.zeros(40))
    x = np.arange(40, dtype=float)
    dx = orc.rhs(m, x)
    for i in range(40):
        expect = (x[(i + 1) % 40] - x[(i - 2) % 40]) * x[(i - 1) % 40] - x[i] + 8.0
        assert dx[i] == expect


def test_pz_rhs_hand_values(orc):
    m = orc.make_model("pz", c=0.25, e=0.3, m_l=0.1, m_q=0.1)
    P, Z, alpha = 2.0, 3.0, 0.7
    dx = orc.rhs(m, [P, Z], alpha=alpha)
    assert dx[0] == pytest.approx(alpha * P - 0.25 * P * Z, rel=1e-15)
    assert dx[1] == pytest.approx(0.3 * 0.25 * P * Z - 0.1 * Z - 0.1 * Z * Z, rel=1e-15)


def _decay_model(orc, integrator, **kw):
    return orc.make_model("lg", integrator=integrator, lam=1.0, **kw)


def test_rk4_exact_decay_and_order(orc):
    errs = []
    for h in (0.2, 0.1, 0.05):
        m = _decay_model(orc, orc.RK4, h0=h)
        x, n = orc.flow(m, [1.0], 0.0, 1.0)
        assert n == round(1.0 / h)
        errs.append(abs(x[0] - math.exp(-1.0)))
    slopes = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.8 < s < 4.2 for s in slopes), slopes


def test_rk4_truncates_last_step(orc):
    m = _decay_model(orc, orc.RK4, h0=0.3)
    x, n = orc.flow(m, [1.0], 0.0, 1.0)   # 0.3, 0.3, 0.3 then 0.1
    assert n == 4
    ref = 1.0
    for h in (0.3, 0.3, 0.3, 1.0 - 0.8999999999999999):
        z = -h
        ref *= 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert x[0] == pytest.approx(ref, rel=1e-13)


def _dopri5_two_steps(orc, h, quirk):
    # huge tolerances: every step is accepted and the second is truncated,
    # so exactly two steps of size h cover [0, 2h]
    m = _decay_model(orc, orc.DOPRI5, h0=h, atol=1e10, rtol=1e10, host_quirk=quirk)
    x, n = orc.flow(m, [1.0], 0.0, 2 * h)
    assert n == 2
    return abs(x[0] - math.exp(-2 * h))


def test_dopri5_fifth_order(orc):
    errs = [_dopri5_two_steps(orc, h, 0) for h in (0.4, 0.2, 0.1)]
    slopes = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    # global error O(h^5) over a fixed number of steps shows local order 6
    assert all(5.5 < s < 6.5 for s in slopes), slopes


def test_dopri5_host_quirk_loses_order(orc):
    good = [_dopri5_two_steps(orc, h, 0) for h in (0.4, 0.2, 0.1)]
    bad = [_dopri5_two_steps(orc, h, 1) for h in (0.4, 0.2, 0.1)]
    # k7 taken at x5 makes the reused k1 wrong: the second step drops to low order
    assert all(b > 100 * g for b, g in zip(bad, good))
    slopes = [math.log2(bad[i] / bad[i + 1]) for i in range(2)]
    assert all(s < 4.0 for s in slopes), slopes


def test_dopri5_matches_tolerance_on_decay(orc):
    m = _decay_model(orc, orc.DOPRI5, h0=0.05, atol=1e-9, rtol=1e-9)
    x, n = orc.flow(m, [1.0], 0.0, 2.0)
    assert x[0] == pytest.approx(math.exp(-2.0), rel=1e-8)


def test_dopri5_l96_against_scipy(orc):
    m = orc.make_model("l96", d=40, atol=1e-10, rtol=1e-10, h0=0.05)
    x0 = np.full(40, 8.0)
    x0[0] += 0.01
    x0, _ = orc.flow(m, x0, 0.0, 2.0)  # leave the fixed point
    x1, _ = orc.flow(m, x0, 0.0, 0.05)
    f = lambda t, x: orc.rhs(m, x)
    sol = integrate.solve_ivp(f, (0.0, 0.05), x0, method="DOP853", rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(x1, sol.y[:, -1], rtol=0, atol=1e-8)
    # the benchmark tolerance 1e-5 stays within a few tolerances of the truth
    m5 = orc.make_model("l96", d=40)
    x5, n = orc.flow(m5, x0, 0.0, 0.05)
    assert 1 <= n <= 10
    np.testing.assert_allclose(x5, sol.y[:, -1], rtol=0, atol=1e-3)


def test_dopri5_step_limit(orc):
    m = orc.make_model("l96", d=40, atol=1e-14, rtol=1e-14, nsteps=3)
    x = np.full(40, 8.0)
    x[0] += 1.0
    _, n = orc.flow(m, x, 0.0, 1.0)
    assert n == 3


def test_obs_logdensity_pz_lognormal(orc):
    m = orc.make_model("pz", obs_sd=0.2)
    X = np.array([[1.5, 2.0, 0.7], [1.0, 1.0, 1.0]])
    y = np.array([1.8])
    out = orc.obs_logdensity(m, y, X, np.array([0.5, 0.0, -1.0]))
    expect = stats.lognorm(s=0.2, scale=X[0]).logpdf(1.8) + np.array([0.5, 0.0, -1.0])
    np.testing.assert_allclose(out, expect, rtol=1e-13)


def test_obs_logdensity_l96_gaussian(orc):
    m = orc.make_model("l96", d=8, obs_sd=1.3)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(8, 5))
    y = rng.normal(size=4)
    out = orc.obs_logdensity(m, y, X, np.zeros(5))
    expect = stats.norm(loc=X[0::2], scale=1.3).logpdf(y[:, None]).sum(axis=0)
    np.testing.assert_allclose(out, expect, rtol=1e-13)


def test_init_samples_distribution(orc):
    m = orc.make_model("pz")
    X = orc.init_samples(m, 3, 20000)
    assert abs(np.log(X[0]).mean() - math.log(2.0)) < 0.01
    assert abs(np.log(X[0]).std() - 0.2) < 0.01
    assert abs(np.log(X[1]).std() - 0.1) < 0.005


def test_transition_noise_layout(orc):
    m = orc.make_model("l96", d=8, noise_sd=0.3)
    X = np.tile(np.full(8, 8.0)[:, None], (1, 3))   # fixed point: flow is identity
    X1 = orc.transition(m, 9, 4, 0.0, 0.05, X)
    for p in range(3):
        z = orc.normals(9, p, 4, orc.TAG_NOISE, 8)
        np.testing.assert_allclose(X1[:, p], 8.0 + 0.3 * z, rtol=1e-14)
