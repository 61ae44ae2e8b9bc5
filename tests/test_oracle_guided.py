"""Pins the oracle's guided-proposal step (orc_guided_step): with a network
that outputs zero it is the bootstrap step, and with any network the
exponentiated weight increment is an unbiased estimate of the predictive
likelihood p(y | Phi), which is Gaussian in closed form (CPU only)."""
import numpy as np
import pytest


def _net(orc, dims, seed, scale):
    rng = np.random.default_rng(seed)
    n = orc.mlp_nparams(dims)
    p = np.empty(n)
    off = 0
    for l in range(len(dims) - 1):
        w = dims[l + 1] * dims[l]
        p[off:off + w] = rng.normal(size=w) * scale / np.sqrt(dims[l])
        off += w
        p[off:off + dims[l + 1]] = rng.normal(size=dims[l + 1]) * 0.1
        off += dims[l + 1]
    return p


def test_zero_network_is_the_bootstrap_step(orc):
    d, P = 8, 300
    dims = [d + d // 2 + 1, 16, 2 * d]
    params = np.zeros(orc.mlp_nparams(dims))
    rng = np.random.default_rng(1)
    Phi = rng.normal(2.0, 3.0, size=(d, P))
    y = rng.normal(size=d // 2)
    norm = (1.0, 2.0, 1.0, 8.0, 1.0)
    X, inc = orc.guided_step(dims, params, norm, 8.0, Phi, y, seed=5, step=3, q=0.1)
    z = np.stack([orc.normals(5, p, 3, orc.TAG_NOISE, d, bits=24) for p in range(P)], axis=1)
    np.testing.assert_allclose(X, Phi + 0.1 * z, rtol=0, atol=1e-15)
    m = orc.make_model("l96", d=d, obs_sd=1.0)
    ref = orc.obs_logdensity(m, y, X, np.zeros(P))
    np.testing.assert_allclose(inc, ref, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("bf16", [False, True])
def test_weight_increment_is_unbiased(orc, bf16):
    """E_eps[exp(inc)] = prod_o N(y_o; Phi_{2o}, q^2 + obs_sd^2) for a fixed
    Phi: the proposal's density cancels exactly.  A wrong sign or a dropped
    term of the correction (e.g. the s/2 Jacobian) biases the mean by a
    factor far outside the Monte Carlo error."""
    d, P, q, r = 4, 400_000, 0.6, 1.0
    dims = [d + d // 2 + 1, 24, 2 * d]
    params = _net(orc, dims, 3, 0.8)
    Phi1 = np.array([0.4, -1.1, 0.9, 0.3])
    Phi = np.repeat(Phi1[:, None], P, axis=1)
    y = np.array([0.9, 1.5])
    norm = (0.0, 1.0, 1.0, 8.0, 1.0)
    X, inc = orc.guided_step(dims, params, norm, 8.0, Phi, y, seed=11, step=2, q=q,
                             obs_sd=r, bf16=bf16)
    # the network moves the proposal away from the transition density
    d_ = (X - Phi) / q
    assert np.abs(d_.mean(axis=1)).max() > 0.05
    w = np.exp(inc)
    est, se = w.mean(), w.std() / np.sqrt(P)
    s2 = q * q + r * r
    exact = np.prod(np.exp(-0.5 * (y - Phi1[0::2]) ** 2 / s2) / np.sqrt(2 * np.pi * s2))
    print("estimate", est, "+-", se, "exact", exact)
    assert abs(est - exact) < 4 * se
    assert se < 0.02 * exact
