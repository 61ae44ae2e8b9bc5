"""Pins the oracle's FP64 MLP (the checker of the BF16 MFMA kernels), CPU only:
forward pass against torch float64 with the same weights, the backpropagated
gradient against central finite differences of the loss, and the bfloat16
rounding points against an independent numpy restatement."""
import numpy as np
import pytest

DIMS = [41, 32, 24, 40]


def _params(orc, dims, seed=0, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, orc.mlp_nparams(dims))


def _unpack(dims, p):
    out, off = [], 0
    for l in range(len(dims) - 1):
        W = p[off:off + dims[l + 1] * dims[l]].reshape(dims[l + 1], dims[l])
        off += dims[l + 1] * dims[l]
        b = p[off:off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W, b))
    return out


def test_forward_matches_torch(orc):
    torch = pytest.importorskip("torch")
    p = _params(orc, DIMS)
    X = np.random.default_rng(1).normal(size=(17, DIMS[0]))
    Y = orc.mlp_forward(DIMS, p, X)
    a = torch.tensor(X)
    layers = _unpack(DIMS, p)
    for l, (W, b) in enumerate(layers):
        a = a @ torch.tensor(W).T + torch.tensor(b)
        if l + 1 < len(layers):
            a = a * torch.sigmoid(1.702 * a)  # GELU, sigmoid form
    np.testing.assert_allclose(Y, a.numpy(), rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences(orc):
    dims = [7, 9, 8, 5]
    p = _params(orc, dims, seed=3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(6, dims[0]))
    T = rng.normal(size=(6, dims[-1]))
    _, g = orc.mlp_loss_grad(dims, p, X, T)
    h = 1e-6
    for i in rng.choice(len(p), 60, replace=False):
        q = p.copy()
        q[i] += h
        lp, _ = orc.mlp_loss_grad(dims, q, X, T)
        q[i] -= 2 * h
        lm, _ = orc.mlp_loss_grad(dims, q, X, T)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-9), i


def test_gaussian_loss_gradient_matches_finite_differences(orc):
    """Gaussian negative log-likelihood: outputs (mean, log variance)."""
    dims = [6, 9, 8]  # 4 targets
    p = _params(orc, dims, seed=8, scale=0.3)
    rng = np.random.default_rng(9)
    X = rng.normal(size=(7, dims[0]))
    T = rng.normal(size=(7, 4))
    val, g = orc.mlp_loss_grad(dims, p, X, T, loss=orc.LOSS_GAUSS)
    Y = orc.mlp_forward(dims, p, X)
    mu, lv = Y[:, :4], Y[:, 4:]
    assert val == pytest.approx(np.mean(0.5 * ((T - mu) ** 2 * np.exp(-lv) + lv)), rel=1e-12)
    h = 1e-6
    for i in rng.choice(len(p), 50, replace=False):
        q = p.copy()
        q[i] += h
        lp, _ = orc.mlp_loss_grad(dims, q, X, T, loss=orc.LOSS_GAUSS)
        q[i] -= 2 * h
        lm, _ = orc.mlp_loss_grad(dims, q, X, T, loss=orc.LOSS_GAUSS)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-9), i


def test_bf16_rounding_points(orc):
    """bf16 mode: inputs, weights, hidden activations and back-propagated
    errors rounded to bfloat16, accumulations and biases in FP64."""
    from scipy.special import expit
    p = _params(orc, DIMS, seed=5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(9, DIMS[0]))
    T = rng.normal(size=(9, DIMS[-1]))
    gelu = lambda z: z * expit(1.702 * z)
    dgelu = lambda z: expit(1.702 * z) * (1 + 1.702 * z * (1 - expit(1.702 * z)))
    layers = _unpack(DIMS, p)
    acts, zs = [orc.bf16(X)], []
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ orc.bf16(W).T + b
        zs.append(z)
        if l + 1 < len(layers):
            acts.append(orc.bf16(gelu(z)))
    Y = zs[-1]
    np.testing.assert_allclose(orc.mlp_forward(DIMS, p, X, bf16=True), Y, rtol=1e-12, atol=1e-12)
    n, dout = X.shape[0], DIMS[-1]
    dz = orc.bf16(2.0 * (Y - T) / (n * dout))
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads.append((dz.T @ acts[l], dz.sum(axis=0)))
        if l > 0:
            dz = orc.bf16((dz @ orc.bf16(W)) * dgelu(zs[l - 1]))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads[::-1]])
    loss, g = orc.mlp_loss_grad(DIMS, p, X, T, bf16=True)
    assert loss == pytest.approx(np.mean((Y - T) ** 2), rel=1e-12)
    np.testing.assert_allclose(g, flat, rtol=1e-10, atol=1e-13)


def test_bf16_rounding_is_nearest_even(orc):
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert orc.lib().orc_bf16(1.0 + 2.0 ** -8) == 1.0
    assert orc.lib().orc_bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert orc.lib().orc_bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(0).normal(size=1000)
    np.testing.assert_array_equal(orc.bf16(x), [orc.lib().orc_bf16(v) for v in x])
