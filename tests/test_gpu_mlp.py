"""BF16 MFMA multilayer perceptron (bihip_mlp_*) against the oracle's FP64
MLP with the same bfloat16 rounding points: forward outputs, the loss, and
the backpropagated gradient; and training that reduces the loss."""
import numpy as np
import pytest

from libbi_amd import capi

pytestmark = pytest.mark.gpu

SURR = [41, 256, 256, 256, 40]


def _data(torch, n, din, dout, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, din)).astype(np.float32)
    T = rng.normal(size=(n, dout)).astype(np.float32)
    return X, T, torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()


def test_single_layer_matches_rounded_product(ctx, torch_gpu, orc):
    """One linear layer: the MFMA operand and accumulator layouts give
    X W^T + b exactly as an FP64 product of the bf16-rounded operands."""
    torch = torch_gpu
    dims = [64, 64]
    net = capi.MLP(ctx, dims, seed=3)
    p = net.get_params()
    X, _, Xd, _ = _data(torch, 128, 64, 64, seed=1)
    Y = net.forward(Xd).cpu().numpy()
    W = orc.bf16(p[:64 * 64].reshape(64, 64))
    b = p[64 * 64:].astype(np.float64)
    ref = orc.bf16(X) @ W.T + b
    np.testing.assert_allclose(Y, ref, rtol=1e-5, atol=1e-5)
    print("single layer max err", np.abs(Y - ref).max(), flush=True)
    net.close()


@pytest.mark.parametrize("n", [64, 1000])
def test_forward_matches_oracle(ctx, torch_gpu, orc, n):
    torch = torch_gpu
    net = capi.MLP(ctx, SURR, seed=5)
    p = net.get_params().astype(np.float64)
    X, _, Xd, _ = _data(torch, n, SURR[0], SURR[-1], seed=2)
    Y = net.forward(Xd).cpu().numpy()
    ref = orc.mlp_forward(SURR, p, X.astype(np.float64), bf16=True)
    err = np.abs(Y - ref).max() / np.abs(ref).max()
    print("forward n", n, "max rel err", err, flush=True)
    assert err < 2e-3
    net.close()


def test_gradient_matches_oracle(ctx, torch_gpu, orc):
    torch = torch_gpu
    net = capi.MLP(ctx, SURR, seed=7)
    p = net.get_params().astype(np.float64)
    n = 1000
    X, T, Xd, Td = _data(torch, n, SURR[0], SURR[-1], seed=3)
    loss, g = net.loss_grad(Xd, Td)
    rl, rg = orc.mlp_loss_grad(SURR, p, X.astype(np.float64), T.astype(np.float64), bf16=True)
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print("loss", loss, rl, "grad rel err", rel, flush=True)
    assert loss == pytest.approx(rl, rel=1e-4)
    assert rel < 2e-3
    # per layer, so that a wrong layer cannot hide behind a large one
    off = 0
    for l in range(len(SURR) - 1):
        cnt = SURR[l + 1] * SURR[l] + SURR[l + 1]
        a, b = g[off:off + cnt], rg[off:off + cnt]
        assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b), l
        off += cnt
    # and against exact FP64 arithmetic: bf16 rounding costs well under 5 %
    _, fg = orc.mlp_loss_grad(SURR, p, X.astype(np.float64), T.astype(np.float64), bf16=False)
    assert np.linalg.norm(g - fg) < 0.05 * np.linalg.norm(fg)
    net.close()


def test_training_reduces_loss(ctx, torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(9)
    n, din, dout = 8192, 8, 4
    X = rng.normal(size=(n, din)).astype(np.float32)
    M = rng.normal(size=(din, dout)) / np.sqrt(din)
    T = (np.sin(X @ M) + 0.1 * (X[:, :dout] ** 2)).astype(np.float32)
    net = capi.MLP(ctx, [din, 64, 64, dout], seed=1)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    l0, _ = net.loss_grad(Xd, Td)
    losses = net.train(Xd, Td, batch=256, steps=600, lr=3e-3, seed=2)
    l1, _ = net.loss_grad(Xd, Td)
    print("loss", l0, "->", l1, "minibatch", losses[:3], losses[-3:], flush=True)
    assert np.isfinite(losses).all()
    assert l1 < 0.1 * l0
    net.close()
