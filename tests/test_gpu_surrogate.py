"""The fused surrogate flow (bihip_surrogate_flow) and the filter step that
uses it (bihip_pf_set_surrogate): the fused kernel against the GEMM path and
the oracle's MLP, the filter step against its composition (flow, Philox
noise, observation density), and a trained surrogate against the exact
integrator."""
import numpy as np
import pytest

from libbi_amd import capi

pytestmark = pytest.mark.gpu

D = 40


def _norm():
    return capi.SurrogateNorm(x_shift=2.3, x_scale=3.6, par_shift=8.0, par_scale=1.0,
                              dx_scale=0.3)


def _states(P, seed=0):
    rng = np.random.default_rng(seed)
    return (2.3 + 3.6 * rng.normal(size=(D, P))).astype(np.float32)


@pytest.mark.parametrize("P", [256, 1000])
def test_fused_flow_matches_gemm_path_and_oracle(ctx, torch_gpu, orc, P):
    torch = torch_gpu
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=11)
    norm, F = _norm(), 8.5
    X = _states(P, seed=P)
    Xd = torch.from_numpy(X).cuda()
    out = net.surrogate_flow(norm, F, Xd).cpu().numpy()
    U = np.concatenate([(X.T - np.float32(2.3)) * np.float32(1 / np.float32(3.6)),
                        np.full((P, 1), 0.5, np.float32)], axis=1)
    U = np.ascontiguousarray(U, dtype=np.float32)
    y_gemm = net.forward(torch.from_numpy(U).cuda()).cpu().numpy()
    y_fused = (out - X).T / np.float32(0.3)
    y_orc = orc.mlp_forward(net.dims, net.get_params().astype(np.float64), U.astype(np.float64),
                            bf16=True)
    scale = np.abs(y_orc).max()
    e1 = np.abs(y_fused - y_gemm).max() / scale
    e2 = np.abs(y_fused - y_orc).max() / scale
    print("P", P, "fused vs gemm", e1, "fused vs oracle", e2, flush=True)
    assert e1 < 5e-3 and e2 < 5e-3
    net.close()


def _filter(ctx, P, T, y, mean0, seed=5, ess_rel=0.0):
    m = capi.make_model("l96", d=D, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                        init_sd=1.0)
    pf = capi.ParticleFilter(ctx, m, P, T, 0.05, seed, ess_rel=ess_rel)
    pf.set_observations(y)
    pf.init(mean0)
    return pf


def test_filter_step_is_flow_noise_and_density(ctx, torch_gpu, orc):
    torch = torch_gpu
    from libbi_amd import data
    P, T, seed = 1000, 3, 5
    mean0 = data.l96_spinup(ctx, d=D)
    gm = capi.make_model("l96", d=D, precision=capi.F32)
    _, y = data.simulate(ctx, gm, mean0, T, 0.05, seed=9)
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=12)
    norm = _norm()
    pf = _filter(ctx, P, T, y, mean0, seed=seed)
    pf.set_surrogate(net, norm)
    Xd, ld, _ = pf.state()
    x = Xd[:, :P].clone()
    lws_ref = np.zeros(P)
    om = orc.make_model("l96", d=D)
    for k in range(T):
        pf.step(1)
        flow = net.surrogate_flow(norm, 8.0, x.contiguous()).cpu().numpy().astype(np.float64)
        z = np.stack([orc.normals(seed, p, k + 1, orc.TAG_NOISE, D, bits=24) for p in range(P)],
                     axis=1)
        x_ref = flow + 0.1 * z
        Xd, ld, lws = pf.state()
        got = Xd[:, :P].cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(got, x_ref, rtol=0, atol=2e-5)
        lws_ref = orc.obs_logdensity(om, y[k], got, lws_ref)
        np.testing.assert_allclose(lws.cpu().numpy()[:P], lws_ref, rtol=1e-9, atol=1e-9)
        x = Xd[:, :P].clone()
    pf.close()
    net.close()
