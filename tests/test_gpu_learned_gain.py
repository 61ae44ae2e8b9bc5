"""What the trained networks gain over the method without them.

Guided proposal (Lorenz-96 d=200): trained briefly on the GPU by the
library's own BF16 MFMA kernels, the guided filter keeps a larger effective
sample size than the bootstrap filter at equal N, and its log-likelihood
estimate varies less between seeds.  Both filters estimate the same evidence
(the weights are corrected exactly), so the gain is variance, which is what a
proposal is for.

Surrogate flow (Lorenz-96 d=40): trained on pairs from the exact integrator,
its flow error falls far below the process noise of one interval, from an
untrained network's error far above it."""
import numpy as np
import pytest

from libbi_amd import capi, data, guided, surrogate

pytestmark = pytest.mark.gpu

DT = 0.05


def _filters(ctx, y, mean0, P, seeds, net, norm, T):
    zs, ess = [], []
    for s in seeds:
        pf = capi.ParticleFilter(ctx, guided.model(8.0), P, T, DT, seed=s, ess_rel=0.5)
        pf.set_observations(y)
        pf.init(mean0)
        if net is not None:
            pf.set_guided(net, norm)
        pf.step(T)
        zs.append(pf.term())
        ess.append(pf.outputs()["ess"].mean() / P)
        pf.close()
    return np.array(zs), np.array(ess)


def test_trained_guided_proposal_gains_over_bootstrap(ctx, torch_gpu):
    T, P, seeds = 8, 1 << 15, (1, 2, 3, 4, 5, 6)
    net, norm, info = guided.train(ctx, n=1 << 19, steps=(1500, 1000, 500), seed=3)
    mean0 = data.l96_spinup(ctx, d=guided.D)
    _, y = data.simulate(ctx, guided.model(8.0), mean0, T, DT, seed=31)
    zb, eb = _filters(ctx, y, mean0, P, seeds, None, None, T)
    zg, eg = _filters(ctx, y, mean0, P, seeds, net, norm, T)
    print(f"training {info}; bootstrap: logZ {zb.mean():.2f} sd {zb.std(ddof=1):.3f} ESS "
          f"{eb.mean():.4f}; guided: logZ {zg.mean():.2f} sd {zg.std(ddof=1):.3f} ESS "
          f"{eg.mean():.4f}", flush=True)
    # larger effective sample size at every seed, on average several-fold
    assert (eg > eb).all()
    assert eg.mean() > 3.0 * eb.mean()
    # a less variable evidence estimate
    assert zg.std(ddof=1) < zb.std(ddof=1)
    # the same evidence within the Monte Carlo error of the two
    se = np.sqrt(zb.var(ddof=1) / len(seeds) + zg.var(ddof=1) / len(seeds))
    assert abs(zg.mean() - zb.mean()) < 4 * se + 0.1
    net.close()


def test_trained_surrogate_flow_error_below_process_noise(ctx, torch_gpu):
    X, Y = surrogate.flow_pairs(ctx, 1 << 20, seed=5)
    Xt, Yt = surrogate.flow_pairs(ctx, 1 << 16, seed=99)
    norm0 = surrogate.normalisation(X, Y)
    untrained = capi.MLP(ctx, surrogate.DIMS, seed=7)
    err0, scale = surrogate.flow_error(ctx, untrained, norm0, Xt, Yt)
    net, norm, info = surrogate.train(ctx, steps=(3000, 2000), lrs=(2e-3, 5e-4), seed=7,
                                      pairs=(X, Y))
    err, _ = surrogate.flow_error(ctx, net, norm, Xt, Yt)
    print(f"surrogate flow RMSE on held-out pairs: untrained {err0:.4f}, trained {err:.4f} "
          f"(increments RMS {scale:.3f}, process noise sd 0.1); {info}", flush=True)
    assert err0 > 0.3       # an untrained network is no flow at all
    assert err < 0.05       # half the noise the filter adds every interval
    # the filter's log-likelihood with the surrogate flow against the exact
    # flow's, on one genealogy: the exact FP32 filter's resampling history is
    # imposed on the surrogate filter (same noise, observations, ancestors)
    P, T = 1 << 16, 20
    mean0 = data.l96_spinup(ctx, d=40)
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=41)
    res = {}
    for kind in ("exact", "trained", "untrained"):
        pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=8, ess_rel=0.5, record=kind == "exact")
        pf.set_observations(y)
        if kind != "exact":
            pf.impose_resampling(res["exact"][1]["resampled"], res["exact"][2])
            pf.set_surrogate(net if kind == "trained" else untrained,
                             norm if kind == "trained" else norm0)
        pf.init(mean0)
        pf.step(T)
        out = pf.outputs()
        anc = np.array([pf.ancestors(k) for k in range(T)]) if kind == "exact" else None
        res[kind] = (out["logLikelihood"], out, anc)
        pf.close()
    d_tr = abs(res["trained"][0] - res["exact"][0])
    d_un = abs(res["untrained"][0] - res["exact"][0])
    print(f"log-likelihood over {T} intervals on the exact filter's genealogy: exact "
          f"{res['exact'][0]:.3f}, trained surrogate {res['trained'][0]:.3f} (diff {d_tr:.4f}), "
          f"untrained {res['untrained'][0]:.3f} (diff {d_un:.1f})", flush=True)
    assert d_un > 100.0 * d_tr
    assert d_tr < 5.0
    untrained.close()
    net.close()
