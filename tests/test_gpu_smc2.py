"""SMC^2 (bihip_smc2_*) and the whole-filter operations under it
(bihip_batch_resample_filters, bihip_batch_accept) against the oracle's
MarginalSIR restatement: same keyed draws, same decisions, same evidence."""
import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import oracle_model

pytestmark = pytest.mark.gpu


def _data(orc, kind, T, seed=13):
    m = orc.make_model(kind)
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return np.asarray(y).reshape(T, -1)


def _compare(ctx, orc, gm, N, T, y, prior_mean, prior_sd, prop_sd, seed, nmoves, er_out, er_in,
             use_graph=False, block=256, resampler=capi.SYSTEMATIC):
    om = oracle_model(orc, gm)
    c = orc.smc2_config(len(N), T, prior_mean, prior_sd, prop_sd, seed=seed, nmoves=nmoves,
                        resampler=resampler, ess_rel_inner=er_in, ess_rel_outer=er_out)
    ref = orc.smc2_run(om, c, N, y)
    s = capi.SMC2(ctx, gm, N, T, prior_mean, prior_sd, prop_sd, seed=seed,
                  nmoves=nmoves, resampler=resampler, ess_rel_inner=er_in,
                  ess_rel_outer=er_out, use_graph=use_graph, block=block)
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    print("SMC2", len(N), T, nmoves, er_out, "resampled", ref["resampled"].sum(), "accepts",
          ref["accepts"].sum(), flush=True)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(out["accepts"], ref["accepts"])
    np.testing.assert_allclose(out["theta"], ref["theta"], rtol=1e-15)
    np.testing.assert_allclose(out["lw"], ref["lw"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(out["ess"], ref["ess"], rtol=1e-9)
    assert out["logZ"] == pytest.approx(ref["logZ"], rel=1e-10, abs=1e-10)
    return ref


@pytest.mark.parametrize("er_out,nmoves", [(0.5, 1), (0.9, 2), (1.0, 1)])
def test_smc2_pz_matches_oracle(ctx, orc, er_out, nmoves):
    T = 8
    y = _data(orc, "pz", T)
    gm = capi.make_model("pz", precision=capi.F64)
    N = [64, 300, 17, 1024, 5, 700, 256, 257, 1, 2000, 90, 128, 33, 511, 64, 4000]
    ref = _compare(ctx, orc, gm, N, T, y, [np.log(0.5), np.log(0.2)], [0.3, 0.3], [0.1, 0.1],
                   seed=31, nmoves=nmoves, er_out=er_out, er_in=0.6)
    assert ref["resampled"].sum() >= 1 and ref["accepts"].sum() >= 1


def test_smc2_lg_graph_and_stratified(ctx, orc):
    T = 10
    y = _data(orc, "lg", T)
    gm = capi.make_model("lg", precision=capi.F64)
    N = [200] * 6 + [50] * 6 + [1000] * 4
    ref = _compare(ctx, orc, gm, N, T, y, [np.log(0.5)], [1.0], [0.3], seed=4, nmoves=1,
                   er_out=0.8, er_in=0.5, use_graph=True, block=512,
                   resampler=capi.STRATIFIED)
    assert ref["resampled"].sum() >= 1


def test_batch_resample_filters_then_step(ctx, orc):
    """After resampling of whole filters, filter j holds the source filter's
    state and keeps drawing from its own random streams."""
    T = 6
    y = _data(orc, "pz", T)
    gm = capi.make_model("pz", precision=capi.F64)
    N = [100, 7, 2048, 300]
    theta = np.array([[0.5, 0.2], [0.6, 0.25], [0.45, 0.15], [0.55, 0.3]])
    bf = capi.BatchFilter(ctx, gm, N, theta, T, 1.0, seed=2, ess_rel=0.5,
                          capacity=4 * 2048)
    bf.set_observations(y)
    bf.init()
    bf.step(3)
    before = bf.outputs()["logLikelihood"]
    anc = np.array([2, 2, 0, 3], np.int32)
    bf.resample_filters(anc)
    after = bf.outputs()["logLikelihood"]
    np.testing.assert_array_equal(after, before[anc])
    np.testing.assert_array_equal(bf.N, np.array(N)[anc])
    bf.step(3)
    out = bf.outputs()
    assert (out["status"] == 0).all()
    # filters 0 and 1 hold the same particles at step 3 but draw different noise
    assert out["logLikelihood"][0] != out["logLikelihood"][1]
    bf.close()


@pytest.mark.parametrize("kind", ["pz", "lg"])
def test_pmmh_matches_oracle(ctx, orc, kind):
    T, C = 8, 12
    y = _data(orc, kind, T)
    gm = capi.make_model(kind, precision=capi.F64)
    om = oracle_model(orc, gm)
    if kind == "pz":
        N = [64, 300, 17, 1024, 5, 700, 256, 2000]
        pm, ps, qs = [np.log(0.5), np.log(0.2)], [0.3, 0.3], [0.1, 0.1]
    else:
        N = [200, 50, 1000, 128]
        pm, ps, qs = [np.log(0.5)], [1.0], [0.3]
    c = orc.smc2_config(len(N), T, pm, ps, qs, seed=17, ess_rel_inner=0.6)
    ref = orc.pmmh_run(om, c, C, N, y)
    s = capi.PMMH(ctx, gm, N, T, C, pm, ps, qs, seed=17, ess_rel=0.6, block=256)
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    print("PMMH", kind, "accepted", ref["accepted"], flush=True)
    np.testing.assert_array_equal(out["accepted"], ref["accepted"])
    np.testing.assert_allclose(out["theta"], ref["theta"], rtol=1e-15)
    np.testing.assert_allclose(out["ll"], ref["ll"], rtol=1e-10, atol=1e-10)
    assert (ref["accepted"] > 1).any()
