"""Filter banks (bihip_pfbank_*): B filters of P particles advanced by shared
launches.  Filter j of a bank equals bihip_pf run alone with parameters
theta_j and Philox batch id j -- the same ancestors at every step, bit for
bit, the same states, and log-likelihood increments equal to rounding (the
bank's correct step sums the same partials in another order) -- and equals
the oracle's filter with theta_j and batch id j."""
import numpy as np
import pytest

from libbi_amd import capi, data, guided
from gpu_helpers import oracle_model

pytestmark = pytest.mark.gpu

DT = 0.05


def _pz_data(orc, T, seed=100):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return y


def _pz_thetas(B, seed=3):
    rng = np.random.default_rng(seed)
    return np.stack([np.exp(rng.normal(np.log(0.5), 0.2, B)),
                     np.exp(rng.normal(np.log(0.2), 0.2, B)),
                     np.exp(rng.normal(np.log(0.25), 0.1, B)),
                     np.exp(rng.normal(np.log(0.3), 0.1, B))], axis=1)


def _lone(ctx, gm, theta, P, T, dt, seed, j, resampler, ess_rel, y, mean0=None):
    m = capi.Model.from_buffer_copy(gm)
    for i, v in enumerate(theta):
        m.par[i] = v
    pf = capi.ParticleFilter(ctx, m, P, T, dt, seed, resampler=resampler, ess_rel=ess_rel,
                             batch=j, record=True)
    pf.set_observations(y)
    pf.init(mean0)
    return pf


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.STRATIFIED])
def test_pz_bank_equals_lone_filters_and_oracle(ctx, orc, resampler):
    B, P, T = 6, 3000, 25  # P not a multiple of the tile or of 64
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _pz_thetas(B)
    bank = capi.FilterBank(ctx, gm, P, T, 1.0, 42, theta, resampler=resampler, ess_rel=0.7,
                           record=True)
    bank.set_observations(y)
    bank.init()
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    assert out["resampled"].sum() > 20
    for j in range(B):
        pf = _lone(ctx, gm, theta[j], P, T, 1.0, 42, j, resampler, 0.7, y)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-13,
                                   atol=1e-12)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k), f"{j} {k}")
        Xb, ld, lwb = bank.state(j)
        Xp, _, lwp = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy())
        assert np.array_equal(lwb.cpu().numpy(), lwp.cpu().numpy())
        pf.close()
        if j in (0, B - 1):
            om = oracle_model(orc, gm)
            for i, v in enumerate(theta[j]):
                om.par[i] = v
            ref = orc.pf_run(om, y, P, 1.0, 42, resampler=resampler, ess_rel=0.7, batch=j,
                             want_ancestors=True)
            np.testing.assert_allclose(out["logIncrements"][j], ref["logIncrements"], rtol=1e-9,
                                       atol=1e-9)
            np.testing.assert_array_equal(out["resampled"][j], ref["resampled"])
            for k in range(T):
                np.testing.assert_array_equal(bank.ancestors(j, k), ref["ancestors"][k])
    bank.close()


def test_bank_of_one_filter_runs_its_row(ctx, orc):
    """A bank of one filter (SMC^2 or PMMH with one theta-particle or chain)
    steps with its own parameter row, not the model's: it equals the lone
    filter with that theta."""
    P, T = 2000, 10
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _pz_thetas(1, seed=9)
    bank = capi.FilterBank(ctx, gm, P, T, 1.0, 42, theta, ess_rel=0.7, record=True)
    bank.set_observations(y)
    bank.init()
    bank.step(T)
    out = bank.outputs()
    pf = _lone(ctx, gm, theta[0], P, T, 1.0, 42, 0, capi.SYSTEMATIC, 0.7, y)
    pf.step(T)
    po = pf.outputs()
    np.testing.assert_array_equal(out["resampled"][0], po["resampled"])
    np.testing.assert_allclose(out["logIncrements"][0], po["logIncrements"], rtol=1e-13,
                               atol=1e-12)
    for k in range(T):
        np.testing.assert_array_equal(bank.ancestors(0, k), pf.ancestors(k))
    pf.close()
    bank.close()


def test_l96_d200_fp32_bank_equals_lone_filters(ctx, torch_gpu):
    """Eight-lane Lorenz-96 d=200 in FP32, one forcing per filter (the
    configuration of 64 forcings x 2^18 particles, here 4 x 2^14)."""
    B, P, T = 4, 1 << 14, 12
    Fs = np.array([7.7, 7.9, 8.1, 8.3])
    gm = guided.model(8.0)
    mean0 = data.l96_spinup(ctx, d=gm.d, F=8.0)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=31)
    bank = capi.FilterBank(ctx, gm, P, T, DT, 5, Fs, ess_rel=0.5, record=True, batch=100)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    assert out["resampled"].sum() >= 4
    for j in range(B):
        pf = _lone(ctx, gm, [Fs[j]], P, T, DT, 5, 100 + j, capi.SYSTEMATIC, 0.5, y, mean0)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-13)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k), f"{j} {k}")
        Xb, _, _ = bank.state(j)
        Xp, _, _ = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy()), j
        pf.close()
    # the forcing matters: the filters' likelihoods differ
    assert len(set(np.round(out["logLikelihood"], 6))) == B
    bank.close()


def test_bank_graph_replay_equals_eager(ctx, orc):
    B, P, T = 3, 5000, 15
    y = _pz_data(orc, T, seed=7)
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _pz_thetas(B, seed=8)
    outs = []
    for use_graph in (False, True):
        bank = capi.FilterBank(ctx, gm, P, T, 1.0, 9, theta, ess_rel=0.8, record=True,
                               use_graph=use_graph)
        bank.set_observations(y)
        bank.init()
        for _ in range(T):  # one replay at a time, with synchronisations between them
            bank.step(1)
            ctx.sync()
        o = bank.outputs()
        o["anc"] = np.array([[bank.ancestors(j, k) for k in range(T)] for j in range(B)])
        outs.append(o)
        bank.close()
    for key in ("logIncrements", "resampled", "anc", "logLikelihood"):
        np.testing.assert_array_equal(outs[0][key], outs[1][key], key)


def test_generated_bank_graph_replay_equals_eager(ctx):
    """A bank of filters on Lorenz96d40.bi (four lanes per particle, its
    interval loop over delta) replayed as a captured hipGraph, one step at a
    time with synchronisations between replays, equals the eager bank."""
    B, P, T = 3, 2048, 8
    par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
    g = capi.generated_model("Lorenz96d40", precision=capi.F32, **par)
    h = capi.make_model("l96", d=40, precision=capi.F32, atol=g.atol, rtol=g.rtol, h0=g.h0,
                        **par)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, h, mean0, T, DT, seed=19)
    theta = np.array([[7.9], [8.0], [8.1]])
    outs = []
    for use_graph in (False, True):
        bank = capi.FilterBank(ctx, g, P, T, DT, 13, theta, ess_rel=0.6, record=True,
                               use_graph=use_graph)
        bank.set_observations(y)
        bank.init()
        for _ in range(T):
            bank.step(1)
            ctx.sync()
        o = bank.outputs()
        o["anc"] = np.array([[bank.ancestors(j, k) for k in range(T)] for j in range(B)])
        outs.append(o)
        bank.close()
    for key in ("logIncrements", "resampled", "anc", "logLikelihood"):
        np.testing.assert_array_equal(outs[0][key], outs[1][key], key)


def test_pmmh_over_generated_banks_with_the_model_files_prior(ctx, orc):
    """PMMH over lam of LinearGaussian.bi with its parameter block's prior and
    the chains' filters as one bank equals the batched form with equal sizes:
    the same draws, acceptances and log-likelihoods."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, **par)
    h = capi.make_model("lg", precision=capi.F64, **par)
    om = oracle_model(orc, h)
    T, B, C, P = 12, 32, 10, 512
    x0 = orc.init_samples(om, 5, 1)[:, 0]
    _, y = orc.simulate(om, 5, T, 1.0, x0)
    outs = []
    for bank in (None, (B, P)):
        s = capi.PMMH(ctx, g, [P] * B, T, C, None, None, [0.3], seed=3, bank=bank)
        s.set_observations(np.asarray(y).reshape(T, 1))
        s.run()
        outs.append(s.outputs())
        s.close()
    a, b = outs
    assert a["accepted"].sum() > B
    np.testing.assert_array_equal(b["accepted"], a["accepted"])
    np.testing.assert_allclose(b["theta"], a["theta"], rtol=1e-15)
    np.testing.assert_allclose(b["ll"], a["ll"], rtol=1e-10, atol=1e-10)


def _guided_net(ctx):
    net = capi.MLP(ctx, guided.DIMS, seed=23)
    p = net.get_params()
    last = guided.DIMS[-1] * guided.DIMS[-2] + guided.DIMS[-1]
    p[-last:] *= 0.3
    net.set_params(p)
    return net


def test_guided_bank_equals_lone_guided_filters(ctx, torch_gpu):
    B, P, T = 3, 1 << 12, 8
    Fs = np.array([7.6, 8.0, 8.4])
    gm = guided.model(8.0)
    mean0 = data.l96_spinup(ctx, d=gm.d, F=8.0)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=13)
    net = _guided_net(ctx)
    norm = capi.GuidedNorm(x_shift=2.3, x_scale=3.6, innov_scale=1.0, par_shift=8.0,
                           par_scale=1.0)
    bank = capi.FilterBank(ctx, gm, P, T, DT, 3, Fs, ess_rel=0.5, record=True)
    bank.set_guided(net, norm)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    for j in range(B):
        pf = _lone(ctx, gm, [Fs[j]], P, T, DT, 3, j, capi.SYSTEMATIC, 0.5, y, mean0)
        pf.set_guided(net, norm)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-13)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k))
        Xb, _, _ = bank.state(j)
        Xp, _, _ = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy()), j
        pf.close()
    bank.close()
    net.close()


def test_surrogate_bank_equals_lone_surrogate_filters(ctx, torch_gpu):
    B, P, T = 3, 1 << 12, 8
    Fs = np.array([7.5, 8.0, 8.5])
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=17)
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=5)
    norm = capi.SurrogateNorm(x_shift=2.3, x_scale=3.6, dx_scale=0.05, par_shift=8.0,
                              par_scale=1.0)
    bank = capi.FilterBank(ctx, gm, P, T, DT, 4, Fs, ess_rel=0.5, record=True)
    bank.set_surrogate(net, norm)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    for j in range(B):
        pf = _lone(ctx, gm, [Fs[j]], P, T, DT, 4, j, capi.SYSTEMATIC, 0.5, y, mean0)
        pf.set_surrogate(net, norm)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-13)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k))
        pf.close()
    bank.close()
    net.close()


def test_bank_rejects_what_it_does_not_run(ctx):
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _pz_thetas(2)
    with pytest.raises(capi.BihipError, match="systematically or stratified"):
        capi.FilterBank(ctx, gm, 100, 5, 1.0, 1, theta, resampler=capi.MULTINOMIAL)
    gen = capi.generated_model("Lorenz96d40", precision=capi.F32)
    gbank = capi.FilterBank(ctx, gen, 256, 5, DT, 1, np.array([[8.0], [8.1]]))
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=1)
    norm = capi.SurrogateNorm(x_shift=0.0, x_scale=1.0, dx_scale=0.05, par_shift=8.0,
                              par_scale=1.0)
    with pytest.raises(capi.BihipError, match="serves FP32 Lorenz-96 d=40"):
        gbank.set_surrogate(net, norm)  # the learned flow is the built-in model's
    gbank.close()
    net.close()
    l96 = capi.make_model("l96", d=40, precision=capi.F32)
    bank = capi.FilterBank(ctx, l96, 256, 5, DT, 1, np.array([[8.0, 0.1], [8.0, 0.2]]))
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=1)
    norm = capi.SurrogateNorm(x_shift=0.0, x_scale=1.0, dx_scale=0.05, par_shift=8.0,
                              par_scale=1.0)
    with pytest.raises(capi.BihipError, match="differ only in par"):
        bank.set_surrogate(net, norm)
    bank.close()
    net.close()


def test_bank_whole_filter_operations_equal_batch(ctx, orc):
    """Outer resampling (filter j takes filter anc[j]) and accepted moves
    (filter j of one bank takes filter j of another where accepted) on a bank
    equal the same operations on bihip_batch filters of the same size, theta
    and streams, step for step afterwards."""
    B, P, T = 5, 1000, 12
    y = _pz_data(orc, T, seed=3)
    gm = capi.make_model("pz", precision=capi.F64)
    th1, th2 = _pz_thetas(B, seed=1), _pz_thetas(B, seed=2)
    anc = np.array([4, 4, 0, 2, 1], np.int32)
    acc = np.array([1, 0, 0, 1, 1], np.int32)
    runs = []
    for kind in ("bank", "batch"):
        def make(th, base):
            if kind == "bank":
                f = capi.FilterBank(ctx, gm, P, T, 1.0, 8, th, ess_rel=0.6)
                f.set_stream_base(base)
            else:
                f = capi.BatchFilter(ctx, gm, [P] * B, th, T, 1.0, seed=8, ess_rel=0.6, block=256,
                                     stream_base=base)
            f.set_observations(y)
            f.init()
            return f
        a, b = make(th1, 0), make(th2, 100)
        a.step(4)
        b.step(4)
        a.resample_filters(anc)
        a.step(2)
        b.step(2)
        a.accept(b, acc)
        a.step(T - 6)
        o = a.outputs()
        runs.append((o["logLikelihood"], np.array([a.increments(k) for k in range(T)])))
        a.close()
        b.close()
    np.testing.assert_allclose(runs[0][0], runs[1][0], rtol=1e-12)
    np.testing.assert_allclose(runs[0][1], runs[1][1], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("er_out,nmoves,use_graph", [(0.5, 1, False), (0.9, 2, True)])
def test_smc2_over_banks_matches_oracle(ctx, orc, er_out, nmoves, use_graph):
    """SMC^2 whose inner filters are filter banks against the oracle's
    MarginalSIR with equal inner sizes: same decisions, same evidence."""
    B, P, T = 16, 1500, 8
    y = _pz_data(orc, T, seed=13)
    gm = capi.make_model("pz", precision=capi.F64)
    pm, ps, pp = [np.log(0.5), np.log(0.2)], [0.3, 0.3], [0.1, 0.1]
    c = orc.smc2_config(B, T, pm, ps, pp, seed=31, nmoves=nmoves, resampler=capi.SYSTEMATIC,
                        ess_rel_inner=0.6, ess_rel_outer=er_out)
    ref = orc.smc2_run(oracle_model(orc, gm), c, [P] * B, y)
    s = capi.SMC2(ctx, gm, None, T, pm, ps, pp, seed=31, nmoves=nmoves, ess_rel_inner=0.6,
                  ess_rel_outer=er_out, use_graph=use_graph, bank=(B, P))
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    print("SMC2 over banks: resampled", ref["resampled"].sum(), "accepts",
          ref["accepts"].sum(), flush=True)
    assert ref["resampled"].sum() >= 1 and ref["accepts"].sum() >= 1
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(out["accepts"], ref["accepts"])
    np.testing.assert_allclose(out["theta"], ref["theta"], rtol=1e-15)
    np.testing.assert_allclose(out["lw"], ref["lw"], rtol=1e-9, atol=1e-9)
    assert out["logZ"] == pytest.approx(ref["logZ"], rel=1e-10, abs=1e-10)


@pytest.mark.parametrize("which", ["PZ", "Lorenz96d40"])
def test_generated_model_bank_equals_lone_filters(ctx, orc, which):
    """A bank on a model compiled from its LibBi file (PZ.bi in FP64;
    Lorenz96d40.bi in FP32 over four lanes per particle, which the batched
    filters do not run) equals the lone generated filters with the same
    parameters and batch ids: the same ancestors, states and log-weights."""
    if which == "PZ":
        B, P, T, dt, mean0 = 4, 3000, 10, 1.0, None
        g = capi.generated_model("PZ", precision=capi.F64, mu=0.5, sigma=0.2)
        theta = _pz_thetas(B)[:, :2]  # (mu, sigma), the file's first params
        y = _pz_data(orc, T)
    else:
        B, P, T, dt = 3, 2048, 8, DT
        par = dict(F=8.0, noise_sd=0.1, obs_sd=1.0, init_sd=1.0)
        g = capi.generated_model("Lorenz96d40", precision=capi.F32, **par)
        h = capi.make_model("l96", d=40, precision=capi.F32, atol=g.atol, rtol=g.rtol, h0=g.h0,
                            **par)
        mean0 = data.l96_spinup(ctx, d=40)
        _, y = data.simulate(ctx, h, mean0, T, DT, seed=5)
        theta = np.array([[7.8], [8.0], [8.2]])  # F
    bank = capi.FilterBank(ctx, g, P, T, dt, 42, theta, ess_rel=0.6, record=True)
    bank.set_observations(y)
    bank.init(mean0)
    bank.step(T)
    out = bank.outputs()
    assert (out["status"] == 0).all()
    assert out["resampled"].sum() >= 2
    for j in range(B):
        pf = _lone(ctx, g, theta[j], P, T, dt, 42, j, capi.SYSTEMATIC, 0.6, y, mean0)
        pf.step(T)
        po = pf.outputs()
        np.testing.assert_array_equal(out["resampled"][j], po["resampled"])
        np.testing.assert_allclose(out["logIncrements"][j], po["logIncrements"], rtol=1e-12,
                                   atol=1e-10)
        for k in range(T):
            np.testing.assert_array_equal(bank.ancestors(j, k), pf.ancestors(k), f"{j} {k}")
        Xb, _, lwb = bank.state(j)
        Xp, _, lwp = pf.state()
        assert np.array_equal(Xb[:, :P].cpu().numpy(), Xp[:, :P].cpu().numpy())
        assert np.array_equal(lwb.cpu().numpy(), lwp.cpu().numpy())
        pf.close()
    bank.close()


def test_smc2_over_generated_banks_equals_batched(ctx, orc):
    """SMC^2 over lam of LinearGaussian.bi with the prior of its parameter
    block, the inner filters as filter banks and as the batched filters of
    equal sizes: the same resamplings, moves and draws, the same evidence."""
    par = dict(lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0)
    g = capi.generated_model("LinearGaussian", precision=capi.F64, atol=1e-10, rtol=1e-10,
                             h0=0.25, **par)
    h = capi.make_model("lg", precision=capi.F64, atol=1e-10, rtol=1e-10, h0=0.25, **par)
    om = oracle_model(orc, h)
    B, P, T = 64, 512, 12
    x0 = orc.init_samples(om, 21, 1)[:, 0]
    _, y = orc.simulate(om, 21, T, 1.0, x0)
    y = np.asarray(y).reshape(T, 1)
    outs = []
    for bank in (None, (B, P)):
        s = capi.SMC2(ctx, g, [P] * B, T, None, None, [0.3], seed=3, nmoves=1,
                      ess_rel_outer=0.8, bank=bank)
        s.set_observations(y)
        s.run()
        outs.append(s.outputs())
        s.close()
    a, b = outs
    assert a["resampled"].sum() >= 1 and a["accepts"].sum() >= 1
    np.testing.assert_array_equal(b["resampled"], a["resampled"])
    np.testing.assert_array_equal(b["accepts"], a["accepts"])
    np.testing.assert_allclose(b["theta"], a["theta"], rtol=1e-15)
    np.testing.assert_allclose(b["lw"], a["lw"], rtol=1e-9, atol=1e-9)
    assert b["logZ"] == pytest.approx(a["logZ"], rel=1e-10, abs=1e-10)


def test_pmmh_over_banks_matches_oracle(ctx, orc):
    """PMMH whose chains' proposals are filtered as one filter bank against the
    oracle's MarginalMH with equal inner sizes: the same acceptances, draws and
    log-likelihoods."""
    B, P, T, C = 8, 1500, 8, 12
    y = _pz_data(orc, T, seed=13)
    gm = capi.make_model("pz", precision=capi.F64)
    pm, ps, qs = [np.log(0.5), np.log(0.2)], [0.3, 0.3], [0.1, 0.1]
    c = orc.smc2_config(B, T, pm, ps, qs, seed=17, ess_rel_inner=0.6)
    ref = orc.pmmh_run(oracle_model(orc, gm), c, C, [P] * B, y)
    s = capi.PMMH(ctx, gm, None, T, C, pm, ps, qs, seed=17, ess_rel=0.6, bank=(B, P))
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    print("PMMH over banks: accepted", ref["accepted"], flush=True)
    assert (ref["accepted"] > 1).any()
    np.testing.assert_array_equal(out["accepted"], ref["accepted"])
    np.testing.assert_allclose(out["theta"], ref["theta"], rtol=1e-15)
    np.testing.assert_allclose(out["ll"], ref["ll"], rtol=1e-10, atol=1e-10)


def test_pmmh_over_banks_runs_lorenz96_d40(ctx):
    """PMMH on the Lorenz-96 d=40 forcing (FP32 state, four lanes per particle,
    which the batched filters do not run): each chain's first log-likelihood
    equals the lone filter's with that forcing and batch id, and the chains
    move inside the prior's support."""
    B, P, T, C = 8, 2048, 6, 4
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=5)
    s = capi.PMMH(ctx, gm, None, T, C, [np.log(8.0)], [0.05], [0.02], dt=DT, seed=3,
                  bank=(B, P), mean0=mean0)
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    assert np.isfinite(out["ll"]).all()
    assert (out["theta"] > 6.0).all() and (out["theta"] < 10.5).all()
    for j in (0, B - 1):
        pf = _lone(ctx, gm, [out["theta"][0, j, 0]], P, T, DT, 3, j, capi.SYSTEMATIC, 0.5, y,
                   mean0)
        pf.step(T)
        assert pf.term() == pytest.approx(out["ll"][0, j], rel=1e-9)
        pf.close()


@pytest.mark.parametrize("learned", ["guided", "surrogate"])
def test_pmmh_over_learned_banks_equals_lone_learned_filters(ctx, torch_gpu, learned):
    """PMMH on the Lorenz-96 forcing with the learned proposal (d=200) or the
    learned flow (d=40) in the chains' filters: each chain's first
    log-likelihood equals the lone learned filter's with that forcing and batch
    id, and SMC^2 over the same learned banks runs inside the prior's support."""
    B, P, T, C = 6, 1 << 12, 5, 3
    if learned == "guided":
        gm = guided.model(8.0)
        mean0 = data.l96_spinup(ctx, d=gm.d, F=8.0)
        net = _guided_net(ctx)
        norm = capi.GuidedNorm(x_shift=2.3, x_scale=3.6, innov_scale=1.0, par_shift=8.0,
                               par_scale=1.0)
    else:
        gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                             init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
        mean0 = data.l96_spinup(ctx, d=40)
        net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=5)
        p = net.get_params()
        p *= 0.1  # a near-zero flow correction
        net.set_params(p)
        norm = capi.SurrogateNorm(x_shift=2.3, x_scale=3.6, dx_scale=0.05, par_shift=8.0,
                                  par_scale=1.0)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=9)
    s = capi.PMMH(ctx, gm, None, T, C, [np.log(8.0)], [0.05], [0.02], dt=DT, seed=3,
                  bank=(B, P), mean0=mean0)
    getattr(s, "set_" + learned)(net, norm)
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    assert np.isfinite(out["ll"]).all()
    for j in (0, B - 1):
        pf = _lone(ctx, gm, [out["theta"][0, j, 0]], P, T, DT, 3, j, capi.SYSTEMATIC, 0.5, y,
                   mean0)
        getattr(pf, "set_" + learned)(net, norm)
        pf.step(T)
        assert pf.term() == pytest.approx(out["ll"][0, j], rel=1e-9)
        pf.close()
    s = capi.SMC2(ctx, gm, None, T, [np.log(8.0)], [0.05], [0.02], dt=DT, seed=3,
                  ess_rel_outer=0.8, bank=(B, P), mean0=mean0)
    getattr(s, "set_" + learned)(net, norm)
    s.set_observations(y)
    s.run()
    o2 = s.outputs()
    s.close()
    net.close()
    assert np.isfinite(o2["logZ"])
    assert (o2["theta"] > 6.0).all() and (o2["theta"] < 10.5).all()


def test_smc2_over_surrogate_banks_runs_lorenz96(ctx, torch_gpu):
    """SMC^2 over the Lorenz-96 forcing with the surrogate flow in every inner
    filter: the outer loop runs (resamplings and accepted moves) and keeps the
    forcing inside the prior's support."""
    B, P, T = 16, 1024, 10
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=5)
    net = capi.MLP(ctx, [41, 256, 256, 256, 40], seed=5)
    p = net.get_params()
    p *= 0.1  # a near-zero flow correction: x' close to x
    net.set_params(p)
    norm = capi.SurrogateNorm(x_shift=2.3, x_scale=3.6, dx_scale=0.05, par_shift=8.0,
                              par_scale=1.0)
    s = capi.SMC2(ctx, gm, None, T, [np.log(8.0)], [0.05], [0.02], dt=DT, seed=3,
                  ess_rel_outer=0.8, use_graph=True, bank=(B, P), mean0=mean0)
    s.set_surrogate(net, norm)
    s.set_observations(y)
    s.run()
    out = s.outputs()
    s.close()
    net.close()
    print("surrogate SMC2: logZ", out["logZ"], "resampled", out["resampled"].sum(), "accepts",
          out["accepts"].sum(), flush=True)
    assert np.isfinite(out["logZ"])
    assert out["resampled"].sum() >= 1
    assert (out["theta"] > 6.0).all() and (out["theta"] < 10.5).all()
