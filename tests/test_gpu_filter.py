"""Whole-filter parity of the device-resident bootstrap particle filter
(bihip_pf_*) with the oracle's filter on identical Philox inputs."""
import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import l96_spun_up, oracle_model

pytestmark = pytest.mark.gpu


def _pz_data(orc, T, seed=100):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return y


def _run_gpu(ctx, gm, y, P, dt, seed, resampler, ess_rel, mean0=None, record=True,
             use_graph=False):
    T = len(y)
    pf = capi.ParticleFilter(ctx, gm, P, T, dt, seed, resampler=resampler, ess_rel=ess_rel,
                             record=record, use_graph=use_graph)
    pf.set_observations(y)
    pf.init(mean0)
    pf.step(T)
    logZ = pf.term()
    out = pf.outputs()
    anc = np.array([pf.ancestors(k) for k in range(T)]) if record else None
    pf.close()
    return logZ, out, anc


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.STRATIFIED, capi.MULTINOMIAL])
def test_pz_filter_fp64_parity(ctx, orc, resampler):
    """Predator-prey bootstrap PF, N=4096, FP64, RK4 h=0.1, resample every step, T=100."""
    y = _pz_data(orc, 100)
    gm = capi.make_model("pz", precision=capi.F64)
    om = oracle_model(orc, gm)
    P = 4096
    logZ, out, anc = _run_gpu(ctx, gm, y, P, 1.0, 42, resampler, 1.0)
    ref = orc.pf_run(om, y, P, 1.0, 42, resampler=resampler, ess_rel=1.0, want_ancestors=True)
    # FP64 everywhere: the estimate agrees to rounding (relative 1e-10)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    # ancestors are bit-exact at every step
    np.testing.assert_array_equal(anc, ref["ancestors"])


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.STRATIFIED, capi.MULTINOMIAL])
def test_pz_filter_large_population_parity(ctx, orc, resampler):
    """N = 2^21 + 5: more than 1024 tiles of 2048 particles, so the resampling
    chain scans its tile sums in a launch of their own; ancestors stay bit-exact."""
    y = _pz_data(orc, 3)
    gm = capi.make_model("pz", precision=capi.F64)
    om = oracle_model(orc, gm)
    P = (1 << 21) + 5
    logZ, out, anc = _run_gpu(ctx, gm, y, P, 1.0, 43, resampler, 1.0)
    ref = orc.pf_run(om, y, P, 1.0, 43, resampler=resampler, ess_rel=1.0, want_ancestors=True)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_l96_filter_fp64_parity(ctx, orc):
    """Lorenz-96 d=40, DOPRI5 rtol=atol=1e-5, FP64 state, N=4096, T=20, ESS<0.5N."""
    om0 = orc.make_model("l96", d=40)
    mean0 = l96_spun_up(orc)
    _, y = orc.simulate(om0, 7, 20, 0.05, mean0)
    gm = capi.make_model("l96", d=40, precision=capi.F64)
    om = oracle_model(orc, gm)
    P = 4096
    logZ, out, anc = _run_gpu(ctx, gm, y, P, 0.05, 9, capi.SYSTEMATIC, 0.5, mean0=mean0)
    ref = orc.pf_run(om, y, P, 0.05, 9, ess_rel=0.5, mean0=mean0, want_ancestors=True)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-9)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_l96_d200_filter_fp64_parity(ctx, orc):
    """Lorenz-96 d=200 (eight lanes per particle, 25 dimensions each: the
    observed even dimensions alternate between lanes), FP64, N=2048, T=8."""
    om0 = orc.make_model("l96", d=200)
    mean0 = l96_spun_up(orc, d=200)
    _, y = orc.simulate(om0, 7, 8, 0.05, mean0)
    gm = capi.make_model("l96", d=200, precision=capi.F64)
    om = oracle_model(orc, gm)
    P = 2048
    logZ, out, anc = _run_gpu(ctx, gm, y, P, 0.05, 9, capi.SYSTEMATIC, 0.5, mean0=mean0)
    ref = orc.pf_run(om, y, P, 0.05, 9, ess_rel=0.5, mean0=mean0, want_ancestors=True)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-9)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_l96_d200_filter_fp32_mixed_precision(ctx, orc):
    """d=200 with FP32 state (eight lanes per particle, the forcing carried in the
    stage constants) and no resampling: the FP64 oracle's logZ and ESS to FP32
    rounding."""
    om0 = orc.make_model("l96", d=200)
    mean0 = l96_spun_up(orc, d=200)
    _, y = orc.simulate(om0, 7, 5, 0.05, mean0)
    gm = capi.make_model("l96", d=200, precision=capi.F32)
    om = oracle_model(orc, gm)
    P = 2048
    logZ, out, _ = _run_gpu(ctx, gm, y, P, 0.05, 9, capi.SYSTEMATIC, 0.0, mean0=mean0,
                            record=False)
    ref = orc.pf_run(om, y, P, 0.05, 9, ess_rel=0.0, mean0=mean0)
    assert not out["resampled"].any()
    assert logZ == pytest.approx(ref["logZ"], rel=1e-5)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-4, atol=1e-2)
    np.testing.assert_allclose(out["ess"], ref["ess"], rtol=1e-3)


@pytest.mark.parametrize("T", [1, 10])
def test_l96_filter_fp32_mixed_precision(ctx, orc, T):
    """FP32 transitions with FP64 log-weights and logZ against the FP64 oracle on
    identical 24-bit Philox noise.  Without resampling the two filters carry
    the same particles, so the difference is FP32 rounding alone.  (With
    resampling, a single ancestor decided differently by rounding sends the
    filters down different Monte Carlo paths; at N=4096 this model's logZ
    spreads over hundreds of nats between seeds, so that comparison is not a
    precision test.)"""
    om0 = orc.make_model("l96", d=40)
    mean0 = l96_spun_up(orc)
    _, y = orc.simulate(om0, 7, T, 0.05, mean0)
    gm = capi.make_model("l96", d=40, precision=capi.F32)
    om = oracle_model(orc, gm)
    P = 4096
    logZ, out, _ = _run_gpu(ctx, gm, y, P, 0.05, 9, capi.SYSTEMATIC, 0.0, mean0=mean0,
                            record=False)
    ref = orc.pf_run(om, y, P, 0.05, 9, ess_rel=0.0, mean0=mean0)
    assert not out["resampled"].any()
    # FP32 state: relative error of logZ below 1e-6 after one step, 1e-5 after ten
    tol = 1e-6 if T == 1 else 1e-5
    assert logZ == pytest.approx(ref["logZ"], rel=tol)
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=tol * 10,
                               atol=tol * 100)


def test_lg_filter_matches_kalman(ctx, orc):
    gm = capi.make_model("lg", precision=capi.F64, lam=0.5, noise_sd=0.5, obs_sd=0.5,
                         atol=1e-10, rtol=1e-10, h0=0.25)
    om = oracle_model(orc, gm)
    x0 = orc.init_samples(om, 11, 1)[:, 0]
    _, y = orc.simulate(om, 11, 40, 1.0, x0)
    exact = orc.kalman_loglik_lg(om, 1.0, y)
    zs = [_run_gpu(ctx, gm, y, 1 << 16, 1.0, s, capi.SYSTEMATIC, 0.5, record=False)[0]
          for s in (1, 2)]
    assert abs(np.mean(zs) - exact) < 0.05, (zs, exact)


def test_graph_replay_equals_eager(ctx, orc):
    y = _pz_data(orc, 30)
    gm = capi.make_model("pz", precision=capi.F64)
    a = _run_gpu(ctx, gm, y, 2048, 1.0, 5, capi.SYSTEMATIC, 0.7, use_graph=False)
    b = _run_gpu(ctx, gm, y, 2048, 1.0, 5, capi.SYSTEMATIC, 0.7, use_graph=True)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1]["logIncrements"], b[1]["logIncrements"])
    np.testing.assert_array_equal(a[2], b[2])


def test_degenerate_weights_do_not_fail(ctx, orc):
    gm = capi.make_model("lg", precision=capi.F64)
    om = oracle_model(orc, gm)
    x0 = orc.init_samples(om, 3, 1)[:, 0]
    _, y = orc.simulate(om, 3, 5, 1.0, x0)
    y = y.copy()
    y[1] = np.inf
    logZ, out, _ = _run_gpu(ctx, gm, y, 256, 1.0, 3, capi.SYSTEMATIC, 1.0, record=False)
    assert logZ == -np.inf
    assert np.isnan(out["ess"][1]) and not out["resampled"][2:].any()


def test_step_beyond_schedule_is_rejected(ctx, orc):
    gm = capi.make_model("pz", precision=capi.F64)
    pf = capi.ParticleFilter(ctx, gm, 64, 3, 1.0, 1)
    pf.set_observations(np.ones(3))
    with pytest.raises(capi.BihipError):
        pf.step(1)  # not initialised
    pf.init()
    pf.step(3)
    with pytest.raises(capi.BihipError):
        pf.step(1)
    pf.close()


def test_paths_match_oracle_lineages(ctx, orc):
    """record = 2: particle paths traced through the device ancestry equal the
    oracle's lineages (AncestryCache::readPath), and samplePath draws the
    particle the keyed uniform selects from the final weights."""
    T, P, seed = 10, 512, 42
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    om = oracle_model(orc, gm)
    pf = capi.ParticleFilter(ctx, gm, P, T, 1.0, seed, ess_rel=0.7, record="paths")
    pf.set_observations(y)
    pf.init()
    pf.step(T)
    idx = np.array([0, 7, 100, 511, 256], np.int32)
    got = pf.paths(idx)
    ref = orc.pf_run(om, y, P, 1.0, seed, ess_rel=0.7, want_ancestors=True)
    states = [orc.init_samples(om, seed, P)]
    for k in range(1, T + 1):
        states.append(orc.pf_run(om, y[:k], P, 1.0, seed, ess_rel=0.7, want_state=True)["X"])
    for q, p in enumerate(idx):
        for k in range(T, -1, -1):
            np.testing.assert_allclose(got[q, k], states[k][:, p], rtol=1e-12, atol=1e-12)
            if k > 0:
                p = ref["ancestors"][k - 1][p]
    pick, path = pf.sample_path()
    _, _, lws = pf.state()
    lw = lws.cpu().numpy()
    Ws = np.cumsum(np.exp(lw - lw.max()))
    u = orc.uniform_at(seed, 0, T, capi.TAG_PATH, 0)
    assert pick == min(int(np.searchsorted(Ws, Ws[-1] * u, side="right")), P - 1)
    np.testing.assert_array_equal(path, pf.paths([pick])[0])
    pf.close()


def test_netcdf_output_of_the_recorded_filter(ctx, orc, tmp_path):
    """bihip_pf_write_netcdf: the recorded history in LibBi's ParticleFilter
    schema, read back with scipy's NetCDF reader: times t0 + k dt, the oracle's
    states and ancestors at every output (output 0: the initial samples with
    identity ancestors), log-weights whose log-mean-exp after step k is the
    running log-likelihood (the sum of the log-increments), and logZ."""
    from scipy.io import netcdf_file
    T, P, seed = 8, 256, 11
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    om = oracle_model(orc, gm)
    pf = capi.ParticleFilter(ctx, gm, P, T, 1.0, seed, ess_rel=0.7, record="paths")
    pf.set_observations(y)
    pf.init()
    pf.step(T)
    path = tmp_path / "pf.nc"
    pf.write_netcdf(path)
    out = pf.outputs()
    logZ = pf.term()
    _, _, lws = pf.state()
    pf.close()
    f = netcdf_file(str(path), "r", mmap=False)
    assert f.dimensions == {"nr": T + 1, "np": P}
    assert f._attributes["libbi_schema"] == b"ParticleFilter"
    v = {k: np.array(x.data) for k, x in f.variables.items()}
    f.close()
    np.testing.assert_array_equal(v["time"], np.arange(T + 1, dtype=float))
    ref = orc.pf_run(om, y, P, 1.0, seed, ess_rel=0.7, want_ancestors=True)
    np.testing.assert_array_equal(v["ancestor"][0], np.arange(P))
    np.testing.assert_array_equal(v["ancestor"][1:], ref["ancestors"])
    states = [orc.init_samples(om, seed, P)]
    for k in range(1, T + 1):
        states.append(orc.pf_run(om, y[:k], P, 1.0, seed, ess_rel=0.7, want_state=True)["X"])
    for k in range(T + 1):
        np.testing.assert_allclose(v["P"][k], states[k][0], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(v["Z"][k], states[k][1], rtol=1e-12, atol=1e-12)
    lw = v["logweight"]
    np.testing.assert_array_equal(lw[0], np.zeros(P))
    np.testing.assert_array_equal(lw[T], lws.cpu().numpy())
    running = np.cumsum(out["logIncrements"])
    for k in range(1, T + 1):
        m = lw[k].max()
        lme = m + np.log(np.exp(lw[k] - m).sum()) - np.log(P)
        assert lme == pytest.approx(running[k - 1], rel=1e-12, abs=1e-9)
    assert float(v["loglikelihood"]) == logZ
    # the noise variable alpha (R_VAR): zero before the first step, then the
    # draw alpha = mu + sigma z of each step, z the particle's first noise normal
    np.testing.assert_array_equal(v["alpha"][0], np.zeros(P))
    mu, sigma = gm.par[0], gm.par[1]
    for k in (1, 4, T):
        z = np.array([orc.normals(seed, p, k, orc.TAG_NOISE, 1, bits=32)[0] for p in range(P)])
        np.testing.assert_allclose(v["alpha"][k], mu + sigma * z, rtol=1e-14, atol=1e-14)


def test_recorded_noise_of_lorenz96_and_unchanged_results(ctx, orc, tmp_path):
    """Lorenz-96 d=8 FP64 with record="paths": the NetCDF file's w[n] holds each
    step's standard normals (x <- x + noise_sd w), and recording changes no
    result of the filter (the same increments and states, bit for bit)."""
    from scipy.io import netcdf_file
    T, P, seed, dt = 6, 300, 5, 0.05
    gm = capi.make_model("l96", d=8, precision=capi.F64)
    om = oracle_model(orc, gm)
    mean0 = l96_spun_up(orc, d=8)
    x0 = orc.init_samples(om, 3, 1, mean0=mean0)[:, 0]
    _, y = orc.simulate(om, 3, T, dt, x0)
    runs = {}
    for rec in (False, "paths"):
        pf = capi.ParticleFilter(ctx, gm, P, T, dt, seed, ess_rel=0.5, record=rec)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(T)
        if rec:
            pf.write_netcdf(tmp_path / "l96.nc")
        X, _, lws = pf.state()
        runs[rec] = (pf.outputs()["logIncrements"], X[:, :P].cpu().numpy(), lws.cpu().numpy())
        pf.close()
    for a, b in zip(runs[False], runs["paths"]):
        np.testing.assert_array_equal(a, b)
    f = netcdf_file(str(tmp_path / "l96.nc"), "r", mmap=False)
    assert f.variables["w"].dimensions == ("nr", "n", "np")
    w = np.array(f.variables["w"].data)
    f.close()
    np.testing.assert_array_equal(w[0], np.zeros((8, P)))
    for k in (1, T):
        for p in (0, 1, P // 2, P - 1):
            z = orc.normals(seed, p, k, orc.TAG_NOISE, 8, bits=32)
            np.testing.assert_allclose(w[k][:, p], z, rtol=1e-14, atol=1e-14)


def test_netcdf_output_needs_recorded_paths(ctx, orc, tmp_path):
    y = _pz_data(orc, 3)
    pf = capi.ParticleFilter(ctx, capi.make_model("pz", precision=capi.F64), 64, 3, 1.0, 1)
    pf.set_observations(y)
    pf.init()
    pf.step(3)
    with pytest.raises(capi.BihipError):
        pf.write_netcdf(tmp_path / "x.nc")
    pf.close()


def _irregular(T, seed=3, p_unobserved=0.3):
    rng = np.random.default_rng(seed)
    times = np.concatenate([[0.0], np.cumsum(rng.uniform(0.3, 1.7, size=T))])
    observed = (rng.uniform(size=T) > p_unobserved).astype(np.int32)
    observed[0] = 1
    return times, observed


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.MULTINOMIAL])
@pytest.mark.parametrize("use_graph", [False, True])
def test_scheduled_filter_matches_oracle(ctx, orc, resampler, use_graph):
    """Irregular observation times with unobserved steps (bihip_pf_set_schedule):
    bit-exact ancestors and resampling decisions, log-increments (zero where
    unobserved) and logZ against the oracle's scheduled filter, FP64."""
    T, P, seed = 40, 4096, 21
    y = _pz_data(orc, T)
    times, observed = _irregular(T)
    gm = capi.make_model("pz", precision=capi.F64)
    om = oracle_model(orc, gm)
    pf = capi.ParticleFilter(ctx, gm, P, T, 1.0, seed, resampler=resampler, ess_rel=0.8,
                             record=True, use_graph=use_graph)
    pf.set_schedule(times, observed)
    pf.set_observations(y)
    pf.init()
    pf.step(T)
    logZ, out = pf.term(), pf.outputs()
    anc = np.array([pf.ancestors(k) for k in range(T)])
    pf.close()
    ref = orc.pf_run(om, y, P, 1.0, seed, resampler=resampler, ess_rel=0.8,
                     want_ancestors=True, times=times, observed=observed)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_allclose(out["logIncrements"], ref["logIncrements"], rtol=1e-9, atol=1e-9)
    assert np.all(out["logIncrements"][observed == 0] == 0.0)
    np.testing.assert_array_equal(anc, ref["ancestors"])
    assert out["resampled"].any()


def test_scheduled_lg_filter_matches_kalman(ctx, orc):
    """The scheduled filter's logZ on the linear-Gaussian model against the exact
    Kalman log-likelihood over the same schedule (N = 2^20, four seeds)."""
    T = 30
    times, observed = _irregular(T, seed=5)
    om = orc.make_model("lg", lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0, s0=1.0, atol=1e-10,
                        rtol=1e-10, h0=0.25)
    x0 = orc.init_samples(om, 11, 1)[:, 0]
    _, y = orc.simulate(om, 11, T, 1.0, x0)
    exact = orc.kalman_loglik_lg_sched(om, times, observed, y)
    gm = capi.make_model("lg", precision=capi.F64, lam=0.5, noise_sd=0.5, obs_sd=0.5, m0=0.0,
                         s0=1.0, atol=1e-10, rtol=1e-10, h0=0.25)
    zs = []
    for seed in range(4):
        pf = capi.ParticleFilter(ctx, gm, 1 << 20, T, 1.0, 100 + seed, ess_rel=0.5)
        pf.set_schedule(times, observed)
        pf.set_observations(y)
        pf.init()
        pf.step(T)
        zs.append(pf.term())
        pf.close()
    print("scheduled LG logZ", np.mean(zs), np.std(zs), "Kalman", exact, flush=True)
    assert abs(np.mean(zs) - exact) < 0.02


def test_schedule_rejects_bad_times(ctx, orc):
    pf = capi.ParticleFilter(ctx, capi.make_model("pz", precision=capi.F64), 64, 3, 1.0, 1)
    with pytest.raises(capi.BihipError):
        pf.set_schedule([0.0, 1.0, 1.0, 2.0])
    with pytest.raises(ValueError):
        pf.set_schedule([0.0, 1.0, 2.0])
    pf.close()


def test_netcdf_times_follow_the_schedule(ctx, orc, tmp_path):
    from scipy.io import netcdf_file
    T = 6
    y = _pz_data(orc, T)
    times, observed = _irregular(T, seed=9)
    pf = capi.ParticleFilter(ctx, capi.make_model("pz", precision=capi.F64), 128, T, 1.0, 3,
                             record="paths")
    pf.set_schedule(times, observed)
    pf.set_observations(y)
    pf.init()
    pf.step(T)
    pf.write_netcdf(tmp_path / "s.nc")
    pf.close()
    f = netcdf_file(str(tmp_path / "s.nc"), "r", mmap=False)
    np.testing.assert_array_equal(np.array(f.variables["time"].data), times)
    f.close()


def test_inputs_match_oracle(ctx, orc):
    """Per-step parameters (bihip_pf_set_inputs, LibBi's inputs): a Lorenz-96
    forcing F_k that varies over the steps, FP64 d=40 with a schedule too:
    bit-exact ancestors and logZ against the oracle with the same inputs."""
    om0 = orc.make_model("l96", d=40)
    mean0 = l96_spun_up(orc)
    T, P, seed = 15, 2048, 9
    _, y = orc.simulate(om0, 7, T, 0.05, mean0)
    gm = capi.make_model("l96", d=40, precision=capi.F64)
    om = oracle_model(orc, gm)
    pf = capi.ParticleFilter(ctx, gm, P, T, 0.05, seed, ess_rel=0.5, record=True)
    F = 8.0 + 1.5 * np.sin(np.arange(T))
    inputs = pf.inputs_from(F=F, noise_sd=0.1 + 0.02 * np.arange(T) / T)
    times = np.concatenate([[0.0], np.cumsum(np.random.default_rng(2).uniform(0.03, 0.08, T))])
    pf.set_schedule(times)
    pf.set_inputs(inputs)
    pf.set_observations(y)
    pf.init(mean0)
    pf.step(T)
    logZ, out = pf.term(), pf.outputs()
    anc = np.array([pf.ancestors(k) for k in range(T)])
    pf.close()
    ref = orc.pf_run(om, y, P, 0.05, seed, ess_rel=0.5, mean0=mean0, want_ancestors=True,
                     times=times, inputs=inputs)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-9)
    np.testing.assert_array_equal(out["resampled"], ref["resampled"])
    np.testing.assert_array_equal(anc, ref["ancestors"])


def test_inputs_equal_to_the_parameters_change_nothing(ctx, orc):
    y = _pz_data(orc, 20)
    gm = capi.make_model("pz", precision=capi.F64)
    res = []
    for use in (False, True):
        pf = capi.ParticleFilter(ctx, gm, 1024, 20, 1.0, 4, ess_rel=0.7, record=True)
        if use:
            pf.set_inputs(pf.inputs_from())
        pf.set_observations(y)
        pf.init()
        pf.step(20)
        res.append((pf.term(), np.array([pf.ancestors(k) for k in range(20)])))
        pf.close()
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])
