"""Batched inner filters (bihip_batch_*): every inner filter equals the
oracle's bootstrap filter run with its own theta_j, N_j and Philox batch id j."""
import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import oracle_model

pytestmark = pytest.mark.gpu

# ragged particle counts, including the edges of the workgroup rows
NS = [1, 2, 3, 255, 256, 257, 511, 1000, 1024, 1025, 4096, 16384, 700, 64, 5000, 129]


def _pz_data(orc, T, seed=100):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, seed, 1)[:, 0]
    _, y = orc.simulate(m, seed, T, 1.0, x0)
    return y


def _thetas(B, seed=3):
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.normal(np.log(0.5), 0.2, B))
    sigma = np.exp(rng.normal(np.log(0.2), 0.2, B))
    c = np.exp(rng.normal(np.log(0.25), 0.1, B))
    e = np.exp(rng.normal(np.log(0.3), 0.1, B))
    return np.stack([mu, sigma, c, e], axis=1)


def _oracle_filter(orc, gm, theta, N, y, j, resampler, ess_rel, seed):
    om = oracle_model(orc, gm)
    for i, v in enumerate(theta):
        om.par[i] = v
    return orc.pf_run(om, y, int(N), 1.0, seed, resampler=resampler, ess_rel=ess_rel, batch=j,
                      want_ancestors=True)


@pytest.mark.parametrize("resampler", [capi.SYSTEMATIC, capi.STRATIFIED])
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_batch_matches_oracle(ctx, orc, resampler, block):
    T = 20
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _thetas(len(NS))
    bf = capi.BatchFilter(ctx, gm, NS, theta, T, 1.0, seed=42, resampler=resampler, ess_rel=0.7,
                          record=True, block=block)
    bf.set_observations(y)
    bf.init()
    bf.step(T)
    out = bf.outputs()
    assert (out["status"] == 0).all()
    for j, N in enumerate(NS):
        ref = _oracle_filter(orc, gm, theta[j], N, y, j, resampler, 0.7, 42)
        assert out["logLikelihood"][j] == pytest.approx(ref["logIncrements"].sum(), rel=1e-10,
                                                        abs=1e-10), j
        np.testing.assert_allclose(out["logIncrements"][j], ref["logIncrements"], rtol=1e-9,
                                   atol=1e-9)
        np.testing.assert_array_equal(out["resampled"][j], ref["resampled"])
        for k in range(T):
            np.testing.assert_array_equal(bf.ancestors(j, k), ref["ancestors"][k], f"{j} {k}")
    bf.close()


def test_batch_equals_single_filter(ctx, orc):
    """Inner filter j of a batch equals bihip_pf run alone with batch = j."""
    T = 15
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    Ns = [300, 2048, 7000]
    theta = _thetas(3, seed=9)
    bf = capi.BatchFilter(ctx, gm, Ns, theta, T, 1.0, seed=5, ess_rel=0.6)
    bf.set_observations(y)
    bf.init()
    bf.step(T)
    out = bf.outputs()
    for j, N in enumerate(Ns):
        m = capi.make_model("pz", precision=capi.F64, mu=theta[j, 0], sigma=theta[j, 1],
                            c=theta[j, 2], e=theta[j, 3])
        pf = capi.ParticleFilter(ctx, m, N, T, 1.0, 5, ess_rel=0.6, batch=j)
        pf.set_observations(y)
        pf.init()
        pf.step(T)
        z = pf.term()
        o = pf.outputs()
        pf.close()
        np.testing.assert_allclose(out["logIncrements"][j], o["logIncrements"], rtol=1e-10,
                                   atol=1e-10)
        assert out["logLikelihood"][j] == pytest.approx(z, rel=1e-10)


def test_batch_autotune_and_graph(ctx, orc):
    """The autotuned size is one of the candidates; graph replay at that size
    reproduces the direct launches bit for bit (the workgroup size sets the
    order of the weight sums, so both runs use the same size)."""
    T = 12
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    rng = np.random.default_rng(1)
    Ns = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), 64))).astype(int)
    theta = _thetas(64, seed=2)
    res = []
    block, items, oversub = 0, 0, 0
    for use_graph in (False, True):
        bf = capi.BatchFilter(ctx, gm, Ns, theta, T, 1.0, seed=8, ess_rel=0.5,
                              use_graph=use_graph, block=block, items=items, oversub=oversub)
        bf.set_observations(y)
        bf.init()
        bf.step(T)
        blk, grid, ms = bf.tuning()
        if block == 0:
            assert blk in (256, 512, 1024) and grid > 0 and (ms > 0).all()
            it, ov, table = bf.tuning_table()
            # workgroup size x tile width (2, 4, 8 for PZ FP64 RK4) x grid factor (1, 2)
            assert len(table) == 18
            assert {c[1] for c, _ in table} == {2, 4, 8} and {c[2] for c, _ in table} == {1, 2}
            best = min(table, key=lambda r: r[1])
            assert best[0] == (blk, it, ov)
            print("autotuner:", sorted(table, key=lambda r: r[1])[:4], flush=True)
            block, items, oversub = blk, it, ov
        res.append(bf.outputs())
        bf.close()
    np.testing.assert_array_equal(res[0]["logIncrements"], res[1]["logIncrements"])


@pytest.mark.parametrize("items,oversub", [(2, 1), (8, 2), (4, 2)])
def test_batch_tile_widths_give_the_same_bits(ctx, orc, items, oversub):
    """The resampling passes' tile width and the grid's oversubscription change
    no result: the weight sums are exact integers and each filter's floating
    point sums run in the same order within its workgroup."""
    T = 10
    y = _pz_data(orc, T)
    gm = capi.make_model("pz", precision=capi.F64)
    theta = _thetas(len(NS))
    outs = []
    for it, ov in ((4, 1), (items, oversub)):
        bf = capi.BatchFilter(ctx, gm, NS, theta, T, 1.0, seed=42, ess_rel=0.7, record=True,
                              block=512, items=it, oversub=ov)
        bf.set_observations(y)
        bf.init()
        bf.step(T)
        outs.append((bf.outputs(), [bf.ancestors(j, T - 1) for j in range(len(NS))]))
        bf.close()
    (o0, a0), (o1, a1) = outs
    np.testing.assert_array_equal(o0["logIncrements"], o1["logIncrements"])
    for j in range(len(NS)):
        np.testing.assert_array_equal(a0[j], a1[j])


def test_batch_rejects_uncompiled_tile_width(ctx):
    gm = capi.make_model("lg", precision=capi.F64)
    with pytest.raises(capi.BihipError):
        capi.BatchFilter(ctx, gm, [100, 200], np.zeros((2, 0)), 4, 1.0, seed=1, block=256,
                         items=8)


def test_batch_rejects_oversized_filter(ctx):
    gm = capi.make_model("pz", precision=capi.F64)
    with pytest.raises(capi.BihipError):
        capi.BatchFilter(ctx, gm, [16385], np.zeros((1, 4)), 5, 1.0, seed=1)
