"""The pruned ancestry tree (record = 3, AncestryCache::writeState / prune /
readPath, share/src/bi/cache/AncestryCache.hpp:58-470) against the full
history (record = 2, whose paths are pinned against the oracle's states and
ancestors in test_gpu_filter.py) on the same seeded runs:

- paths read from the tree equal the full history's, bit for bit;
- the tree's live nodes equal a brute-force count from the recorded
  ancestors: every (generation, particle) that is an ancestor of a current
  particle, and nothing else;
- a history whose full record would not fit in HBM (Lorenz-96 d = 200,
  N = 2^20, T = 500 in FP32: 420 GB) runs in a tree of a few generations,
  and its paths end at the current particles."""
import numpy as np
import pytest

from libbi_amd import capi, data, guided

pytestmark = pytest.mark.gpu


def lineage_nodes(anc, K):
    """Live nodes after K steps: sum over generations k = 0..K of the number
    of distinct ancestors at generation k of the particles at K."""
    P = anc.shape[1]
    live = np.ones(P, bool)
    total = P
    for k in range(K - 1, -1, -1):
        prev = np.zeros(P, bool)
        prev[anc[k][live]] = True
        live = prev
        total += int(live.sum())
    return total


def _pz_filters(ctx, record, resampler=capi.SYSTEMATIC, ess_rel=1.0, P=4096, T=100):
    gm = capi.make_model("pz", precision=capi.F64)
    _, y = data.simulate(ctx, gm, [2.0, 2.0], T, 1.0, seed=2024)
    pf = capi.ParticleFilter(ctx, gm, P, T, 1.0, seed=77, resampler=resampler, ess_rel=ess_rel,
                             record=record)
    pf.set_observations(y)
    pf.init()
    return pf


@pytest.mark.parametrize("resampler,ess_rel", [(capi.SYSTEMATIC, 1.0), (capi.MULTINOMIAL, 0.5)])
def test_tree_paths_equal_full_history_and_nodes_equal_lineages(ctx, torch_gpu, resampler,
                                                                ess_rel):
    P, T = 4096, 100
    full = _pz_filters(ctx, "paths", resampler, ess_rel, P, T)
    tree = _pz_filters(ctx, "tree", resampler, ess_rel, P, T)
    rng = np.random.default_rng(5)
    idx = rng.choice(P, 64, replace=False)
    for K in (1, 2, 10, 37, 100):
        n = K - full.position()
        full.step(n)
        tree.step(n)
        np.testing.assert_array_equal(tree.paths(idx), full.paths(idx))
        anc = np.array([full.ancestors(k) for k in range(K)])
        live, cap = tree.cache_nodes()
        expect = lineage_nodes(anc, K)
        print(f"resampler {resampler} K={K}: live nodes {live} (lineages {expect}), capacity "
              f"{cap}, full history {(K + 1) * P}", flush=True)
        assert live == expect
    assert tree.term() == full.term()
    p0, path0 = full.sample_path()
    p1, path1 = tree.sample_path()
    assert p0 == p1
    np.testing.assert_array_equal(path0, path1)
    full.close()
    tree.close()


def test_tree_on_fp32_lorenz96_equals_full_history(ctx, torch_gpu):
    P, T, DT = 1 << 18, 120, 0.05
    gm = capi.make_model("l96", d=40, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    mean0 = data.l96_spinup(ctx, d=40)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=2024)
    runs = {}
    for rec in ("paths", "tree"):
        pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=77, ess_rel=0.5, record=rec)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(T)
        runs[rec] = pf
    idx = np.arange(0, P, P // 32)
    np.testing.assert_array_equal(runs["tree"].paths(idx), runs["paths"].paths(idx))
    live, cap = runs["tree"].cache_nodes()
    print(f"L96 d=40 N=2^18 T={T}: live nodes {live} ({live / P:.2f} P), capacity {cap / P:.0f} P,"
          f" full history {T + 1} P", flush=True)
    assert live < 0.1 * (T + 1) * P
    for pf in runs.values():
        pf.close()


def test_tree_holds_a_history_beyond_hbm(ctx, torch_gpu):
    """L96 d=200 FP32, N=2^20, T=500: the full record would take
    (T + 1) x 200 x 2^20 x 4 B = 420 GB, more than the GPU's 288 GB."""
    torch = torch_gpu
    P, T, DT = 1 << 20, 500, 0.05
    gm = guided.model(8.0)
    full_bytes = (T + 1) * gm.d * P * 4
    assert full_bytes > torch.cuda.get_device_properties(0).total_memory
    mean0 = data.l96_spinup(ctx, d=gm.d)
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=7)
    pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=11, ess_rel=0.5, record="tree")
    pf.set_observations(y)
    pf.init(mean0)
    peak = 0
    for k in range(0, T, 50):
        pf.step(50)
        live, cap = pf.cache_nodes()
        peak = max(peak, cap)
    out = pf.outputs()
    idx = np.array([0, 1, P // 2, P - 1])
    paths = pf.paths(idx)
    X, _, _ = pf.state()
    cur = X[:, idx.tolist()].double().cpu().numpy().T
    print(f"L96 d=200 N=2^20 T={T}: resampled {int(out['resampled'].sum())} steps; live nodes "
          f"{live} ({live / P:.2f} P), node capacity {peak} ({peak * (gm.d * 4 + 8) / 1e9:.1f} GB)"
          f" against a full record of {full_bytes / 1e9:.0f} GB", flush=True)
    assert out["resampled"].sum() > 100
    assert paths.shape == (4, T + 1, gm.d)
    np.testing.assert_array_equal(paths[:, T, :], cur)  # each path ends at its particle
    # and starts from an initial draw x0 ~ N(mean0, I)
    assert np.abs(paths[:, 0, :] - mean0).max() < 7.0
    assert peak * (gm.d * 4 + 8) < 0.1 * full_bytes
    pf.close()
