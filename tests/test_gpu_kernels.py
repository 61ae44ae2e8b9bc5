"""Kernel-level parity of the HIP path (through the C ABI) with the oracle.

Integer and index work (Philox words, uniforms, offspring, ancestors,
permutations, gathers) must be bit-exact; floating-point results are
checked to the tolerance written in each test."""
import ctypes

import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import dev, empty, host, l96_spun_up, oracle_model

pytestmark = pytest.mark.gpu
L = capi.lib
_LIVE = []  # tensors passed to the library stay alive until the test ends


@pytest.fixture(autouse=True)
def _keep_alive():
    yield
    _LIVE.clear()


def _c(x):
    _LIVE.append(x)
    return ctypes.c_void_p(x.data_ptr())


# ---------------------------------------------------------------- random numbers
def test_philox_bit_exact(ctx, orc, torch_gpu):
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, size=(2000, 4), dtype=np.uint64).astype(np.uint32)
    ctr[0] = 0
    ctr[1] = 0xFFFFFFFF
    key = (0xA4093822, 0x299F31D0)
    c = dev(ctr.view(np.int32))
    out = empty((2000, 4), torch_gpu.int32)
    capi.check(L().bihip_philox(ctx.h, _c(c), key[0], key[1], 2000, _c(out)))
    got = host(out).view(np.uint32)
    for i in range(0, 2000, 7):
        assert list(got[i]) == orc.philox(list(map(int, ctr[i])), key), i


def test_uniforms_bit_exact(ctx, orc, torch_gpu):
    out = empty(4099, torch_gpu.float64)
    capi.check(L().bihip_uniforms(ctx.h, 12345678901, 17, capi.TAG_MULTINOMIAL, 3, 4099, _c(out)))
    np.testing.assert_array_equal(host(out), orc.uniforms(12345678901, 17, orc.TAG_MULTINOMIAL,
                                                          4099, batch=3))


@pytest.mark.parametrize("precision", [capi.F64, capi.F32])
def test_normals(ctx, orc, torch_gpu, precision):
    n = 1001
    dt = torch_gpu.float64 if precision == capi.F64 else torch_gpu.float32
    out = empty(n, dt)
    capi.check(L().bihip_normals(ctx.h, 77, 5, 9, capi.TAG_NOISE, 1, n, precision, _c(out)))
    bits = 32 if precision == capi.F64 else 24
    ref = orc.normals(77, 5, 9, orc.TAG_NOISE, n, bits=bits, batch=1)
    tol = 1e-13 if precision == capi.F64 else 2e-6
    np.testing.assert_allclose(host(out), ref, rtol=tol, atol=tol)


# ---------------------------------------------------------------- model kernels
MODELS = [
    ("pz", capi.RK4, capi.F64, 0),
    ("pz", capi.DOPRI5, capi.F64, 0),
    ("pz", capi.RK4, capi.F32, 0),
    ("lg", capi.DOPRI5, capi.F64, 0),
    ("l96", capi.DOPRI5, capi.F64, 0),
    ("l96", capi.DOPRI5, capi.F64, 1),
    ("l96", capi.RK4, capi.F64, 0),
    ("l96", capi.DOPRI5, capi.F32, 0),
    ("l96", capi.DOPRI5, capi.F32, 1),
]


def _setup(orc, kind, integ, prec, quirk, P):
    kw = dict(integrator=integ, precision=prec, host_quirk=quirk)
    if kind == "l96":
        kw["d"] = 40
        if integ == capi.RK4:
            kw["h0"] = 0.01
    if kind == "pz" and integ == capi.DOPRI5:
        kw["h0"] = 0.1
    gm = capi.make_model(kind, **kw)
    om = oracle_model(orc, gm)
    mean0 = l96_spun_up(orc) if kind == "l96" else None
    X0 = orc.init_samples(om, 21, P, mean0=mean0)
    return gm, om, X0, mean0


def _tol(prec, kind):
    if prec == capi.F64:
        return 1e-10
    return 2e-3 if kind == "l96" else 2e-5


@pytest.mark.parametrize("kind,integ,prec,quirk", MODELS)
def test_init_samples(ctx, orc, torch_gpu, kind, integ, prec, quirk):
    P = 1000
    gm, om, X0, mean0 = _setup(orc, kind, integ, prec, quirk, P)
    dt = torch_gpu.float64 if prec == capi.F64 else torch_gpu.float32
    X = empty((gm.d, P), dt)
    m0 = None if mean0 is None else mean0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    capi.check(L().bihip_init_samples(ctx.h, ctypes.byref(gm), 21, 0, P, _c(X), P, m0))
    tol = 1e-13 if prec == capi.F64 else 2e-6
    np.testing.assert_allclose(host(X).astype(np.float64), X0, rtol=tol, atol=tol)


@pytest.mark.parametrize("kind,integ,prec,quirk", MODELS)
def test_transition_parity(ctx, orc, torch_gpu, kind, integ, prec, quirk):
    P = 2000
    gm, om, X0, _ = _setup(orc, kind, integ, prec, quirk, P)
    npdt = np.float64 if prec == capi.F64 else np.float32
    dt = torch_gpu.float64 if prec == capi.F64 else torch_gpu.float32
    X0 = X0.astype(npdt)
    Xin = dev(X0)
    Xout = empty((gm.d, P), dt)
    ns = empty(P, torch_gpu.int32)
    t1, t2 = (3.0, 4.0) if kind in ("pz", "lg") else (0.25, 0.30)
    capi.check(L().bihip_transition(ctx.h, ctypes.byref(gm), 5, 0, 4, t1, t2, P, _c(Xin),
                                    _c(Xout), P, None, _c(ns)))
    ref, rns = orc.transition(om, 5, 4, t1, t2, X0.astype(np.float64), return_steps=True)
    got = host(Xout).astype(np.float64)
    tol = _tol(prec, kind)
    np.testing.assert_allclose(got, ref, rtol=tol, atol=tol)
    if prec == capi.F64:
        np.testing.assert_array_equal(host(ns), rns)


@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_l96_d200_transition_and_density(ctx, orc, torch_gpu, prec):
    """d=200: eight lanes per particle with 25-dimension slices."""
    P = 700
    gm = capi.make_model("l96", d=200, precision=prec)
    om = oracle_model(orc, gm)
    mean0 = l96_spun_up(orc, d=200)
    npdt = np.float64 if prec == capi.F64 else np.float32
    dt = torch_gpu.float64 if prec == capi.F64 else torch_gpu.float32
    X0 = orc.init_samples(om, 21, P, mean0=mean0).astype(npdt)
    Xin, Xout = dev(X0), empty((200, P), dt)
    capi.check(L().bihip_transition(ctx.h, ctypes.byref(gm), 5, 0, 4, 0.25, 0.30, P, _c(Xin),
                                    _c(Xout), P, None, None))
    ref = orc.transition(om, 5, 4, 0.25, 0.30, X0.astype(np.float64))
    tol = _tol(prec, "l96")
    got = host(Xout).astype(np.float64)
    np.testing.assert_allclose(got, ref, rtol=tol, atol=tol)
    rng = np.random.default_rng(4)
    y = rng.normal(size=100) + 2.0
    lws0 = rng.normal(size=P)
    lws = dev(lws0)
    capi.check(L().bihip_obs_logdensity(ctx.h, ctypes.byref(gm), _c(dev(y)), P, _c(Xout), P,
                                        _c(lws)))
    ref_l = orc.obs_logdensity(om, y, got, lws0)
    np.testing.assert_allclose(host(lws), ref_l, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_transition_gathers_rows(ctx, orc, torch_gpu, prec):
    P = 3000
    gm, om, X0, _ = _setup(orc, "l96", capi.DOPRI5, prec, 0, P)
    npdt = np.float64 if prec == capi.F64 else np.float32
    dt = torch_gpu.float64 if prec == capi.F64 else torch_gpu.float32
    rng = np.random.default_rng(3)
    o = rng.multinomial(P, np.ones(P) / P)
    a = orc.offspring_to_ancestors_permute(o)
    X0 = X0.astype(npdt)
    Xin, Xout, as_ = dev(X0), empty((gm.d, P), dt), dev(a)
    capi.check(L().bihip_transition(ctx.h, ctypes.byref(gm), 8, 0, 2, 0.0, 0.05, P, _c(Xin),
                                    _c(Xout), P, _c(as_), None))
    ref = orc.transition(om, 8, 2, 0.0, 0.05, X0.astype(np.float64)[:, a])
    tol = _tol(prec, "l96")
    np.testing.assert_allclose(host(Xout).astype(np.float64), ref, rtol=tol, atol=tol)


@pytest.mark.parametrize("kind", ["pz", "l96", "lg"])
@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_obs_logdensity_parity(ctx, orc, torch_gpu, kind, prec):
    P = 1500
    gm, om, X0, _ = _setup(orc, kind, capi.DOPRI5 if kind != "pz" else capi.RK4, prec, 0, P)
    npdt = np.float64 if prec == capi.F64 else np.float32
    X0 = X0.astype(npdt)
    rng = np.random.default_rng(4)
    y = np.abs(rng.normal(size=capi.nobs(gm))) + 0.5
    lws0 = rng.normal(size=P)
    X, yd, lws = dev(X0), dev(y), dev(lws0)
    capi.check(L().bihip_obs_logdensity(ctx.h, ctypes.byref(gm), _c(yd), P, _c(X), P, _c(lws)))
    ref = orc.obs_logdensity(om, y, X0.astype(np.float64), lws0)
    np.testing.assert_allclose(host(lws), ref, rtol=1e-13, atol=1e-12)


# ---------------------------------------------------------------- reductions and scan
def _weights(kind, P, rng):
    if kind == "normal":
        return rng.normal(scale=3.0, size=P)
    if kind == "nan":
        w = rng.normal(size=P)
        w[rng.random(P) < 0.3] = np.nan
        return w
    if kind == "neginf":
        w = rng.normal(size=P)
        w[rng.random(P) < 0.5] = -np.inf
        return w
    if kind == "onehot":
        w = np.full(P, -np.inf)
        w[P // 2] = 1.5
        return w
    raise ValueError(kind)


@pytest.mark.parametrize("P", [1, 5, 2048, 2049, 100003, 1 << 20])
@pytest.mark.parametrize("kind", ["normal", "nan", "neginf", "onehot"])
def test_reductions(ctx, orc, torch_gpu, P, kind):
    rng = np.random.default_rng(P)
    w = _weights(kind, P, rng)
    if kind == "nan" and np.all(np.isnan(w)):
        w[0] = 0.0
    d = dev(w)
    ess, lW, lse, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    capi.check(L().bihip_ess_reduce(ctx.h, _c(d), P, ctypes.byref(ess), ctypes.byref(lW)))
    capi.check(L().bihip_logsumexp_reduce(ctx.h, _c(d), P, ctypes.byref(lse)))
    capi.check(L().bihip_max_reduce(ctx.h, _c(d), P, ctypes.byref(mx)))
    re, rl = orc.ess_reduce(w)
    assert ess.value == pytest.approx(re, rel=1e-12)
    assert lW.value == pytest.approx(rl, rel=1e-13, abs=1e-13)
    assert lse.value == pytest.approx(orc.logsumexp_reduce(w), rel=1e-13, abs=1e-13)
    assert mx.value == orc.max_reduce(w)


@pytest.mark.parametrize("P", [1, 7, 2048, 2049, 65537, 1 << 20, (1 << 20) + 5, (1 << 22) + 3])
def test_sumexp_scan_parity_and_determinism(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(P + 1)
    w = rng.normal(scale=2.0, size=P)
    w[rng.random(P) < 0.05] = -np.inf
    d = dev(w)
    Ws = empty(P, torch_gpu.float64)
    mx = ctypes.c_double()
    capi.check(L().bihip_sumexpu_inclusive_scan(ctx.h, _c(d), P, _c(Ws), ctypes.byref(mx)))
    got = host(Ws)
    ref, rmx = orc.sumexpu_inclusive_scan(w)
    assert mx.value == rmx
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    assert np.all(np.diff(got) >= 0)
    for _ in range(3):  # bit-identical on every run
        capi.check(L().bihip_sumexpu_inclusive_scan(ctx.h, _c(d), P, _c(Ws), None))
        np.testing.assert_array_equal(host(Ws), got)


def test_sumexp_scan_paths_agree_bit_for_bit(ctx, torch_gpu):
    """A grid resident at once runs the single-pass look-back scan, a larger
    one the tiled reduce-then-scan; with the maximum in the first 2^20
    elements both scan those elements to the same bits."""
    P, Q = (1 << 22) + 3, 1 << 20
    rng = np.random.default_rng(77)
    w = rng.normal(scale=2.0, size=P)
    w[0] = w.max() + 1.0
    d = dev(w)
    full = empty(P, torch_gpu.float64)
    head = empty(Q, torch_gpu.float64)
    capi.check(L().bihip_sumexpu_inclusive_scan(ctx.h, _c(d), P, _c(full), None))
    capi.check(L().bihip_sumexpu_inclusive_scan(ctx.h, _c(d), Q, _c(head), None))
    np.testing.assert_array_equal(host(full)[:Q], host(head))


# ---------------------------------------------------------------- resamplers (bit-exact)
def _Ws(orc, P, rng, scale=2.0):
    Ws, _ = orc.sumexpu_inclusive_scan(rng.normal(scale=scale, size=P))
    return Ws


@pytest.mark.parametrize("P", [1, 2, 3, 100, 4097, 1 << 20])
def test_systematic_bit_exact(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(10 + P)
    Ws = _Ws(orc, P, rng)
    d = dev(Ws)
    O = empty(P, torch_gpu.int32)
    for a in (0.0, 0.5, float(rng.random()), 1.0 - 2.0 ** -53):
        capi.check(L().bihip_systematic_cumulative_offspring(ctx.h, _c(d), P, P, a, _c(O)))
        ref, rc = orc.systematic_cumulative_offspring(Ws, a)
        np.testing.assert_array_equal(host(O), ref)


@pytest.mark.parametrize("P", [1, 5, 4097, 1 << 18])
def test_stratified_bit_exact(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(20 + P)
    Ws = _Ws(orc, P, rng)
    al = rng.random(P)
    O = empty(P, torch_gpu.int32)
    capi.check(L().bihip_stratified_cumulative_offspring(ctx.h, _c(dev(Ws)), P, P, _c(dev(al)),
                                                         _c(O)))
    ref, _ = orc.stratified_cumulative_offspring(Ws, al)
    np.testing.assert_array_equal(host(O), ref)


@pytest.mark.parametrize("P", [1, 5, 4097, 1 << 18])
def test_multinomial_bit_exact(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(30 + P)
    Ws = _Ws(orc, P, rng, scale=4.0)
    U = rng.random(P)
    U[0] = 0.0
    U[-1] = 1.0 - 2.0 ** -53
    a = empty(P, torch_gpu.int32)
    capi.check(L().bihip_multinomial_ancestors(ctx.h, _c(dev(Ws)), P, P, _c(dev(U)), _c(a)))
    ref, _ = orc.multinomial_ancestors(Ws, U)
    np.testing.assert_array_equal(host(a), ref)


def _ws_cases(P, rng):
    """Cumulative weights of every shape the chunked search treats differently:
    spread, one particle holding everything, a few heavy particles among
    light ones (chunks with few draws next to chunks with many), and weights
    concentrated at the end."""
    w = rng.random(P)
    yield "spread", np.cumsum(w)
    w = np.zeros(P)
    w[P // 3] = 1.0
    yield "one", np.cumsum(w)
    w = np.full(P, 1e-9)
    w[rng.choice(P, min(P, 5), replace=False)] = 1.0
    yield "few heavy", np.cumsum(w)
    w = np.exp(np.linspace(-30.0, 0.0, P))
    yield "rising", np.cumsum(w)


@pytest.mark.parametrize("P,n", [(1, 1000), (7, 5), (2048, 2048), (2049, 70000),
                                 ((1 << 20) + 3, 1 << 20), ((1 << 22) + 7, 300000)])
def test_multinomial_chunked_search_bit_exact(ctx, orc, torch_gpu, P, n):
    """Draws sorted by chunk of the cumulative weights and searched in LDS or
    in global memory give the oracle's ancestors draw for draw, and the
    filter's offspring counts (through the C ABI's ancestors) for every shape
    of weights, including n != P and the last uniform below 1."""
    rng = np.random.default_rng(P + n)
    for name, Ws in _ws_cases(P, rng):
        U = rng.random(n)
        U[0] = 0.0
        U[-1] = 1.0 - 2.0 ** -53
        a = empty(n, torch_gpu.int32)
        capi.check(L().bihip_multinomial_ancestors(ctx.h, _c(dev(Ws)), P, n, _c(dev(U)), _c(a)))
        ref, _ = orc.multinomial_ancestors(Ws, U)
        np.testing.assert_array_equal(host(a), ref, name)


def test_degenerate_weights_raise(ctx, torch_gpu):
    Ws = dev(np.zeros(10))
    O = empty(10, torch_gpu.int32)
    rc = L().bihip_systematic_cumulative_offspring(ctx.h, _c(Ws), 10, 10, 0.5, _c(O))
    assert rc == capi.DEGENERATE


def _offspring_cases(P, rng):
    yield rng.multinomial(P, np.ones(P) / P)
    yield np.ones(P, dtype=np.int64)
    o = np.zeros(P, dtype=np.int64)
    o[P // 3] = P
    yield o
    o = np.zeros(P, dtype=np.int64)
    o[-1] = P
    yield o
    yield rng.multinomial(P, rng.dirichlet(np.full(P, 0.05)))


@pytest.mark.parametrize("P", [1, 2, 3, 31, 2048, 2049, 100000, 1 << 20, (3 << 20) + 5,
                               4 << 20, (8 << 20) + 5])
def test_permute_bit_exact(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(40 + P)
    for o in _offspring_cases(P, rng):
        O = np.cumsum(o).astype(np.int32)
        a = empty(P, torch_gpu.int32)
        capi.check(L().bihip_cumulative_offspring_to_ancestors_permute(ctx.h, _c(dev(O)), P,
                                                                       _c(a)))
        ref = orc.cumulative_offspring_to_ancestors_permute(O)
        np.testing.assert_array_equal(host(a), ref)
        capi.check(L().bihip_offspring_to_ancestors_permute(ctx.h, _c(dev(o, np.int32)), P,
                                                            _c(a)))
        np.testing.assert_array_equal(host(a), ref)


@pytest.mark.parametrize("P", [1, 9, 4096, 100003])
def test_offspring_conversions(ctx, orc, torch_gpu, P):
    rng = np.random.default_rng(50 + P)
    anc = rng.integers(0, P, size=P).astype(np.int32)
    os_ = empty(P, torch_gpu.int32)
    capi.check(L().bihip_ancestors_to_offspring(ctx.h, _c(dev(anc)), P, P, _c(os_)))
    ref_os = orc.ancestors_to_offspring(anc, P)
    np.testing.assert_array_equal(host(os_), ref_os)
    O = empty(P, torch_gpu.int32)
    capi.check(L().bihip_offspring_to_cumulative(ctx.h, _c(os_), P, _c(O)))
    np.testing.assert_array_equal(host(O), orc.offspring_to_cumulative(ref_os))
    a = empty(P, torch_gpu.int32)
    capi.check(L().bihip_cumulative_offspring_to_ancestors(ctx.h, _c(O), P, _c(a)))
    np.testing.assert_array_equal(host(a), orc.cumulative_offspring_to_ancestors(
        orc.offspring_to_cumulative(ref_os)))


@pytest.mark.parametrize("elem", [4, 8])
def test_gather_rows_in_place(ctx, orc, torch_gpu, elem):
    P, d = 5000, 7
    rng = np.random.default_rng(60)
    a = orc.offspring_to_ancestors_permute(rng.multinomial(P, np.ones(P) / P))
    X = rng.normal(size=(d, P)).astype(np.float64 if elem == 8 else np.float32)
    Xd = dev(X)
    capi.check(L().bihip_gather_rows(ctx.h, _c(dev(a)), P, d, _c(Xd), P, elem))
    np.testing.assert_array_equal(host(Xd), X[:, a])


@pytest.mark.parametrize("kind,integ,prec", [("l96", capi.DOPRI5, capi.F64),
                                             ("l96", capi.DOPRI5, capi.F32),
                                             ("pz", capi.RK4, capi.F64),
                                             ("lg", capi.DOPRI5, capi.F64)])
def test_flow_is_noise_free_transition(ctx, orc, torch_gpu, kind, integ, prec):
    """bihip_flow (RK4Integrator / DOPRI5Integrator::update) equals the oracle's
    transition with the noise switched off (noise_sd = 0; PZ: sigma = 0)."""
    P = 1500
    gm, om, X0, _ = _setup(orc, kind, integ, prec, 0, P)
    if kind == "pz":
        om.par[1] = 0.0
    else:
        om.par[1] = 0.0
    npdt = np.float64 if prec == capi.F64 else np.float32
    X = dev(X0.astype(npdt))
    ns = empty(P, torch_gpu.int32)
    t1, t2 = (0.0, 1.0) if kind != "l96" else (0.0, 0.05)
    capi.check(L().bihip_flow(ctx.h, ctypes.byref(gm), t1, t2, P, _c(X), P, None, _c(ns)))
    ref, rns = orc.transition(om, 1, 1, t1, t2, X0.astype(npdt).astype(np.float64),
                              return_steps=True)
    tol = _tol(prec, kind)
    np.testing.assert_allclose(host(X).astype(np.float64), ref, rtol=tol, atol=tol)
    if prec == capi.F64:
        np.testing.assert_array_equal(host(ns), rns)
