"""Pins the oracle's reductions, scans and resamplers (CPU only)."""
import itertools
from fractions import Fraction

import numpy as np
import pytest
from scipy import special, stats


# ---------------------------------------------------------------- reductions
def test_ess_equal_weights(orc):
    ess, lW = orc.ess_reduce(np.full(100, -3.25))
    assert ess == pytest.approx(100.0, rel=1e-14)
    assert lW == pytest.approx(-3.25, abs=1e-14)


def test_ess_one_hot(orc):
    lws = np.full(50, -np.inf)
    lws[17] = 2.0
    ess, lW = orc.ess_reduce(lws)
    assert ess == 1.0
    assert lW == pytest.approx(2.0 - np.log(50), rel=1e-15)


def test_ess_long_double_brute_force(orc):
    rng = np.random.default_rng(1)
    lws = rng.normal(scale=5.0, size=257)
    ess, lW = orc.ess_reduce(lws)
    w = np.exp(lws.astype(np.longdouble) - lws.max())
    assert ess == pytest.approx(float(w.sum() ** 2 / (w * w).sum()), rel=1e-13)
    assert lW == pytest.approx(float(special.logsumexp(lws) - np.log(257)), rel=1e-13)


def test_ess_shift_invariance(orc):
    rng = np.random.default_rng(2)
    lws = rng.normal(size=64)
    e1, l1 = orc.ess_reduce(lws)
    e2, l2 = orc.ess_reduce(lws + 1000.0)
    assert e1 == pytest.approx(e2, rel=1e-12)
    assert l2 - l1 == pytest.approx(1000.0, rel=1e-13)


def test_nan_weighs_zero_and_is_lowest(orc):
    lws = np.array([np.nan, 1.0, np.nan, 0.0])
    assert orc.max_reduce(lws) == 1.0
    ess, lW = orc.ess_reduce(lws)
    w = np.exp(np.array([1.0, 0.0]) - 1.0)
    assert ess == pytest.approx(w.sum() ** 2 / (w * w).sum(), rel=1e-15)
    assert lW == pytest.approx(1.0 + np.log(w.sum()) - np.log(4), rel=1e-15)
    assert np.isnan(orc.max_reduce(np.array([np.nan, np.nan])))


def test_all_minus_inf_is_degenerate(orc):
    ess, lW = orc.ess_reduce(np.full(8, -np.inf))
    assert np.isnan(ess) and lW == -np.inf
    Ws, mx = orc.sumexpu_inclusive_scan(np.full(8, -np.inf))
    assert mx == -np.inf and np.all(Ws == 0.0)
    _, rc = orc.systematic_cumulative_offspring(Ws, 0.5)
    assert rc == orc.DEGENERATE


def test_logsumexp_scipy(orc):
    rng = np.random.default_rng(3)
    x = rng.normal(scale=20, size=1000)
    assert orc.logsumexp_reduce(x) == pytest.approx(special.logsumexp(x), rel=1e-14)


def test_scan_exact_integer_weights(orc):
    w = np.array([1, 3, 0, 2, 5, 1, 0, 4], dtype=float)
    with np.errstate(divide="ignore"):
        lws = np.log(w)   # log 0 = -inf
    Ws, mx = orc.sumexpu_inclusive_scan(lws)
    assert mx == np.log(5.0)
    np.testing.assert_allclose(Ws * 5.0, np.cumsum(w), rtol=1e-15)
    assert np.all(np.diff(Ws) >= 0)


# ---------------------------------------------------------------- offspring
def _points_below(Ws, n, a):
    """#{k : (k + 1 - a) W / n <= Ws_i} in exact rationals, capped at n."""
    W = Fraction(Ws[-1])
    out = []
    for w in Ws:
        w = Fraction(w)
        c = 0
        for k in range(n):
            if (k + 1 - Fraction(a)) * W / n <= w:
                c += 1
        out.append(min(c, n))
    return out


def test_systematic_exact_rationals(orc):
    rng = np.random.default_rng(4)
    for trial in range(30):
        P = int(rng.integers(1, 9))
        w = rng.integers(0, 5, size=P).astype(float)
        if w.sum() == 0:
            w[0] = 1
        Ws = np.cumsum(w) / 8.0  # dyadic, exact
        for a in rng.random(size=5):
            O, rc = orc.systematic_cumulative_offspring(Ws, float(a))
            assert rc == orc.OK
            assert list(O) == _points_below(Ws, P, a)


def test_systematic_invariants(orc):
    rng = np.random.default_rng(5)
    for P in (1, 2, 7, 100, 4096):
        lws = rng.normal(scale=2, size=P)
        Ws, _ = orc.sumexpu_inclusive_scan(lws)
        O, rc = orc.systematic_cumulative_offspring(Ws, float(rng.random()))
        assert O[-1] == P and np.all(np.diff(O) >= 0) and O[0] >= 0
        o = np.diff(np.concatenate([[0], O]))
        expect = P * np.exp(lws - lws.max()) / np.exp(lws - lws.max()).sum()
        assert np.all(o >= np.floor(expect) - 1e-9) and np.all(o <= np.ceil(expect) + 1e-9)


def test_stratified_invariants(orc):
    rng = np.random.default_rng(6)
    for P in (1, 3, 50, 1000):
        lws = rng.normal(size=P)
        Ws, _ = orc.sumexpu_inclusive_scan(lws)
        alphas = rng.random(P)
        O, rc = orc.stratified_cumulative_offspring(Ws, alphas)
        assert rc == orc.OK and O[-1] == P and np.all(np.diff(O) >= 0)
        # one point per stratum k at (k + 1 - alpha_k) W / n, each counted once
        W = Ws[-1]
        pts = (np.arange(P) + 1 - alphas) / P * W
        counts = np.searchsorted(Ws, pts, side="left")
        np.testing.assert_array_equal(np.bincount(counts, minlength=P),
                                      np.diff(np.concatenate([[0], O])))


def test_multinomial_definition_and_frequencies(orc):
    rng = np.random.default_rng(7)
    w = np.array([0.1, 0.0, 0.5, 0.2, 0.2])
    Ws = np.cumsum(w)
    U = rng.random(200000)
    a, rc = orc.multinomial_ancestors(Ws, U)
    assert rc == orc.OK
    np.testing.assert_array_equal(a, np.searchsorted(Ws, Ws[-1] * U, side="right"))
    counts = np.bincount(a, minlength=5)
    assert counts[1] == 0
    keep = w > 0
    chi = stats.chisquare(counts[keep], f_exp=len(U) * w[keep] / w.sum())
    assert chi.pvalue > 1e-4


# ---------------------------------------------------------------- permutation
def _reference_loop(Os):
    """Literal restatement of the reference host loop with its k == 0 case
    (share/src/bi/host/resampler/ResamplerHost.hpp:141-167)."""
    P = len(Os)
    a = [None] * P
    k = 0
    for i in range(P):
        O1 = Os[i - 1] if i > 0 else 0
        o = Os[i] - O1
        if o > 0:
            a[i] = i
            o -= 1
        if k == 0 and o > 0:
            if Os[k] > 0:
                k += 1
            else:
                a[k] = i
                k += 1
                o -= 1
        for _ in range(o):
            while Os[k] - Os[k - 1] > 0:
                k += 1
            a[k] = i
            k += 1
    return a


def _rank_formula(o):
    """The GPU's formulation: exclusive scans E of max(o-1,0) and F of [o==0]."""
    o = np.asarray(o)
    e = np.maximum(o - 1, 0)
    f = (o == 0).astype(int)
    E = np.concatenate([[0], np.cumsum(e)[:-1]])
    F = np.concatenate([[0], np.cumsum(f)[:-1]])
    a = np.empty(len(o), int)
    for k in range(len(o)):
        a[k] = k if o[k] > 0 else np.searchsorted(E, F[k], side="right") - 1
    return a


def _compositions(P):
    for cut in itertools.combinations(range(2 * P - 1), P - 1):
        parts, prev = [], -1
        for c in cut + (2 * P - 1,):
            parts.append(c - prev - 1)
            prev = c
        yield parts


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5, 6])
def test_permute_all_offspring_vectors(orc, P):
    n = 0
    for o in _compositions(P):
        O = np.cumsum(o).astype(np.int32)
        a = orc.cumulative_offspring_to_ancestors_permute(O)
        assert list(a) == _reference_loop(list(O))
        np.testing.assert_array_equal(a, _rank_formula(o))
        np.testing.assert_array_equal(a, orc.offspring_to_ancestors_permute(np.array(o)))
        # invariants: own slot kept, offspring multiset preserved, gather-safe
        for i in range(P):
            if o[i] > 0:
                assert a[i] == i
            assert a[a[i]] == a[i]
        np.testing.assert_array_equal(np.bincount(a, minlength=P), o)
        n += 1
    assert n == len(list(_compositions(P)))


def test_swap_permute_invariants(orc):
    rng = np.random.default_rng(8)
    for _ in range(200):
        P = int(rng.integers(1, 12))
        a = rng.integers(0, P, size=P)
        b = orc.permute(a)
        np.testing.assert_array_equal(np.sort(b), np.sort(a))
        for i in range(P):
            if i in set(a):
                assert b[i] == i


def test_cumulative_to_ancestors(orc):
    O = np.array([0, 2, 2, 5, 6], np.int32)
    np.testing.assert_array_equal(orc.cumulative_offspring_to_ancestors(O), [1, 1, 3, 3, 3, 4])


def test_gather_in_place_equals_out_of_place(orc):
    rng = np.random.default_rng(9)
    for _ in range(50):
        P = int(rng.integers(1, 40))
        o = rng.multinomial(P, np.ones(P) / P)
        a = orc.offspring_to_ancestors_permute(o)
        X = rng.normal(size=(3, P))
        np.testing.assert_array_equal(orc.gather_rows(a, X), X[:, a])
