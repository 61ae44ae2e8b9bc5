"""The model-language front end (libbi_amd/dsl): parsing as share/bi.lex and
share/bi.yp define it, the checks of the generator, the generated functors'
form, and the library's description of the generated models (host calls
only, no GPU)."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

from libbi_amd import capi
from libbi_amd.dsl import codegen, parser
from libbi_amd.dsl.codegen import GenError
from libbi_amd.dsl.parser import BiSyntaxError

HERE = os.path.dirname(os.path.abspath(__file__))
BI = os.path.join(os.path.dirname(HERE), "libbi_amd", "bi")


def _gen(text):
    return codegen.analyse_text(text)


def test_tokens_follow_the_lexer():
    toks = parser.tokenize("x <- a**2 .* b // c\n/* d */ 1.5e-3 .5 2 'RK4' dx/dt")
    kinds = [(t.kind, t.text) for t in toks[:-1]]
    assert kinds == [("ident", "x"), ("op", "<-"), ("ident", "a"), ("op", "**"),
                     ("number", "2"), ("op", ".*"), ("ident", "b"), ("number", "1.5e-3"),
                     ("number", ".5"), ("number", "2"), ("string", "'RK4'"), ("ident", "dx"),
                     ("op", "/"), ("keyword", "dt")]
    assert toks[7].line == 2


def test_precedence_and_associativity():
    """'**' binds tighter than '*', which binds tighter than '+'; binary
    operators associate to the left (share/bi.yp:330-420)."""
    m = parser.parse("model M { state x\n sub initial { x <- 1 - 2 - 3 * 4 ** 2 ** 3 } }")
    e = m.block("initial").items[0].expr
    # ((1 - 2) - (3 * ((4 ** 2) ** 3)))
    assert isinstance(e, parser.Binary) and e.op == "-"
    assert e.left.op == "-" and e.left.left.value == 1 and e.left.right.value == 2
    r = e.right
    assert r.op == "*" and r.left.value == 3
    assert r.right.op == "**" and r.right.left.op == "**" and r.right.right.value == 3
    # the analyser's constant evaluation agrees with Python on the same tree
    assert codegen._const_eval(e, {}, 0) == (1 - 2) - 3 * ((4 ** 2) ** 3)


def test_statements_without_semicolons():
    m = parser.parse("""model M {
      const a = 2, b = 3
      param p
      state x, y; noise w
      sub transition { w ~ gaussian(0, p)
        x <- x + w
        y <- -x }
    }""")
    assert m.consts[0][0] == "a" and m.consts[1][0] == "b"
    assert m.names_of("state") == ["x", "y"]
    acts = m.block("transition").items
    assert [a.target for a in acts] == ["w", "x", "y"]
    assert isinstance(acts[2].expr, parser.Unary)


def test_syntax_errors_name_the_line():
    with pytest.raises(BiSyntaxError, match="line 3"):
        parser.parse("model M {\n state x\n sub initial { x <- ( }\n}")
    with pytest.raises(BiSyntaxError):
        parser.parse("model M { state x } extra")
    with pytest.raises(BiSyntaxError, match="dx/dt"):
        parser.parse("model M { state x\n sub transition { ode { x/dt = 1 } } }")


def test_pz_model_of_the_user_guide():
    g = codegen.load_models([os.path.join(BI, "PZ.bi")])[0]
    assert (g.params, g.states, g.noises, g.obs) == (["mu", "sigma"], ["P", "Z"], ["alpha"],
                                                       ["P_obs"])
    assert g.consts == {"c": 0.25, "e": 0.3, "m_l": 0.1, "m_q": 0.1}
    assert g.integrator == capi.RK4 and g.h == 0.1 and g.delta == 1.0
    assert [s.target for s in g.pre] == ["alpha"] and not g.post
    assert {s.target for s in g.ode} == {"P", "Z"}
    assert [s.target for s in g.observation] == ["P_obs"] and g.observation[0].draw.log


def test_generated_expressions_keep_the_tree():
    """Every operation is parenthesised as parsed, so the device evaluates the
    same tree; constants are compiled in, functions of constants folded."""
    g = codegen.load_models([os.path.join(BI, "PZ.bi")])[0]
    src = codegen.generate_model(g)
    assert "k[0] = ((w_alpha * x[0]) - ((T(0.25) * x[0]) * x[1]));" in src
    assert ("k[1] = (((((T(0.3) * T(0.25)) * x[0]) * x[1]) - (T(0.1) * x[1])) - "
            "((T(0.1) * x[1]) * x[1]));") in src
    assert "w_alpha = (pt_mu + (pt_sigma * zq0[0]));" in src
    # log(2.0) folded to the correctly rounded double
    assert f"exp(({float(np.log(2.0))!r} + (0.2 * (double)zq0[0])))" in src
    assert "lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sg * xy);" in src


def test_draw_counters():
    """The k-th normal draw of a block uses Philox block k/4, word pair k%4;
    uniforms use their own blocks."""
    g = _gen("""model M { param a; state x, y, z, u, v; noise w
      sub initial { x ~ gaussian() y ~ gaussian() z ~ uniform(0, 2) u ~ gaussian()
                    v ~ normal() }
      sub transition { w ~ gaussian() } }""")
    draws = [s.draw for s in g.initial]
    assert [(d.dist, d.index) for d in draws] == [("gaussian", 0), ("gaussian", 1),
                                                  ("uniform", 0), ("gaussian", 2),
                                                  ("gaussian", 3)]
    assert g.normals["initial"] == 4 and g.uniforms["initial"] == 1
    src = codegen.generate_model(g)
    assert "normals4<T>(seed, p, 0u, TAG_INIT, batch, 0u, zq0);" in src
    assert "zq1" not in src.split("void init")[1].split("};")[0]
    assert "xd[4] = (0.0 + (1.0 * (double)zq0[3]));" in src


@pytest.mark.parametrize("text,msg", [
    ("model M { dim n(4)\n dim k(2)\n state x[n, k] }", "more than one dimension"),
    ("model M { dim n(4)\n state x[n]\n sub initial { x[i] <- x[i+1] } }", "outside"),
    ("model M { dim n(4)\n state x[n], y\n sub initial { y <- x } }", "without an index"),
    ("model M { dim n(4, boundary = 'odd')\n state x[n] }", "boundary"),
    ("model M { dim n(2.5)\n state x[n] }", "positive integer"),
    ("model M { dim n(40)\n dim k(30)\n state x[n], z[k] }", "at most 64"),
    ("model M { state x\n sub transition { x ~ wishart(1.0) } }", "not supported"),
    ("model M { state x\n sub transition { ode { dx/dt = -x } } }", "RK4\\(3\\)"),
    ("model M { state x\n sub transition { ode(alg = 'RK4') { dx/dt = y } } }", "unknown name"),
    ("model M { state x; obs y\n sub transition { x <- y } }", "cannot be read"),
    ("model M { state x; obs y\n sub observation { x ~ gaussian() } }", "cannot be the target"),
    ("model M { param " + ", ".join(f"p{i}" for i in range(17)) + "\n state x }", "at most 16"),
    ("model M { state x\n sub transition { x <- x\n ode(alg='RK4') { dx/dt = 1 }\n"
     " ode(alg='RK4') { dx/dt = 2 } } }", "one ode block"),
    ("model M { state x\n sub transition { x <- f(x) } }", "unknown function"),
    ("model M { noise w }", "no state"),
])
def test_unsupported_models_are_rejected(text, msg):
    with pytest.raises(GenError, match=msg):
        _gen(text)


def test_registry_models_compile_into_the_library():
    """Every file in libbi_amd/bi generates; the library knows each model by
    name with the file's variables, ode settings and observation count."""
    paths = sorted(glob.glob(os.path.join(BI, "*.bi")))
    gs = codegen.load_models(paths)
    for g in gs:
        m = capi.generated_model(g.name)
        assert m.model >= capi.MODEL_GENERATED
        assert m.d == g.D == sum(capi.model_var_sizes(m, capi.VAR_STATE))
        assert capi.nobs(m) == g.NOBS
        assert capi.model_vars(m, capi.VAR_PARAM) == g.params
        assert capi.model_vars(m, capi.VAR_STATE) == g.states
        assert capi.model_vars(m, capi.VAR_OBS) == g.obs
        assert capi.model_vars(m, capi.VAR_NOISE) == g.noises
        assert (m.integrator, m.h0, m.atol, m.rtol) == (g.integrator, g.h, g.atol, g.rtol)
        assert capi.lib().bihip_model_name(ctypes.byref(m)).decode() == g.name
    with pytest.raises(capi.BihipError, match="no generated model"):
        capi.generated_model("NoSuchModel")
    with pytest.raises(KeyError):
        capi.generated_model("PZ", nonsense=1.0)
    # hand-written models are not generated ones
    assert capi.model_vars(capi.make_model("pz"), capi.VAR_STATE) == []


def test_netcdf_output_uses_the_model_files_variables(tmp_path):
    from scipy.io import netcdf_file
    m = capi.generated_model("PZ", precision=capi.F64)
    o = capi.PFOutput(tmp_path / "pz.nc", m, 2, 3)
    for k in range(2):
        o.write_time(k, float(k))
        o.write_state(k, np.full((2, 3), float(k)), np.arange(3))
    o.close()
    f = netcdf_file(str(tmp_path / "pz.nc"), "r", mmap=False)
    assert set(f.variables) >= {"P", "Z", "time", "ancestor"}
    f.close()


# Minimal host declarations of what the kernels provide, so that generated
# functors compile with g++ (syntax checks and host evaluation of densities).
HOST_STUB = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#define __device__
#define __forceinline__ inline
#define BIHIP_RK4 0
#define BIHIP_F64 1
#define BIHIP_EINVAL 2
#define BIHIP_MODEL_GENERATED 64
using std::exp; using std::log; using std::sqrt; using std::floor; using std::fmin;
using std::pow; using std::fabs; using std::lgamma; using std::log1p; using std::cos;
namespace bihip {
struct ModelArgs { double h0; double par[16]; };
struct bihip_model_ { int model, precision; };
struct u32x4 { uint32_t x, y, z, w; };
struct GenInfo { const char* name; int d, nobs, integrator, lpp; double h, atol, rtol, delta;
  const char* const* names[5]; const int* sizes[5]; const char* const* state_dims;
  const char* const* noise_dims; };
constexpr double HALF_LOG_TWO_PI = 0.91893853320467274178;
constexpr uint32_t TAG_INIT = 1, TAG_NOISE = 2, TAG_OBS = 6;
inline int set_error(int c, const char*) { return c; }
inline double gen_scaled(double t, double d) { return t / d; }
template <class T> void normals4(uint64_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                 T (&z)[4]) { z[0] = z[1] = z[2] = z[3] = T(0); }
// host stand-in for the keyed Philox words (splitmix64 of the counter): good
// enough words for statistical tests of the samplers
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline u32x4 philox_block(uint64_t seed, uint32_t i, uint32_t s, uint32_t t, uint32_t b,
                          uint32_t blk) {
  uint64_t h = mix64(seed ^ mix64(((uint64_t)i << 32 | s) ^ mix64(((uint64_t)t << 32 | b) ^
                                                                     mix64(blk))));
  uint64_t h2 = mix64(h);
  return {(uint32_t)h, (uint32_t)(h >> 32), (uint32_t)h2, (uint32_t)(h2 >> 32)};
}
inline double cospi(double x) { return std::cos(3.14159265358979323846 * x); }
template <class T> T uniform_word(const u32x4&) { return T(0); }
template <class T, int DL> void slice_normals(uint64_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                              int, T (&z)[DL]) { for (int i = 0; i < DL; ++i) z[i] = T(0); }
struct Grp { int g = 0;
  template <class T> T prev(T v) const { return v; }
  template <class T> T next(T v) const { return v; } };
template <class M, class G, class T> int integrate_rk4(const M&, const G&, T (&)[M::DL],
                                                       double, double, double) { return 0; }
template <class M, class G, class T> int integrate_dopri5(const M&, const G&, T (&)[M::DL],
                                                          double, double, const ModelArgs&) {
  return 0; }
}
#include "GEN_SAMPLERS"
using bihip_model = bihip::bihip_model_;
"""


def _gxx(tmp_path, text, run=False):
    src = tmp_path / "gen.cpp"
    samplers = os.path.join(os.path.dirname(HERE), "libbi_amd", "csrc", "gen_samplers.hpp")
    src.write_text(HOST_STUB.replace("GEN_SAMPLERS", samplers) + text)
    exe = tmp_path / "gen"
    args = ["g++", "-std=c++17", "-O1", "-Wall", str(src), "-o", str(exe)]
    if not run:
        args = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", str(src)]
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    if run:
        return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


def test_generated_header_is_standalone_cxx(tmp_path):
    """The generated text is what build() compiles: it parses as C++ with
    minimal declarations of what the kernels provide (host syntax check)."""
    gs = codegen.load_models(sorted(glob.glob(os.path.join(BI, "*.bi"))))
    _gxx(tmp_path, codegen.generate(gs) + r"""
namespace bihip {
struct Runner { template <class M> int run() {
  M m(ModelArgs{});
  double x[M::DL] = {}; typename M::real_t xt[M::DL] = {}, k[M::DL], xf[M::D] = {};
  Grp g;
  m.rhs(g, xt, k);
  m.template transition<0>(g, xt, 0.0, 1.0, ModelArgs{}, 1, 0, 0, 1);
  double y[M::NOBS > 0 ? M::NOBS : 1] = {};
  (void)m.obs(xt, y, 0);
  m.obs_sample(xf, 1, 0, 0, 1, y);
  m.init(x, 1, 0, 0, 0);
  return 0; } };
int check(const bihip_model* m) { Runner r; return with_generated(m, r); }
}
""")


# one observation of each distribution, given a state x and params a, b
DENSITIES = {
    "poisson": "poisson(x)",
    "negbin": "negbin(x, a)",
    "binomial": "binomial(b, x/(1.0 + x))",
    "gamma": "gamma(a, x)",
    "inverse_gamma": "inverse_gamma(a, x)",
    "exponential": "exponential(x)",
    "beta": "beta(a, x)",
    "gaussian": "gaussian(x, a)",
    "log_normal": "log_normal(x, a)",
    "uniform": "uniform(x, x + a)",
}


def _scipy_logpdf(name, yv, x, a, b):
    from scipy import stats
    if name == "poisson":
        return stats.poisson.logpmf(yv, x)
    if name == "negbin":  # mean x, shape a: n = a, p = a / (a + x)
        return stats.nbinom.logpmf(yv, a, a / (a + x))
    if name == "binomial":
        return stats.binom.logpmf(yv, b, x / (1 + x))
    if name == "gamma":
        return stats.gamma.logpdf(yv, a, scale=x)
    if name == "inverse_gamma":
        return stats.invgamma.logpdf(yv, a, scale=x)
    if name == "exponential":
        return stats.expon.logpdf(yv, scale=1 / x)
    if name == "beta":
        return stats.beta.logpdf(yv, a, x)
    if name == "gaussian":
        return stats.norm.logpdf(yv, x, a)
    if name == "log_normal":
        return stats.lognorm.logpdf(yv, a, scale=np.exp(x))
    if name == "uniform":
        return stats.uniform.logpdf(yv, x, a)
    raise KeyError(name)


def test_observation_densities_against_scipy(tmp_path):
    """The generated log-densities (share/src/bi/pdf/functor.hpp formulas),
    compiled for the host and evaluated on a grid, against scipy.stats."""
    models = []
    for name, dist in DENSITIES.items():
        models.append(codegen.analyse_text(
            f"model D_{name} {{ param a, b\n state x\n obs y\n"
            f" sub observation {{ y ~ {dist} }} }}"))
    cases = {"poisson": ([0, 1, 3, 17], [0.5, 4.0]), "negbin": ([0, 1, 5, 40], [0.7, 12.0]),
             "binomial": ([0, 2, 7], [0.3, 2.5]), "gamma": ([0.2, 1.5, 6.0], [0.8, 3.0]),
             "inverse_gamma": ([0.2, 1.5, 6.0], [0.8, 3.0]),
             "exponential": ([0.0, 0.4, 3.0], [0.5, 2.0]), "beta": ([0.1, 0.5, 0.93], [0.6, 2.5]),
             "gaussian": ([-1.0, 0.3, 2.0], [0.1, 1.2]), "log_normal": ([0.2, 1.0, 4.0], [-0.3, 0.9]),
             "uniform": ([0.35, 0.9, 2.0], [0.3, 0.8])}
    a, b = 1.7, 7.0
    body = ["int main() {", "  bihip::ModelArgs ma{}; ma.par[0] = %r; ma.par[1] = %r;" % (a, b)]
    for g in models:
        name = g.name[2:]
        ys, xs = cases[name]
        for xv in xs:
            for yv in ys:
                body.append(f"  {{ bihip::{g.cname}<double> m(ma); double x[1] = {{{xv!r}}};"
                            f" double y[1] = {{{float(yv)!r}}};"
                            f' std::printf("{name} %.17g %.17g %.17g\\n", x[0], y[0],'
                            f" m.obs(x, y, 0)); }}")
    body.append("  return 0;\n}")
    src = "\n".join(codegen.generate_model(g) for g in models)
    out = _gxx(tmp_path, "namespace bihip {\n" + src + "\n}\n" + "\n".join(body), run=True)
    n = 0
    for line in out.splitlines():
        name, xv, yv, lp = line.split()
        ref = _scipy_logpdf(name, float(yv), float(xv), a, b)
        assert float(lp) == pytest.approx(ref, rel=1e-11, abs=1e-11), (name, xv, yv)
        n += 1
    assert n == sum(len(y) * len(x) for y, x in cases.values())


def test_draws_from_every_distribution():
    """Exponential draws invert a uniform; the rejection samplers take a keyed
    stream per draw element (fire-dependent in the transition)."""
    g = _gen("model M { param l\n state x\n noise w, v[n2]\n dim n2(2)\n obs c\n"
             " sub transition { w ~ exponential(l)\n v ~ gamma(2.0, l)\n x <- x + w }\n"
             " sub observation { c ~ poisson(x) } }")
    src = codegen.generate_model(g)
    assert "w_w = ((-log(T(1.0) - u0)) / pt_l);" in src
    assert ("t_[1] = T(gen_gamma((double)(T(2.0)), (double)(pt_l), DrawStream(seed, p, step, "
            "TAG_NOISE, batch, 0x20000000u + ((fire * 2u + 1u) << 12))));") in src
    assert ("y[0] = gen_poisson((double)((double)x[0]), DrawStream(seed, p, step, TAG_OBS, "
            "batch, 0x20000000u + ((0u + 0u) << 12)));") in src


def test_netcdf_output_of_a_vector_model(tmp_path):
    """A generated model's state over a dimension is written as (nr, dim, np)."""
    from scipy.io import netcdf_file
    m = capi.generated_model("Lorenz96", precision=capi.F32)
    o = capi.PFOutput(tmp_path / "l96.nc", m, 2, 3)
    X = np.arange(8 * 3, dtype=np.float32).reshape(8, 3)
    for k in range(2):
        o.write_time(k, 0.05 * k)
        o.write_state(k, X + k, np.arange(3))
    o.close()
    f = netcdf_file(str(tmp_path / "l96.nc"), "r", mmap=False)
    assert f.dimensions["n"] == 8 and f.variables["x"].dimensions == ("nr", "n", "np")
    np.testing.assert_array_equal(np.array(f.variables["x"].data)[1], X + 1)
    f.close()


def _scipy_prior(name, a, b):
    from scipy import stats
    return {"gaussian": lambda: stats.norm(a, b), "log_normal": lambda: stats.lognorm(b, scale=np.exp(a)),
            "uniform": lambda: stats.uniform(a, b - a), "exponential": lambda: stats.expon(scale=1 / a),
            "gamma": lambda: stats.gamma(a, scale=b), "inverse_gamma": lambda: stats.invgamma(a, scale=b),
            "beta": lambda: stats.beta(a, b)}[name]()


PRIORS = {  # the parameter blocks of the model files, per param
    "LinearGaussian": [("gamma", 2.0, 0.25), ("inverse_gamma", 3.0, 1.0),
                       ("log_normal", np.log(0.5), 0.2), ("gaussian", 0.0, 1.0),
                       ("uniform", 0.5, 2.0)],
    "SIR": [("gamma", 4.0, 0.15), ("beta", 2.0, 6.0), ("exponential", 0.1, 0.0),
            ("uniform", 0.0, 0.3)],
    "PZ": [("uniform", 0.0, 1.0), ("uniform", 0.0, 0.5)],
}


@pytest.mark.parametrize("model", sorted(PRIORS))
def test_parameter_block_priors_against_scipy(model):
    """Draws from a model file's parameter block (host, keyed) pass a
    Kolmogorov-Smirnov test against scipy's distribution, and the log-density
    equals scipy's."""
    from scipy import stats
    m = capi.generated_model(model)
    spec = PRIORS[model]
    n = len(spec)
    draws = np.array([capi.prior_sample(m, n, 11, j) for j in range(4000)])
    for i, (name, a, b) in enumerate(spec):
        p = stats.kstest(draws[:, i], _scipy_prior(name, a, b).cdf).pvalue
        assert p > 1e-3, (model, i, name, p)
    # deterministic: the same key gives the same draw
    np.testing.assert_array_equal(capi.prior_sample(m, n, 11, 5), draws[5])
    for th in draws[:20]:
        ref = sum(_scipy_prior(name, a, b).logpdf(v) for (name, a, b), v in zip(spec, th))
        assert capi.prior_logdensity(m, th) == pytest.approx(ref, rel=1e-12, abs=1e-12)
    # outside the support
    bad = draws[0].copy()
    bad[0] = -1.0 if spec[0][0] != "gaussian" else bad[0]
    if spec[0][0] != "gaussian":
        assert capi.prior_logdensity(m, bad) == -np.inf


def test_priors_must_exist_and_be_positive_for_the_samplers():
    with pytest.raises(capi.BihipError, match="no prior"):
        capi.prior_sample(capi.generated_model("Lorenz96"), 1, 1, 0)
    with pytest.raises(capi.BihipError, match="not generated"):
        capi.prior_sample(capi.make_model("pz"), 1, 1, 0)
    with pytest.raises(GenError, match="cannot be read"):
        _gen("model M { param a, b\n state x\n sub parameter { a ~ gaussian(b, 1.0) } }")


SAMPLER_CASES = [  # (distribution, expression of the draw, scipy distribution)
    ("gamma", "gen_gamma(2.5, 1.5, w)", lambda st: st.gamma(2.5, scale=1.5)),
    ("gamma_small_shape", "gen_gamma(0.4, 2.0, w)", lambda st: st.gamma(0.4, scale=2.0)),
    ("inverse_gamma", "gen_inverse_gamma(3.0, 2.0, w)", lambda st: st.invgamma(3.0, scale=2.0)),
    ("beta", "gen_beta(2.0, 5.0, w)", lambda st: st.beta(2.0, 5.0)),
    ("poisson_small", "gen_poisson(3.7, w)", lambda st: st.poisson(3.7)),
    ("poisson_large", "gen_poisson(57.3, w)", lambda st: st.poisson(57.3)),
    ("negbin", "gen_negbin(12.0, 3.0, w)", lambda st: st.nbinom(3.0, 3.0 / 15.0)),
    ("negbin_large", "gen_negbin(140.0, 20.0, w)", lambda st: st.nbinom(20.0, 20.0 / 160.0)),
    ("binomial_small", "gen_binomial(20.0, 0.3, w)", lambda st: st.binom(20, 0.3)),
    ("binomial_large", "gen_binomial(400.0, 0.35, w)", lambda st: st.binom(400, 0.35)),
    ("binomial_flip", "gen_binomial(90.0, 0.8, w)", lambda st: st.binom(90, 0.8)),
    # inversion over a count of 1e9: the walk stops at the tail, not at n
    ("binomial_huge_n", "gen_binomial(1.0e9, 3.0e-9, w)", lambda st: st.binom(10**9, 3.0e-9)),
]


def test_rejection_samplers_against_scipy(tmp_path):
    """The draws of gen_samplers.hpp (compiled for the host with stand-in
    random words) follow their distributions: Kolmogorov-Smirnov for the
    continuous ones, chi-square over the support for the counts."""
    from scipy import stats
    n = 20000
    body = ["int main() {"]
    for name, expr, _ in SAMPLER_CASES:
        body.append(f"  for (uint32_t j = 0; j < {n}u; ++j) {{ bihip::DrawStream w(7, j, 1, 2, 0, "
                    f"0x20000000u); std::printf(\"{name} %.17g\\n\", bihip::{expr.replace(', w)', ', w)')}); }}")
    body.append("  return 0;\n}")
    out = _gxx(tmp_path, "\n".join(body), run=True)
    draws = {}
    for line in out.splitlines():
        k, v = line.split()
        draws.setdefault(k, []).append(float(v))
    for name, _, dist in SAMPLER_CASES:
        x = np.array(draws[name])
        d = dist(stats)
        if name.startswith(("poisson", "negbin", "binomial")):
            assert np.all(x == np.floor(x)) and x.min() >= 0
            lo, hi = int(d.ppf(1e-4)), int(d.ppf(1 - 1e-4))
            ks = np.arange(lo, hi + 1)
            obs = np.array([np.sum(x == k) for k in ks], float)
            obs[0] += np.sum(x < lo)
            obs[-1] += np.sum(x > hi)
            exp = d.pmf(ks) * n
            exp[0] += d.cdf(lo - 1) * n
            exp[-1] += d.sf(hi) * n
            keep = exp > 5
            chi = np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep])
            p = stats.chi2.sf(chi, keep.sum() - 1)
        else:
            p = stats.kstest(x, d.cdf).pvalue
        assert p > 1e-4, (name, p)


LANE_OBS_MODEL = """
model LaneObs {
  const N = 40
  dim n(size = N, boundary = 'cyclic')
  dim m(size = N/2)
  param F, obs_sd
  state x[n]
  obs y[m]
  obs z[m]
  obs c[m]
  sub transition(delta = 0.05) {
    ode(alg = 'RK5(4)', h = 0.05, atoler = 1.0e-5, rtoler = 1.0e-5) {
      dx[i]/dt = (x[i+1] - x[i-2])*x[i-1] - x[i] + F
    }
  }
  sub observation {
    y[j] ~ gaussian(x[2*j], obs_sd)
    z[j] ~ gaussian(x[3], obs_sd + 0.01*j)
    c[j] ~ gaussian(F, obs_sd)
  }
}
"""


def test_lane_group_observations_equal_one_lane(tmp_path):
    """A vector observation whose state index does not move with its element
    (z[j] ~ gaussian(x[3], ...)), or that reads no state (c[j] ~ gaussian(F,
    ...)), adds every element's term in the lane group layout: the lanes'
    obs() summed over a group equal the one-lane functor's obs()."""
    import copy
    g = codegen.analyse_text(LANE_OBS_MODEL)
    assert g.lpp == 4
    g1 = copy.deepcopy(g)
    g1.lpp = 1
    g1.name = g.name + "One"
    src = codegen.generate_model(g) + "\n" + codegen.generate_model(g1)
    rng = np.random.default_rng(5)
    x = rng.normal(8.0, 2.0, size=40)
    y = rng.normal(size=60)
    xs = ", ".join(repr(float(v)) for v in x)
    ys = ", ".join(repr(float(v)) for v in y)
    body = f"""
int main() {{
  bihip::ModelArgs ma{{}}; ma.par[0] = 8.0; ma.par[1] = 0.7;
  const double x[40] = {{{xs}}};
  const double y[60] = {{{ys}}};
  bihip::{g.cname}<double> m(ma);
  bihip::{g1.cname}<double> m1(ma);
  double lanes = 0.0;
  for (int l = 0; l < 4; ++l) {{
    double xl[10];
    for (int i = 0; i < 10; ++i) xl[i] = x[10 * l + i];
    lanes += m.obs(xl, y, 10 * l);
  }}
  double one[40];
  for (int i = 0; i < 40; ++i) one[i] = x[i];
  std::printf("%.17g %.17g\\n", lanes, m1.obs(one, y, 0));
  return 0;
}}
"""
    out = _gxx(tmp_path, "namespace bihip {\n" + src + "\n}\n" + body, run=True)
    lanes, one = (float(v) for v in out.split())
    from scipy import stats
    ref = (stats.norm.logpdf(y[:20], x[0:40:2], 0.7).sum() +
           stats.norm.logpdf(y[20:40], x[3], 0.7 + 0.01 * np.arange(20)).sum() +
           stats.norm.logpdf(y[40:60], 8.0, 0.7).sum())
    assert one == pytest.approx(ref, rel=1e-12)
    assert lanes == pytest.approx(ref, rel=1e-12)
