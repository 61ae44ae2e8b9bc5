"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared against.
Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module.  The product (libbi_amd/) never does.

Parity status of each oracle function (how it is pinned) is listed in the
module-level ``PINNED`` table; the tests under tests/ implement each check.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

MODEL_PZ, MODEL_L96, MODEL_LG = 0, 1, 2
RK4, DOPRI5 = 0, 1
SYSTEMATIC, STRATIFIED, MULTINOMIAL = 0, 1, 2
TAG_INIT, TAG_NOISE, TAG_SYSTEMATIC, TAG_STRATIFIED, TAG_MULTINOMIAL, TAG_OBS = 1, 2, 3, 4, 5, 6
TAG_THETA, TAG_PROPOSAL, TAG_ACCEPT, TAG_OUTER = 7, 8, 9, 10
OK, DEGENERATE = 0, 1

# function -> how its parity is pinned (each entry is exercised in tests/)
PINNED = {
    "orc_philox4x32_10": "Random123 known-answer vectors + rocRAND philox4x32_10 fixture",
    "orc_normals": "Box-Muller closed form on fixed words; moments/KS on 1e6 draws",
    "orc_rhs": "closed forms (L96 fixed point x=F, PZ hand values, LG linear)",
    "orc_flow(RK4)": "x'=-x closed form; 4th-order convergence slope",
    "orc_flow(DOPRI5)": "x'=-x closed form within tolerance; scipy RK45 on L96; "
                        "5th-order slope; host-quirk order drop",
    "orc_obs_logdensity": "scipy.stats norm/lognorm logpdf",
    "orc_ess_reduce": "closed forms (equal weights, one-hot), long-double brute force, "
                      "scipy.special.logsumexp, NaN/-inf rules",
    "orc_sumexpu_inclusive_scan": "exact integer weights; numpy cumsum",
    "orc_systematic_cumulative_offspring": "O_{P-1}=P, floor/ceil bounds, exhaustive offset grid",
    "orc_stratified_cumulative_offspring": "O_{P-1}=P, per-stratum count bounds",
    "orc_multinomial_ancestors": "searchsorted(side=right) definition; chi-square frequency test",
    "orc_cumulative_offspring_to_ancestors_permute": "brute force over all offspring vectors P<=6 "
                                                     "against a literal restatement of the "
                                                     "reference loop",
    "orc_permute": "in-place property and multiset preservation, brute force P<=6",
    "orc_gather_rows": "in-place == out-of-place gather",
    "orc_pf_run": "Kalman filter exact logZ on the LG model (essRel 0, 0.5, 1)",
    "orc_pf_step": "equals orc_pf_run step by step (orc_pf_run is built on it)",
    "orc_pf_run(schedule)": "uniform times with every step observed reproduce the plain run "
                            "bit for bit; LG model over irregular times with unobserved steps "
                            "against the scheduled Kalman log-likelihood; bookkeeping (zero "
                            "increments, no resampling after an unobserved step)",
    "orc_pf_run(inputs)": "inputs equal to the model's parameters reproduce the plain run bit "
                          "for bit; LG model with a per-step decay rate against the Kalman "
                          "likelihood with the same per-step transition coefficients",
    "orc_pf_run(imposed)": "the filter's own resampling history imposed reproduces its run bit "
                           "for bit (three resamplers); a history without resampling equals "
                           "the essRel = 0 run; one ancestor row laid out differently changes "
                           "the estimate from that step on and not before",
    "orc_pf_set_state/orc_pf_step_imposed": "restarting from the run's own state after k steps "
                                            "and imposing its resampling continues the run bit "
                                            "for bit",
    "orc_kalman_loglik_lg_sched": "the joint Gaussian log-density of the observed y built from "
                                  "the model's covariances (scipy multivariate_normal); equals "
                                  "orc_kalman_loglik_lg on a uniform schedule",
    "orc_mlp_forward": "torch float64 forward of the same network (GELU z sigmoid(1.702 z))",
    "orc_mlp_loss_grad": "central finite differences of the FP64 loss; bf16 mode: rounding "
                         "points checked against numpy bf16 rounding and torch with the same "
                         "rounded weights",
    "orc_pmmh_run": "first iterations rebuilt from orc_pf_run and the keyed draws; LG model: "
                    "posterior mean of lambda from the chains against Gauss-Hermite "
                    "quadrature of the exact Kalman likelihood",
    "orc_smc2_run": "no moves and no outer resampling: lw_j equals orc_pf_run with theta_j and "
                    "batch j; logZ = logsumexp(lw) - log B; LG model: evidence and posterior "
                    "mean of lambda against Gauss-Hermite quadrature of the Kalman likelihood; "
                    "acceptance decisions against proposal filters run through orc_pf_run",
}


class Model(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int),
        ("d", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("nsteps", ctypes.c_int),
        ("host_quirk", ctypes.c_int),
        ("noise_bits", ctypes.c_int),
        ("h0", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("par", ctypes.c_double * 16),
    ]


class PFConfig(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int),
        ("T", ctypes.c_int),
        ("resampler", ctypes.c_int),
        ("essRel", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("batch", ctypes.c_uint32),
        ("times", ctypes.POINTER(ctypes.c_double)),
        ("observed", ctypes.POINTER(ctypes.c_int)),
        ("inputs", ctypes.POINTER(ctypes.c_double)),
        ("imposed_resampled", ctypes.POINTER(ctypes.c_int)),
        ("imposed_ancestors", ctypes.POINTER(ctypes.c_int)),
    ]


class SMC2Config(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int),
        ("T", ctypes.c_int),
        ("ntheta", ctypes.c_int),
        ("nmoves", ctypes.c_int),
        ("resampler", ctypes.c_int),
        ("essRel_inner", ctypes.c_double),
        ("essRel_outer", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("prior_mean", ctypes.c_double * 8),
        ("prior_sd", ctypes.c_double * 8),
        ("prop_sd", ctypes.c_double * 8),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def _declare(L):
    c = ctypes
    L.orc_philox4x32_10.argtypes = [_U32, _U32, _U32]
    L.orc_philox_block.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                   c.c_uint32, _U32]
    L.orc_u53.argtypes = [c.c_uint32, c.c_uint32]
    L.orc_u53.restype = c.c_double
    L.orc_uniforms.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_int, _D]
    L.orc_normals.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                              c.c_int, c.c_int, _D]
    L.orc_nobs.argtypes = [c.POINTER(Model)]
    L.orc_nobs.restype = c.c_int
    L.orc_init_samples.argtypes = [c.POINTER(Model), c.c_uint64, c.c_uint32, c.c_int, _D, c.c_int,
                                   _D]
    L.orc_transition.argtypes = [c.POINTER(Model), c.c_uint64, c.c_uint32, c.c_uint32, c.c_double,
                                 c.c_double, c.c_int, _D, c.c_int, _I]
    L.orc_obs_logdensity.argtypes = [c.POINTER(Model), _D, c.c_int, _D, c.c_int, _D]
    L.orc_obs_sample.argtypes = [c.POINTER(Model), c.c_uint64, c.c_uint32, c.c_uint32, _D, _D]
    L.orc_flow.argtypes = [c.POINTER(Model), c.c_double, c.c_double, c.c_double, _D]
    L.orc_flow.restype = c.c_int
    L.orc_rhs.argtypes = [c.POINTER(Model), c.c_double, c.c_double, _D, _D]
    for name in ("orc_max_reduce", "orc_logsumexp_reduce"):
        getattr(L, name).argtypes = [_D, c.c_int]
        getattr(L, name).restype = c.c_double
    L.orc_ess_reduce.argtypes = [_D, c.c_int, _D]
    L.orc_ess_reduce.restype = c.c_double
    L.orc_sumexpu_inclusive_scan.argtypes = [_D, c.c_int, _D]
    L.orc_sumexpu_inclusive_scan.restype = c.c_double
    L.orc_systematic_cumulative_offspring.argtypes = [_D, c.c_int, c.c_int, c.c_double, _I]
    L.orc_stratified_cumulative_offspring.argtypes = [_D, c.c_int, c.c_int, _D, _I]
    L.orc_multinomial_ancestors.argtypes = [_D, c.c_int, c.c_int, _D, _I]
    L.orc_ancestors_to_offspring.argtypes = [_I, c.c_int, c.c_int, _I]
    L.orc_offspring_to_cumulative.argtypes = [_I, c.c_int, _I]
    L.orc_cumulative_offspring_to_ancestors.argtypes = [_I, c.c_int, _I]
    L.orc_cumulative_offspring_to_ancestors_permute.argtypes = [_I, c.c_int, _I]
    L.orc_offspring_to_ancestors_permute.argtypes = [_I, c.c_int, _I]
    L.orc_permute.argtypes = [_I, c.c_int]
    L.orc_gather_rows.argtypes = [_I, c.c_int, c.c_int, _D, c.c_int]
    L.orc_pf_run.argtypes = [c.POINTER(Model), c.POINTER(PFConfig), _D, _D, _D, _D, _D, _I, _I, _D]
    L.orc_pmmh_run.argtypes = [c.POINTER(Model), c.POINTER(SMC2Config), c.c_int, _I, _D, _D, _D,
                               _I]
    L.orc_uniform_at.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32, _D]
    L.orc_smc2_run.argtypes = [c.POINTER(Model), c.POINTER(SMC2Config), _I, _D, _D, _D, _D, _I,
                               _D, _I]
    L.orc_bf16.argtypes = [c.c_double]
    L.orc_bf16.restype = c.c_double
    L.orc_mlp_nparams.argtypes = [c.c_int, _I]
    L.orc_mlp_nparams.restype = c.c_size_t
    L.orc_mlp_forward.argtypes = [c.c_int, _I, _D, _D, c.c_int, c.c_int, _D]
    L.orc_mlp_loss_grad.argtypes = [c.c_int, _I, _D, _D, _D, c.c_int, c.c_int, c.c_int, _D]
    L.orc_mlp_loss_grad.restype = c.c_double
    L.orc_pf_new.argtypes = [c.POINTER(Model), c.c_int, c.c_int, c.c_double, c.c_double,
                             c.c_double, c.c_uint64, c.c_uint32]
    L.orc_pf_new.restype = c.c_void_p
    L.orc_pf_free.argtypes = [c.c_void_p]
    L.orc_pf_init.argtypes = [c.c_void_p, _D]
    L.orc_pf_impose.argtypes = [c.c_void_p, _I, _I]
    L.orc_pf_set_state.argtypes = [c.c_void_p, c.c_int, _D, _D, c.c_double, c.c_double]
    L.orc_pf_get_state.argtypes = [c.c_void_p, _D, _D]
    L.orc_pf_step.argtypes = [c.c_void_p, _D, _D, _I, _I]
    L.orc_pf_step_imposed.argtypes = [c.c_void_p, _D, c.c_int, _I, _D]
    L.orc_pf_loglikelihood.argtypes = [c.c_void_p]
    L.orc_pf_loglikelihood.restype = c.c_double
    L.orc_pf_ess.argtypes = [c.c_void_p]
    L.orc_pf_ess.restype = c.c_double
    L.orc_pf_term.argtypes = [c.c_void_p]
    L.orc_pf_term.restype = c.c_double
    L.orc_guided_step.argtypes = [c.c_int, _I, _D, c.c_int, _D, c.c_double, c.c_int, c.c_int,
                                  _D, _D, c.c_uint64, c.c_uint32, c.c_uint32, c.c_double,
                                  c.c_double, c.c_int, _D, _D]
    L.orc_kalman_loglik_lg.argtypes = [c.POINTER(Model), c.c_double, c.c_int, _D]
    L.orc_kalman_loglik_lg.restype = c.c_double
    L.orc_kalman_loglik_lg_sched.argtypes = [c.POINTER(Model), _D, c.POINTER(c.c_int), c.c_int,
                                             _D]
    L.orc_kalman_loglik_lg_sched.restype = c.c_double


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- models
def make_model(kind, d=None, integrator=None, h0=None, atol=1e-5, rtol=1e-5, nsteps=1000,
               host_quirk=0, noise_bits=32, **par):
    """Model configuration; par names follow the model (see oracle.h)."""
    m = Model()
    if kind == "pz":
        m.model, m.d = MODEL_PZ, 2
        m.integrator = RK4 if integrator is None else integrator
        m.h0 = 0.1 if h0 is None else h0
        vals = [par.get("mu", 0.5), par.get("sigma", 0.2), par.get("c", 0.25), par.get("e", 0.3),
                par.get("m_l", 0.1), par.get("m_q", 0.1), par.get("obs_sd", 0.2),
                par.get("P0_logmean", np.log(2.0)), par.get("P0_logsd", 0.2),
                par.get("Z0_logmean", np.log(2.0)), par.get("Z0_logsd", 0.1)]
    elif kind == "l96":
        m.model, m.d = MODEL_L96, 40 if d is None else d
        m.integrator = DOPRI5 if integrator is None else integrator
        m.h0 = 0.05 if h0 is None else h0
        vals = [par.get("F", 8.0), par.get("noise_sd", 0.1), par.get("obs_sd", 1.0),
                par.get("init_sd", 1.0)]
    elif kind == "lg":
        m.model, m.d = MODEL_LG, 1 if d is None else d
        m.integrator = DOPRI5 if integrator is None else integrator
        m.h0 = 0.1 if h0 is None else h0
        vals = [par.get("lam", 0.5), par.get("noise_sd", 0.5), par.get("obs_sd", 0.5),
                par.get("m0", 0.0), par.get("s0", 1.0)]
    else:
        raise ValueError(kind)
    m.nsteps, m.host_quirk, m.noise_bits = nsteps, host_quirk, noise_bits
    m.atol, m.rtol = atol, rtol
    for i, v in enumerate(vals):
        m.par[i] = v
    return m


def nobs(m):
    return lib().orc_nobs(ctypes.byref(m))


# ---------------------------------------------------------------- random
def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def philox_block(seed, index, step, tag, batch, block):
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox_block(seed, index, step, tag, batch, block, o)
    return list(o)


def uniforms(seed, step, tag, n, batch=0):
    u = np.empty(n, np.float64)
    lib().orc_uniforms(seed, step, tag, batch, n, _dp(u))
    return u


def normals(seed, index, step, tag, n, bits=32, batch=0):
    z = np.empty(n, np.float64)
    lib().orc_normals(seed, index, step, tag, batch, n, bits, _dp(z))
    return z


# ---------------------------------------------------------------- model functions
def init_samples(m, seed, P, mean0=None, batch=0):
    X = np.zeros((m.d, P), np.float64)  # column-major P x d == row-major d x P
    mean = _f64(mean0) if mean0 is not None else np.zeros(m.d)
    lib().orc_init_samples(ctypes.byref(m), seed, batch, P, _dp(X), P, _dp(mean))
    return X


def transition(m, seed, step, t1, t2, X, batch=0, return_steps=False):
    X = _f64(X).copy()
    P = X.shape[1]
    ns = np.zeros(P, np.int32)
    lib().orc_transition(ctypes.byref(m), seed, batch, step, t1, t2, P, _dp(X), P, _ip(ns))
    return (X, ns) if return_steps else X


def obs_logdensity(m, y, X, lws):
    X = _f64(X)
    lws = _f64(lws).copy()
    lib().orc_obs_logdensity(ctypes.byref(m), _dp(_f64(y)), X.shape[1], _dp(X), X.shape[1],
                             _dp(lws))
    return lws


def obs_sample(m, seed, step, x, batch=0):
    y = np.empty(nobs(m), np.float64)
    lib().orc_obs_sample(ctypes.byref(m), seed, batch, step, _dp(_f64(x)), _dp(y))
    return y


def flow(m, x, t1, t2, alpha=0.0):
    x = _f64(x).copy()
    n = lib().orc_flow(ctypes.byref(m), alpha, t1, t2, _dp(x))
    return x, n


def rhs(m, x, alpha=0.0, t=0.0):
    x = _f64(x)
    dx = np.empty_like(x)
    lib().orc_rhs(ctypes.byref(m), alpha, t, _dp(x), _dp(dx))
    return dx


def simulate(m, seed, T, dt, x0, t0=0.0):
    """Truth trajectory (T+1 x d) and observations (T x nobs) of one particle."""
    xs = [np.array(x0, np.float64)]
    ys = []
    x = np.array(x0, np.float64).reshape(m.d, 1)
    for k in range(T):
        x = transition(m, seed, k + 1, t0 + k * dt, t0 + (k + 1) * dt, x)
        xs.append(x[:, 0].copy())
        ys.append(obs_sample(m, seed, k + 1, x[:, 0]))
    return np.array(xs), np.array(ys)


# ---------------------------------------------------------------- primitives
def max_reduce(x):
    x = _f64(x)
    return lib().orc_max_reduce(_dp(x), len(x))


def logsumexp_reduce(x):
    x = _f64(x)
    return lib().orc_logsumexp_reduce(_dp(x), len(x))


def ess_reduce(lws):
    lws = _f64(lws)
    lW = ctypes.c_double()
    ess = lib().orc_ess_reduce(_dp(lws), len(lws), ctypes.byref(lW))
    return ess, lW.value


def sumexpu_inclusive_scan(lws):
    lws = _f64(lws)
    Ws = np.empty_like(lws)
    mx = lib().orc_sumexpu_inclusive_scan(_dp(lws), len(lws), _dp(Ws))
    return Ws, mx


def systematic_cumulative_offspring(Ws, a, n=None):
    Ws = _f64(Ws)
    n = len(Ws) if n is None else n
    Os = np.empty(len(Ws), np.int32)
    rc = lib().orc_systematic_cumulative_offspring(_dp(Ws), len(Ws), n, a, _ip(Os))
    return Os, rc


def stratified_cumulative_offspring(Ws, alphas, n=None):
    Ws = _f64(Ws)
    n = len(Ws) if n is None else n
    Os = np.empty(len(Ws), np.int32)
    rc = lib().orc_stratified_cumulative_offspring(_dp(Ws), len(Ws), n, _dp(_f64(alphas)),
                                                   _ip(Os))
    return Os, rc


def multinomial_ancestors(Ws, U):
    Ws, U = _f64(Ws), _f64(U)
    a = np.empty(len(U), np.int32)
    rc = lib().orc_multinomial_ancestors(_dp(Ws), len(Ws), len(U), _dp(U), _ip(a))
    return a, rc


def ancestors_to_offspring(a, P):
    a = _i32(a)
    os_ = np.empty(P, np.int32)
    lib().orc_ancestors_to_offspring(_ip(a), len(a), P, _ip(os_))
    return os_


def offspring_to_cumulative(os_):
    os_ = _i32(os_)
    O = np.empty_like(os_)
    lib().orc_offspring_to_cumulative(_ip(os_), len(os_), _ip(O))
    return O


def cumulative_offspring_to_ancestors(O):
    O = _i32(O)
    a = np.empty(int(O[-1]), np.int32)
    lib().orc_cumulative_offspring_to_ancestors(_ip(O), len(O), _ip(a))
    return a


def cumulative_offspring_to_ancestors_permute(O):
    O = _i32(O)
    a = np.empty(len(O), np.int32)
    lib().orc_cumulative_offspring_to_ancestors_permute(_ip(O), len(O), _ip(a))
    return a


def offspring_to_ancestors_permute(os_):
    os_ = _i32(os_)
    a = np.empty(len(os_), np.int32)
    lib().orc_offspring_to_ancestors_permute(_ip(os_), len(os_), _ip(a))
    return a


def permute(a):
    a = _i32(a).copy()
    lib().orc_permute(_ip(a), len(a))
    return a


def gather_rows(a, X):
    X = _f64(X).copy()
    lib().orc_gather_rows(_ip(_i32(a)), X.shape[1], X.shape[0], _dp(X), X.shape[1])
    return X


# ---------------------------------------------------------------- filter
def pf_run(m, y, P, dt, seed, resampler=SYSTEMATIC, ess_rel=0.5, t0=0.0, mean0=None,
           batch=0, want_ancestors=False, want_state=False, times=None, observed=None,
           inputs=None, imposed_resampled=None, imposed_ancestors=None):
    """times (T + 1) and observed (T flags), both optional, give a schedule in
    place of t0 + k dt with every step observed; inputs (T x 16, optional) the
    model's parameters during each step.  imposed_resampled (T flags) with
    imposed_ancestors (T x P) re-run another filter's resampling history:
    step k resamples iff the flag is set, with those ancestors."""
    y = _f64(y).reshape(len(y), -1)
    T = y.shape[0]
    cfg = PFConfig(P=P, T=T, resampler=resampler, essRel=ess_rel, t0=t0, dt=dt, seed=seed,
                   batch=batch)
    if (imposed_resampled is None) != (imposed_ancestors is None):
        raise ValueError("imposed_resampled and imposed_ancestors go together")
    if imposed_resampled is not None:
        imposed_resampled = _i32(imposed_resampled)
        imposed_ancestors = _i32(imposed_ancestors)
        assert imposed_resampled.shape == (T,) and imposed_ancestors.shape == (T, P)
        cfg.imposed_resampled = _ip(imposed_resampled)
        cfg.imposed_ancestors = _ip(imposed_ancestors)
    if times is not None:
        times = _f64(times)
        assert times.shape == (T + 1,)
        cfg.times = _dp(times)
    if observed is not None:
        observed = np.ascontiguousarray(observed, np.int32)
        assert observed.shape == (T,)
        cfg.observed = _ip(observed)
    if inputs is not None:
        inputs = _f64(inputs)
        assert inputs.shape == (T, 16)
        cfg.inputs = _dp(inputs)
    logZ = ctypes.c_double()
    li = np.zeros(T)
    ess = np.zeros(T)
    res = np.zeros(T, np.int32)
    anc = np.zeros((T, P), np.int32) if want_ancestors else None
    Xo = np.zeros((m.d, P)) if want_state else None
    mean = _f64(mean0) if mean0 is not None else np.zeros(m.d)
    rc = lib().orc_pf_run(ctypes.byref(m), ctypes.byref(cfg), _dp(y), _dp(mean),
                          ctypes.byref(logZ), _dp(li), _dp(ess), _ip(res), _ip(anc), _dp(Xo))
    return dict(rc=rc, logZ=logZ.value, logIncrements=li, ess=ess, resampled=res,
                ancestors=anc, X=Xo)


class StepFilter:
    """The step-wise oracle filter (orc_pf_*): one BootstrapPF stepped from
    Python, whose state can be replaced between steps (set_state) and whose
    resampling can follow another filter's (step with ancestors)."""

    def __init__(self, m, P, dt, seed, resampler=SYSTEMATIC, ess_rel=0.5, t0=0.0, batch=0):
        self.m, self.P = m, P
        self.h = lib().orc_pf_new(ctypes.byref(m), P, resampler, ess_rel, t0, dt, seed, batch)

    def close(self):
        if self.h:
            lib().orc_pf_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, mean0=None):
        mean = _f64(mean0) if mean0 is not None else np.zeros(self.m.d)
        lib().orc_pf_init(self.h, _dp(mean))

    def set_state(self, k, X, lws, log_likelihood, ess):
        """X: d x P (any float type, converted to float64)."""
        X = _f64(X)
        lws = _f64(lws)
        assert X.shape == (self.m.d, self.P) and lws.shape == (self.P,)
        lib().orc_pf_set_state(self.h, k, _dp(X), _dp(lws), log_likelihood, ess)

    def state(self):
        X = np.empty((self.m.d, self.P))
        lws = np.empty(self.P)
        lib().orc_pf_get_state(self.h, _dp(X), _dp(lws))
        return X, lws

    def step(self, y, resampled=None, ancestors=None):
        """One step with observation y; with resampled (0/1) the resampling at
        the step's start is imposed, with the P ancestors when set.  Returns
        (rc, logIncrement)."""
        y = _f64(y).ravel()
        li = ctypes.c_double()
        if resampled is None:
            rc = lib().orc_pf_step(self.h, _dp(y), ctypes.byref(li), None, None)
        else:
            a = _i32(ancestors) if resampled else None
            if resampled:
                assert a.shape == (self.P,)
            rc = lib().orc_pf_step_imposed(self.h, _dp(y), 1 if resampled else 0, _ip(a),
                                           ctypes.byref(li))
        return rc, li.value

    def log_likelihood(self):
        return lib().orc_pf_loglikelihood(self.h)

    def ess(self):
        return lib().orc_pf_ess(self.h)

    def term(self):
        return lib().orc_pf_term(self.h)


def uniform_at(seed, index, step, tag, batch=0):
    u = ctypes.c_double()
    lib().orc_uniform_at(seed, index, step, tag, batch, ctypes.byref(u))
    return u.value


def smc2_config(B, T, prior_mean, prior_sd, prop_sd, dt=1.0, seed=1, nmoves=1,
                resampler=SYSTEMATIC, ess_rel_inner=0.5, ess_rel_outer=0.5, t0=0.0):
    nt = len(prior_mean)
    assert nt <= 8 and len(prior_sd) == nt and len(prop_sd) == nt
    c = SMC2Config(B=B, T=T, ntheta=nt, nmoves=nmoves, resampler=resampler,
                   essRel_inner=ess_rel_inner, essRel_outer=ess_rel_outer, t0=t0, dt=dt,
                   seed=seed)
    for i in range(nt):
        c.prior_mean[i] = prior_mean[i]
        c.prior_sd[i] = prior_sd[i]
        c.prop_sd[i] = prop_sd[i]
    return c


def smc2_run(m, cfg, N, y):
    """SMC^2 (MarginalSIR) over log(theta) with inner filters of N[j] particles."""
    y = _f64(y).reshape(len(y), -1)
    assert y.shape[0] == cfg.T and len(N) == cfg.B
    N = _i32(N)
    logZ = ctypes.c_double()
    theta = np.zeros((cfg.B, cfg.ntheta))
    lw = np.zeros(cfg.B)
    acc = np.zeros(cfg.T + 1, np.int32)
    ess = np.zeros(cfg.T)
    res = np.zeros(cfg.T, np.int32)
    rc = lib().orc_smc2_run(ctypes.byref(m), ctypes.byref(cfg), _ip(N), _dp(y),
                            ctypes.byref(logZ), _dp(theta), _dp(lw), _ip(acc), _dp(ess),
                            _ip(res))
    return dict(rc=rc, logZ=logZ.value, theta=theta, lw=lw, accepts=acc, ess=ess, resampled=res)


def mlp_nparams(dims):
    d = _i32(dims)
    return lib().orc_mlp_nparams(len(d) - 1, _ip(d))


def mlp_forward(dims, params, X, bf16=False):
    d = _i32(dims)
    X = _f64(X).reshape(-1, d[0])
    Y = np.empty((X.shape[0], d[-1]))
    lib().orc_mlp_forward(len(d) - 1, _ip(d), _dp(_f64(params)), _dp(X), X.shape[0],
                          1 if bf16 else 0, _dp(Y))
    return Y


LOSS_MSE, LOSS_GAUSS = 0, 1


def mlp_loss_grad(dims, params, X, T, bf16=False, loss=LOSS_MSE):
    d = _i32(dims)
    X = _f64(X).reshape(-1, d[0])
    T = _f64(T).reshape(-1, d[-1] // 2 if loss == LOSS_GAUSS else d[-1])
    g = np.empty(mlp_nparams(dims))
    val = lib().orc_mlp_loss_grad(len(d) - 1, _ip(d), _dp(_f64(params)), _dp(X), _dp(T),
                                  X.shape[0], loss, 1 if bf16 else 0, _dp(g))
    return val, g


def guided_step(dims, params, norm, F, Phi, y, seed, step, q=0.1, obs_sd=1.0, batch=0,
                bf16=True, noise_bits=24):
    """The learned guided proposal's step on flowed states Phi (d x P):
    returns (X', increments); norm = (x_shift, x_scale, innov_scale,
    par_shift, par_scale)."""
    d_ = _i32(dims)
    Phi = _f64(Phi)
    d, P = Phi.shape
    Xo = np.empty_like(Phi)
    inc = np.empty(P)
    nm = _f64(norm)
    rc = lib().orc_guided_step(len(d_) - 1, _ip(d_), _dp(_f64(params)), 1 if bf16 else 0,
                               _dp(nm), F, d, P, _dp(Phi), _dp(_f64(y)), seed, batch, step, q,
                               obs_sd, noise_bits, _dp(Xo), _dp(inc))
    if rc != OK:
        raise ValueError("guided_step: dims do not fit the state")
    return Xo, inc


def bf16(x):
    """Round to bfloat16 (nearest even), as a float64 array."""
    f = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((f + 0x7FFF + ((f >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)
    return r.astype(np.float64)


def pmmh_run(m, cfg, C, N, y):
    """PMMH (MarginalMH) with cfg.B independent chains of C iterations."""
    y = _f64(y).reshape(len(y), -1)
    assert y.shape[0] == cfg.T and len(N) == cfg.B
    N = _i32(N)
    theta = np.zeros((C, cfg.B, cfg.ntheta))
    ll = np.zeros((C, cfg.B))
    acc = np.zeros(cfg.B, np.int32)
    rc = lib().orc_pmmh_run(ctypes.byref(m), ctypes.byref(cfg), C, _ip(N), _dp(y), _dp(theta),
                            _dp(ll), _ip(acc))
    return dict(rc=rc, theta=theta, ll=ll, accepted=acc)


def kalman_loglik_lg(m, dt, y):
    y = _f64(y).ravel()
    return lib().orc_kalman_loglik_lg(ctypes.byref(m), dt, len(y), _dp(y))


def kalman_loglik_lg_sched(m, times, observed, y):
    """Exact log-likelihood of the LG model over a schedule (times: T + 1,
    observed: T flags; unobserved steps predict only)."""
    y = _f64(y).ravel()
    times = _f64(times)
    obs = np.ascontiguousarray(observed, np.int32)
    return lib().orc_kalman_loglik_lg_sched(ctypes.byref(m), _dp(times), _ip(obs), len(y), _dp(y))
