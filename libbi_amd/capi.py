"""ctypes binding of the C ABI in include/bihip.h (libbi_amd/lib/libbihip.so).

Device buffers are torch tensors on a ROCm device (torch is used for device
memory and streams only); every call passes plain pointers.  The library is
loaded from the package tree; a missing library raises immediately -- there is
no fallback implementation.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbihip.so")

OK, DEGENERATE, EINVAL, EHIP, ESTATE, EIO = 0, 1, 2, 3, 4, 5
MODEL_PZ, MODEL_L96, MODEL_LG = 0, 1, 2
MODEL_GENERATED = 64  # first id of the models generated from libbi_amd/bi/*.bi
VAR_PARAM, VAR_INPUT, VAR_STATE, VAR_OBS, VAR_NOISE = 0, 1, 2, 3, 4
RK4, DOPRI5 = 0, 1
SYSTEMATIC, STRATIFIED, MULTINOMIAL = 0, 1, 2
F32, F64 = 0, 1
LOSS_MSE, LOSS_GAUSS = 0, 1
TAG_INIT, TAG_NOISE, TAG_SYSTEMATIC, TAG_STRATIFIED, TAG_MULTINOMIAL, TAG_OBS = 1, 2, 3, 4, 5, 6
TAG_THETA, TAG_PROPOSAL, TAG_ACCEPT, TAG_OUTER, TAG_PATH = 7, 8, 9, 10, 13
TIMING_ALL, TIMING_TRANSITION = 1, 2
NOALLOC_MAX_P = 1 << 20  # BIHIP_NOALLOC_MAX_P: largest P of the allocation-free permutation

# exported entry points (must match include/bihip.h)
EXPORTS = [
    "bihip_ctx_create", "bihip_ctx_destroy", "bihip_ctx_set_stream", "bihip_ctx_synchronize",
    "bihip_last_error", "bihip_model_nobs", "bihip_model_elem_size",
    "bihip_malloc", "bihip_free", "bihip_memcpy_to_device", "bihip_memcpy_to_host",
    "bihip_memcpy_device", "bihip_fill", "bihip_iota", "bihip_uniform_host",
    "bihip_philox", "bihip_uniforms", "bihip_normals",
    "bihip_init_samples", "bihip_transition", "bihip_flow", "bihip_obs_logdensity",
    "bihip_obs_sample",
    "bihip_ess_reduce", "bihip_logsumexp_reduce", "bihip_max_reduce",
    "bihip_sumexpu_inclusive_scan",
    "bihip_systematic_cumulative_offspring", "bihip_stratified_cumulative_offspring",
    "bihip_multinomial_ancestors", "bihip_ancestors_to_offspring",
    "bihip_offspring_to_cumulative", "bihip_cumulative_offspring_to_ancestors",
    "bihip_cumulative_offspring_to_ancestors_permute", "bihip_offspring_to_ancestors_permute",
    "bihip_cumulative_offspring_to_ancestors_permute_noalloc",
    "bihip_gather_rows",
    "bihip_pf_create", "bihip_pf_destroy", "bihip_pf_set_observations", "bihip_pf_init",
    "bihip_pf_step", "bihip_pf_term", "bihip_pf_status", "bihip_pf_outputs",
    "bihip_pf_ancestors", "bihip_pf_state", "bihip_pf_position", "bihip_pf_set_timing",
    "bihip_pf_kernel_time", "bihip_pf_impose_resampling",
    "bihip_batch_create", "bihip_batch_destroy", "bihip_batch_set_observations",
    "bihip_batch_init", "bihip_batch_step", "bihip_batch_outputs", "bihip_batch_ancestors",
    "bihip_batch_tuning", "bihip_batch_tuning_table", "bihip_batch_set_timing",
    "bihip_batch_kernel_time",
    "bihip_batch_particles", "bihip_batch_set_theta", "bihip_batch_set_sizes",
    "bihip_batch_set_stream_base", "bihip_batch_increments", "bihip_batch_resample_filters",
    "bihip_batch_accept",
    "bihip_smc2_create", "bihip_smc2_destroy", "bihip_smc2_set_observations", "bihip_smc2_run",
    "bihip_smc2_outputs",
    "bihip_mlp_create", "bihip_mlp_destroy", "bihip_mlp_nparams", "bihip_mlp_set_params",
    "bihip_mlp_get_params", "bihip_mlp_forward", "bihip_mlp_loss_grad", "bihip_mlp_train",
    "bihip_surrogate_flow", "bihip_pf_set_surrogate", "bihip_mlp_set_loss",
    "bihip_guided_propose", "bihip_pf_set_guided",
    "bihip_pmmh_create", "bihip_pmmh_destroy", "bihip_pmmh_set_observations", "bihip_pmmh_run",
    "bihip_pmmh_outputs", "bihip_pf_paths", "bihip_pf_sample_path",
    "bihip_pfout_create", "bihip_pfout_write_time", "bihip_pfout_write_state",
    "bihip_pfout_write_log_weights", "bihip_pfout_write_log_likelihood", "bihip_pfout_write_noise",
    "bihip_pfout_write_clock", "bihip_pfout_close", "bihip_pf_write_netcdf",
    "bihip_pf_cache_nodes",
    "bihip_pf_set_schedule", "bihip_pf_set_inputs",
    "bihip_model_generated", "bihip_model_var_name", "bihip_model_name", "bihip_model_var_size",
    "bihip_model_prior_sample", "bihip_model_prior_logdensity",
    "bihip_pfbank_create", "bihip_pfbank_destroy", "bihip_pfbank_set_observations",
    "bihip_pfbank_init", "bihip_pfbank_step", "bihip_pfbank_outputs", "bihip_pfbank_ancestors",
    "bihip_pfbank_state", "bihip_pfbank_position", "bihip_pfbank_set_surrogate",
    "bihip_pfbank_set_guided", "bihip_pfbank_set_timing", "bihip_pfbank_kernel_time",
    "bihip_pfbank_set_theta", "bihip_pfbank_set_stream_base", "bihip_pfbank_increments",
    "bihip_pfbank_resample_filters", "bihip_pfbank_accept", "bihip_smc2_create_bank",
    "bihip_smc2_set_surrogate", "bihip_pmmh_create_bank", "bihip_smc2_set_guided",
    "bihip_pmmh_set_surrogate", "bihip_pmmh_set_guided", "bihip_pfout_create_format",
]


class BihipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bihip error {code}: {msg}")
        self.code = code


class ParticleFilterDegeneratedException(BihipError):
    """Raised for BIHIP_DEGENERATE, as bi::ParticleFilterDegeneratedException
    (share/src/bi/misc/exception.hpp:47-49)."""


class Model(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("d", ctypes.c_int), ("integrator", ctypes.c_int),
        ("nsteps", ctypes.c_int), ("host_quirk", ctypes.c_int), ("precision", ctypes.c_int),
        ("h0", ctypes.c_double), ("atol", ctypes.c_double), ("rtol", ctypes.c_double),
        ("par", ctypes.c_double * 16),
    ]


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int), ("T", ctypes.c_int), ("resampler", ctypes.c_int),
        ("essRel", ctypes.c_double), ("t0", ctypes.c_double), ("dt", ctypes.c_double),
        ("seed", ctypes.c_uint64), ("record", ctypes.c_int), ("use_graph", ctypes.c_int),
        ("block", ctypes.c_int), ("capacity", ctypes.c_long), ("stream_base", ctypes.c_uint32),
        ("items", ctypes.c_int), ("oversub", ctypes.c_int),
    ]


class SMC2Config(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int), ("T", ctypes.c_int), ("ntheta", ctypes.c_int),
        ("nmoves", ctypes.c_int), ("resampler", ctypes.c_int),
        ("essRel_inner", ctypes.c_double), ("essRel_outer", ctypes.c_double),
        ("t0", ctypes.c_double), ("dt", ctypes.c_double), ("seed", ctypes.c_uint64),
        ("prior_mean", ctypes.c_double * 8), ("prior_sd", ctypes.c_double * 8),
        ("prop_sd", ctypes.c_double * 8), ("use_graph", ctypes.c_int), ("block", ctypes.c_int),
        ("prior_from_model", ctypes.c_int),
    ]


class SurrogateNorm(ctypes.Structure):
    _fields_ = [
        ("x_shift", ctypes.c_float), ("x_scale", ctypes.c_float),
        ("par_shift", ctypes.c_double), ("par_scale", ctypes.c_double),
        ("dx_scale", ctypes.c_float),
    ]


class GuidedNorm(ctypes.Structure):
    _fields_ = [
        ("x_shift", ctypes.c_float), ("x_scale", ctypes.c_float),
        ("innov_scale", ctypes.c_double), ("par_shift", ctypes.c_double),
        ("par_scale", ctypes.c_double),
    ]


class PFConfig(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int), ("T", ctypes.c_int), ("resampler", ctypes.c_int),
        ("essRel", ctypes.c_double), ("t0", ctypes.c_double), ("dt", ctypes.c_double),
        ("seed", ctypes.c_uint64), ("batch", ctypes.c_uint32), ("record", ctypes.c_int),
        ("use_graph", ctypes.c_int),
    ]


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with libbi_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    c = ctypes
    P = c.c_void_p
    i, u32, u64, d = c.c_int, c.c_uint32, c.c_uint64, c.c_double
    pd, pi, pl = c.POINTER(c.c_double), c.POINTER(c.c_int), c.POINTER(c.c_long)
    pf = c.POINTER(c.c_float)
    pm, pc = c.POINTER(Model), c.POINTER(PFConfig)
    sig = {
        "bihip_ctx_create": [i, c.POINTER(P)],
        "bihip_ctx_destroy": [P],
        "bihip_ctx_set_stream": [P, P],
        "bihip_ctx_synchronize": [P],
        "bihip_model_nobs": [pm],
        "bihip_model_elem_size": [pm],
        "bihip_malloc": [P, c.c_size_t, c.POINTER(P)],
        "bihip_free": [P, P],
        "bihip_memcpy_to_device": [P, P, P, c.c_size_t],
        "bihip_memcpy_to_host": [P, P, P, c.c_size_t],
        "bihip_memcpy_device": [P, P, P, c.c_size_t],
        "bihip_fill": [P, P, i, d],
        "bihip_iota": [P, P, i],
        "bihip_flow": [P, pm, d, d, i, P, i, P, P],
        "bihip_philox": [P, P, u32, u32, i, P],
        "bihip_uniforms": [P, u64, u32, u32, u32, i, P],
        "bihip_normals": [P, u64, u32, u32, u32, u32, i, i, P],
        "bihip_init_samples": [P, pm, u64, u32, i, P, i, pd],
        "bihip_transition": [P, pm, u64, u32, u32, d, d, i, P, P, i, P, P],
        "bihip_obs_logdensity": [P, pm, P, i, P, i, P],
        "bihip_obs_sample": [P, pm, u64, u32, u32, i, P, i, P],
        "bihip_ess_reduce": [P, P, i, pd, pd],
        "bihip_logsumexp_reduce": [P, P, i, pd],
        "bihip_max_reduce": [P, P, i, pd],
        "bihip_sumexpu_inclusive_scan": [P, P, i, P, pd],
        "bihip_systematic_cumulative_offspring": [P, P, i, i, d, P],
        "bihip_stratified_cumulative_offspring": [P, P, i, i, P, P],
        "bihip_multinomial_ancestors": [P, P, i, i, P, P],
        "bihip_ancestors_to_offspring": [P, P, i, i, P],
        "bihip_offspring_to_cumulative": [P, P, i, P],
        "bihip_cumulative_offspring_to_ancestors": [P, P, i, P],
        "bihip_cumulative_offspring_to_ancestors_permute": [P, P, i, P],
        "bihip_offspring_to_ancestors_permute": [P, P, i, P],
        "bihip_cumulative_offspring_to_ancestors_permute_noalloc": [P, P, i, P],
        "bihip_pf_cache_nodes": [P, pl, pl],
        "bihip_gather_rows": [P, P, i, i, P, i, i],
        "bihip_pf_create": [P, pm, pc, c.POINTER(P)],
        "bihip_pf_destroy": [P],
        "bihip_pf_set_observations": [P, pd],
        "bihip_pf_init": [P, pd],
        "bihip_pf_step": [P, i],
        "bihip_pf_term": [P, pd],
        "bihip_pf_status": [P],
        "bihip_pf_outputs": [P, pd, pd, pi, pd],
        "bihip_pf_ancestors": [P, i, pi],
        "bihip_pf_state": [P, c.POINTER(P), pi, c.POINTER(P)],
        "bihip_pf_position": [P],
        "bihip_pf_set_timing": [P, i],
        "bihip_pf_kernel_time": [P, i, pd, pl],
        "bihip_batch_create": [P, pm, c.POINTER(BatchConfig), pi, pd, i, c.POINTER(P)],
        "bihip_batch_destroy": [P],
        "bihip_batch_set_observations": [P, pd],
        "bihip_batch_init": [P],
        "bihip_batch_step": [P, i],
        "bihip_batch_outputs": [P, pd, pd, pd, pi, pi],
        "bihip_batch_ancestors": [P, i, i, pi],
        "bihip_batch_tuning": [P, pi, pi, pd],
        "bihip_batch_tuning_table": [P, pi, pi, pi, pi, pd],
        "bihip_batch_set_timing": [P, i],
        "bihip_batch_kernel_time": [P, pd, pl],
        "bihip_batch_set_theta": [P, pd],
        "bihip_batch_set_sizes": [P, pi],
        "bihip_batch_set_stream_base": [P, u32],
        "bihip_batch_increments": [P, i, pd],
        "bihip_batch_resample_filters": [P, pi],
        "bihip_batch_accept": [P, P, pi],
        "bihip_smc2_create": [P, pm, c.POINTER(SMC2Config), pi, c.POINTER(P)],
        "bihip_smc2_destroy": [P],
        "bihip_smc2_set_observations": [P, pd],
        "bihip_smc2_run": [P],
        "bihip_smc2_outputs": [P, pd, pd, pd, pi, pd, pi],
        "bihip_mlp_create": [P, i, pi, u64, c.POINTER(P)],
        "bihip_mlp_destroy": [P],
        "bihip_mlp_set_params": [P, pf],
        "bihip_mlp_get_params": [P, pf],
        "bihip_mlp_forward": [P, P, c.c_long, c.c_long, P, c.c_long],
        "bihip_mlp_loss_grad": [P, P, c.c_long, P, c.c_long, i, pd, pf],
        "bihip_mlp_train": [P, P, c.c_long, P, c.c_long, c.c_long, i, i, c.c_float, u64, pf],
        "bihip_surrogate_flow": [P, P, c.POINTER(SurrogateNorm), d, P, P, i, i],
        "bihip_pf_set_surrogate": [P, P, c.POINTER(SurrogateNorm)],
        "bihip_mlp_set_loss": [P, i],
        "bihip_guided_propose": [P, P, c.POINTER(GuidedNorm), d, P, i, i, P, u64, u32, u32, d, d,
                                 P],
        "bihip_pf_set_guided": [P, P, c.POINTER(GuidedNorm)],
        "bihip_pmmh_create": [P, pm, c.POINTER(SMC2Config), i, pi, c.POINTER(P)],
        "bihip_pmmh_destroy": [P],
        "bihip_pmmh_set_observations": [P, pd],
        "bihip_pmmh_run": [P],
        "bihip_pmmh_outputs": [P, pd, pd, pi],
        "bihip_pf_paths": [P, pi, i, pd],
        "bihip_pf_sample_path": [P, pd, pi],
        "bihip_pfout_create": [c.c_char_p, pm, i, i, c.POINTER(P)],
        "bihip_pfout_write_time": [P, i, d],
        "bihip_pfout_write_state": [P, i, P, i, pi],
        "bihip_pfout_write_log_weights": [P, i, pd],
        "bihip_pfout_write_noise": [P, i, P, i],
        "bihip_pfout_write_log_likelihood": [P, d],
        "bihip_pfout_write_clock": [P, d],
        "bihip_pfout_close": [P],
        "bihip_pf_write_netcdf": [P, c.c_char_p],
        "bihip_pf_set_schedule": [P, pd, pi],
        "bihip_pf_set_inputs": [P, pd],
        "bihip_pf_impose_resampling": [P, pi, pi],
        "bihip_pfbank_create": [P, pm, pc, i, pd, i, c.POINTER(P)],
        "bihip_pfbank_destroy": [P],
        "bihip_pfbank_set_observations": [P, pd],
        "bihip_pfbank_init": [P, pd],
        "bihip_pfbank_step": [P, i],
        "bihip_pfbank_outputs": [P, pd, pd, pd, pi, pi],
        "bihip_pfbank_ancestors": [P, i, i, pi],
        "bihip_pfbank_state": [P, i, c.POINTER(P), pi, c.POINTER(P)],
        "bihip_pfbank_position": [P],
        "bihip_pfbank_set_surrogate": [P, P, c.POINTER(SurrogateNorm)],
        "bihip_pfbank_set_guided": [P, P, c.POINTER(GuidedNorm)],
        "bihip_pfbank_set_timing": [P, i],
        "bihip_pfbank_kernel_time": [P, i, pd, pl],
        "bihip_pfbank_set_theta": [P, pd],
        "bihip_pfbank_set_stream_base": [P, u32],
        "bihip_pfbank_increments": [P, i, pd],
        "bihip_pfbank_resample_filters": [P, pi],
        "bihip_pfbank_accept": [P, P, pi],
        "bihip_smc2_create_bank": [P, pm, c.POINTER(SMC2Config), i, pd, c.POINTER(P)],
        "bihip_smc2_set_surrogate": [P, P, c.POINTER(SurrogateNorm)],
        "bihip_pmmh_create_bank": [P, pm, c.POINTER(SMC2Config), i, i, pd, c.POINTER(P)],
        "bihip_smc2_set_guided": [P, P, c.POINTER(GuidedNorm)],
        "bihip_pmmh_set_surrogate": [P, P, c.POINTER(SurrogateNorm)],
        "bihip_pmmh_set_guided": [P, P, c.POINTER(GuidedNorm)],
        "bihip_pfout_create_format": [c.c_char_p, pm, i, i, i, c.POINTER(P)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = c.c_int
    L.bihip_last_error.argtypes = []
    L.bihip_last_error.restype = c.c_char_p
    L.bihip_model_generated.argtypes = [c.c_char_p, c.c_int, P]
    L.bihip_model_var_name.argtypes = [P, c.c_int, c.c_int]
    L.bihip_model_var_name.restype = c.c_char_p
    L.bihip_model_name.argtypes = [P]
    L.bihip_model_name.restype = c.c_char_p
    L.bihip_model_var_size.argtypes = [P, c.c_int, c.c_int]
    L.bihip_model_prior_sample.argtypes = [P, c.c_uint64, c.c_uint32, c.c_int, P]
    L.bihip_model_prior_logdensity.argtypes = [P, c.c_int, P, P]
    L.bihip_uniform_host.argtypes = [u64, u32, u32, u32, u32]
    L.bihip_uniform_host.restype = c.c_double
    L.bihip_batch_particles.argtypes = [P]
    L.bihip_batch_particles.restype = c.c_long
    L.bihip_mlp_nparams.argtypes = [P]
    L.bihip_mlp_nparams.restype = c.c_long


def check(rc):
    if rc != OK:
        msg = lib().bihip_last_error().decode(errors="replace")
        if rc == DEGENERATE:
            raise ParticleFilterDegeneratedException(rc, msg)
        raise BihipError(rc, msg)
    return rc


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------- models
# parameter names -> index in bihip_model::par, per model (make_model's order)
PAR_INDEX = {
    MODEL_PZ: {n: i for i, n in enumerate(["mu", "sigma", "c", "e", "m_l", "m_q", "obs_sd",
                                           "P0_logmean", "P0_logsd", "Z0_logmean", "Z0_logsd"])},
    MODEL_L96: {n: i for i, n in enumerate(["F", "noise_sd", "obs_sd", "init_sd"])},
    MODEL_LG: {n: i for i, n in enumerate(["lam", "noise_sd", "obs_sd", "m0", "s0"])},
}


def make_model(kind, d=None, integrator=None, precision=F64, h0=None, atol=1e-5, rtol=1e-5,
               nsteps=1000, host_quirk=0, **par):
    """Model description with the same parameter names as the oracle's make_model."""
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
    m.precision, m.nsteps, m.host_quirk = precision, nsteps, host_quirk
    m.atol, m.rtol = atol, rtol
    for k, v in enumerate(vals):
        m.par[k] = v
    return m


def model_vars(m, kind):
    """Variable names of one kind (VAR_*) of a generated model ([] otherwise)."""
    out = []
    while True:
        n = lib().bihip_model_var_name(ctypes.byref(m), kind, len(out))
        if n is None:
            return out
        out.append(n.decode())


def model_var_sizes(m, kind):
    """Element counts of the variables of one kind of a generated model."""
    return [lib().bihip_model_var_size(ctypes.byref(m), kind, i)
            for i in range(len(model_vars(m, kind)))]


def par_index(m):
    """Parameter name -> (first) slot of bihip_model::par for a model description."""
    if m.model >= MODEL_GENERATED:
        out, slot = {}, 0
        for kind in (VAR_PARAM, VAR_INPUT):
            for n, k in zip(model_vars(m, kind), model_var_sizes(m, kind)):
                out[n] = slot
                slot += k
        return out
    return PAR_INDEX[m.model]


def generated_model(name, precision=F64, integrator=None, h0=None, atol=None, rtol=None,
                    nsteps=None, **par):
    """Description of a model compiled from a LibBi model file in
    libbi_amd/bi (the `model` name of the file); named values set its params
    and inputs, every other slot is zero.  The ode block's settings are the
    defaults of integrator, h0, atol and rtol."""
    m = Model()
    check(lib().bihip_model_generated(os.fsencode(name), precision, ctypes.byref(m)))
    idx = par_index(m)
    sizes = dict(zip(model_vars(m, VAR_PARAM) + model_vars(m, VAR_INPUT),
                     model_var_sizes(m, VAR_PARAM) + model_var_sizes(m, VAR_INPUT)))
    for k, v in par.items():
        if k not in idx:
            raise KeyError(f"{name} has no param or input {k!r} (has {sorted(idx)})")
        vals = np.atleast_1d(np.asarray(v, np.float64))
        if vals.size != sizes[k]:
            raise ValueError(f"{k} has {sizes[k]} elements, {vals.size} given")
        for j, x in enumerate(vals):
            m.par[idx[k] + j] = x
    if integrator is not None:
        m.integrator = integrator
    if h0 is not None:
        m.h0 = h0
    if atol is not None:
        m.atol = atol
    if rtol is not None:
        m.rtol = rtol
    if nsteps is not None:
        m.nsteps = nsteps
    return m


def prior_sample(m, n, seed, index):
    """Draw `index` of the first n param elements from the model file's prior."""
    th = np.zeros(n)
    check(lib().bihip_model_prior_sample(ctypes.byref(m), seed, index, n,
                                         th.ctypes.data_as(ctypes.c_void_p)))
    return th


def prior_logdensity(m, theta):
    th = np.ascontiguousarray(theta, np.float64)
    lp = ctypes.c_double()
    check(lib().bihip_model_prior_logdensity(ctypes.byref(m), len(th),
                                             th.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.byref(lp)))
    return lp.value


def nobs(m):
    return lib().bihip_model_nobs(ctypes.byref(m))


class Context:
    """bihip_ctx bound to one device and to torch's current stream."""

    def __init__(self, device=0, stream=None):
        import torch
        self.device = device
        self._children = weakref.WeakSet()  # objects that must be destroyed before the context
        h = ctypes.c_void_p()
        check(lib().bihip_ctx_create(device, ctypes.byref(h)))
        self.h = h
        s = stream if stream is not None else torch.cuda.current_stream(device)
        self.stream = s
        check(lib().bihip_ctx_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)))

    def adopt(self, obj):
        self._children.add(obj)
        return obj

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            lib().bihip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().bihip_ctx_synchronize(self.h))


class ParticleFilter:
    """bihip_pf: device-resident bootstrap particle filter."""

    def __init__(self, ctx, model, P, T, dt, seed, resampler=SYSTEMATIC, ess_rel=0.5, t0=0.0,
                 batch=0, record=False, use_graph=False):
        """record: False, True (ancestors), "paths" (ancestors and states at
        every time) or "tree" (the pruned ancestry tree: paths without the
        full history)."""
        self.ctx, self.model = ctx, model
        rec = {"paths": 2, "tree": 3}.get(record, 1 if record else 0)
        self.cfg = PFConfig(P=P, T=T, resampler=resampler, essRel=ess_rel, t0=t0, dt=dt,
                            seed=seed, batch=batch, record=rec,
                            use_graph=1 if use_graph else 0)
        h = ctypes.c_void_p()
        check(lib().bihip_pf_create(ctx.h, ctypes.byref(model), ctypes.byref(self.cfg),
                                    ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.P, self.T, self.nobs = P, T, nobs(model)

    def close(self):
        if self.h:
            lib().bihip_pf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.T, self.nobs)
        check(lib().bihip_pf_set_observations(self.h, y.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def init(self, mean0=None):
        m = None
        if mean0 is not None:
            self._mean0 = np.ascontiguousarray(mean0, dtype=np.float64)
            m = self._mean0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(lib().bihip_pf_init(self.h, m))

    def step(self, n=1):
        check(lib().bihip_pf_step(self.h, n))

    def term(self):
        z = ctypes.c_double()
        check(lib().bihip_pf_term(self.h, ctypes.byref(z)))
        return z.value

    def status(self):
        return check(lib().bihip_pf_status(self.h))

    def outputs(self):
        li = np.zeros(self.T)
        ess = np.zeros(self.T)
        res = np.zeros(self.T, np.int32)
        ll = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        check(lib().bihip_pf_outputs(self.h, li.ctypes.data_as(D), ess.ctypes.data_as(D),
                                     res.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     ctypes.byref(ll)))
        return dict(logIncrements=li, ess=ess, resampled=res, logLikelihood=ll.value)

    def ancestors(self, k=-1):
        a = np.zeros(self.P, np.int32)
        check(lib().bihip_pf_ancestors(self.h, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return a

    def impose_resampling(self, resampled, ancestors):
        """Follow another filter's resampling history: step k resamples iff
        resampled[k], with ancestors[k] (T x P); None removes it."""
        if resampled is None:
            check(lib().bihip_pf_impose_resampling(self.h, None, None))
            return
        r = np.ascontiguousarray(resampled, np.int32)
        a = np.ascontiguousarray(ancestors, np.int32)
        if r.shape != (self.T,) or a.shape != (self.T, self.P):
            raise ValueError(f"need {self.T} flags and {self.T} x {self.P} ancestors")
        self._imposed = (r, a)
        I = ctypes.POINTER(ctypes.c_int)
        check(lib().bihip_pf_impose_resampling(self.h, r.ctypes.data_as(I), a.ctypes.data_as(I)))

    def state_ptr(self):
        X = ctypes.c_void_p()
        ld = ctypes.c_int()
        lws = ctypes.c_void_p()
        check(lib().bihip_pf_state(self.h, ctypes.byref(X), ctypes.byref(ld), ctypes.byref(lws)))
        return X.value, ld.value, lws.value

    def position(self):
        return lib().bihip_pf_position(self.h)

    def paths(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        K = self.position()
        out = np.zeros((len(idx), K + 1, self.model.d))
        check(lib().bihip_pf_paths(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(idx), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def cache_nodes(self):
        """(live nodes, node capacity) of the ancestry tree (record="tree")."""
        live, cap = ctypes.c_long(), ctypes.c_long()
        check(lib().bihip_pf_cache_nodes(self.h, ctypes.byref(live), ctypes.byref(cap)))
        return live.value, cap.value

    def sample_path(self):
        K = self.position()
        out = np.zeros((K + 1, self.model.d))
        p = ctypes.c_int()
        check(lib().bihip_pf_sample_path(self.h, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), ctypes.byref(p)))
        return p.value, out

    def state(self):
        """Device copies of the current state (d x ld torch tensor), ld, and the
        log-weights (P doubles)."""
        import torch
        Xp, ld, lwp = self.state_ptr()
        dt = torch.float32 if self.model.precision == F32 else torch.float64
        X = torch.empty((self.model.d, ld), dtype=dt, device="cuda")
        lws = torch.empty(self.P, dtype=torch.float64, device="cuda")
        check(lib().bihip_memcpy_device(self.ctx.h, X.data_ptr(), Xp, X.numel() * X.element_size()))
        check(lib().bihip_memcpy_device(self.ctx.h, lws.data_ptr(), lwp, 8 * self.P))
        self.ctx.sync()
        return X, ld, lws

    def set_surrogate(self, mlp, norm):
        """Use the learned flow (capi.MLP, SurrogateNorm) in place of the integrator;
        mlp=None restores the integrator."""
        check(lib().bihip_pf_set_surrogate(self.h, mlp.h if mlp is not None else None,
                                           ctypes.byref(norm) if norm is not None else None))

    def set_guided(self, mlp, norm):
        """Use the learned guided proposal (capi.MLP, GuidedNorm); mlp=None restores
        the bootstrap step."""
        check(lib().bihip_pf_set_guided(self.h, mlp.h if mlp is not None else None,
                                        ctypes.byref(norm) if norm is not None else None))

    def set_schedule(self, times, observed=None):
        """Step k runs from times[k] to times[k+1]; observed[k] == 0 skips its
        correction (and the resampling after it)."""
        t = np.ascontiguousarray(times, np.float64)
        if t.shape != (self.T + 1,):
            raise ValueError(f"need {self.T + 1} times")
        o = None
        if observed is not None:
            o = np.ascontiguousarray(observed, np.int32)
            if o.shape != (self.T,):
                raise ValueError(f"need {self.T} observed flags")
        check(lib().bihip_pf_set_schedule(
            self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if o is not None else None))

    def set_inputs(self, inputs):
        """The model's 16 parameters for each step (T x 16), or None to remove
        them; see inputs_from()."""
        if inputs is None:
            check(lib().bihip_pf_set_inputs(self.h, None))
            return
        v = np.ascontiguousarray(inputs, np.float64)
        if v.shape != (self.T, 16):
            raise ValueError(f"need a {self.T} x 16 array of parameters")
        check(lib().bihip_pf_set_inputs(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def inputs_from(self, **series):
        """T x 16 parameters: the model's, with named parameters replaced by
        per-step series (e.g. F=array of T values for the Lorenz-96 forcing)."""
        out = np.tile(np.array(self.model.par[:16], np.float64), (self.T, 1))
        for name, vals in series.items():
            out[:, par_index(self.model)[name]] = vals
        return out

    def write_netcdf(self, path):
        """The recorded history (record="paths") in LibBi's ParticleFilter NetCDF schema."""
        check(lib().bihip_pf_write_netcdf(self.h, os.fsencode(path)))

    def set_timing(self, on, transition_only=False):
        """HIP events around the step's phases (kernel_time), or around the
        transition and proposal kernels only."""
        mode = (TIMING_TRANSITION if transition_only else TIMING_ALL) if on else 0
        check(lib().bihip_pf_set_timing(self.h, mode))

    def kernel_time(self, kind):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        check(lib().bihip_pf_kernel_time(self.h, kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class FilterBank:
    """bihip_pfbank: B bootstrap filters of P particles with parameters theta_j
    (B x ntheta, column i replacing par[i]), advanced by shared launches;
    filter j equals ParticleFilter(batch=batch + j) with those parameters."""

    def __init__(self, ctx, model, P, T, dt, seed, theta, resampler=SYSTEMATIC, ess_rel=0.5,
                 t0=0.0, batch=0, record=False, use_graph=False):
        self.ctx, self.model = ctx, model
        th = np.ascontiguousarray(theta, np.float64)
        if th.ndim == 1:
            th = th[:, None]
        self.B, self.ntheta = th.shape
        self.theta = th
        self.cfg = PFConfig(P=P, T=T, resampler=resampler, essRel=ess_rel, t0=t0, dt=dt,
                            seed=seed, batch=batch, record=1 if record else 0,
                            use_graph=1 if use_graph else 0)
        h = ctypes.c_void_p()
        check(lib().bihip_pfbank_create(ctx.h, ctypes.byref(model), ctypes.byref(self.cfg),
                                        self.B, th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        self.ntheta, ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.P, self.T, self.nobs = P, T, nobs(model)

    def close(self):
        if self.h:
            lib().bihip_pfbank_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.T, self.nobs)
        check(lib().bihip_pfbank_set_observations(self.h, y.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def init(self, mean0=None):
        m = None
        if mean0 is not None:
            self._mean0 = np.ascontiguousarray(mean0, dtype=np.float64)
            m = self._mean0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(lib().bihip_pfbank_init(self.h, m))

    def step(self, n=1):
        check(lib().bihip_pfbank_step(self.h, n))

    def position(self):
        return lib().bihip_pfbank_position(self.h)

    def outputs(self):
        B, T = self.B, self.T
        ll, li, ess = np.zeros(B), np.zeros((B, T)), np.zeros((B, T))
        res, st = np.zeros((B, T), np.int32), np.zeros(B, np.int32)
        D, I = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        check(lib().bihip_pfbank_outputs(self.h, ll.ctypes.data_as(D), li.ctypes.data_as(D),
                                         ess.ctypes.data_as(D), res.ctypes.data_as(I),
                                         st.ctypes.data_as(I)))
        return dict(logLikelihood=ll, logIncrements=li, ess=ess, resampled=res, status=st)

    def ancestors(self, j, k):
        a = np.zeros(self.P, np.int32)
        check(lib().bihip_pfbank_ancestors(self.h, j, k,
                                           a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return a

    def state(self, j):
        """Device copies of filter j's state (d x ld torch tensor), ld and
        log-weights (P doubles)."""
        import torch
        X, ld, lws = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_void_p()
        check(lib().bihip_pfbank_state(self.h, j, ctypes.byref(X), ctypes.byref(ld),
                                       ctypes.byref(lws)))
        dt = torch.float32 if self.model.precision == F32 else torch.float64
        Xt = torch.empty((self.model.d, ld.value), dtype=dt, device="cuda")
        lw = torch.empty(self.P, dtype=torch.float64, device="cuda")
        check(lib().bihip_memcpy_device(self.ctx.h, Xt.data_ptr(), X.value,
                                        Xt.numel() * Xt.element_size()))
        check(lib().bihip_memcpy_device(self.ctx.h, lw.data_ptr(), lws.value, 8 * self.P))
        self.ctx.sync()
        return Xt, ld.value, lw

    def set_surrogate(self, mlp, norm):
        check(lib().bihip_pfbank_set_surrogate(self.h, mlp.h if mlp is not None else None,
                                               ctypes.byref(norm) if norm is not None else None))

    def set_guided(self, mlp, norm):
        check(lib().bihip_pfbank_set_guided(self.h, mlp.h if mlp is not None else None,
                                            ctypes.byref(norm) if norm is not None else None))

    def set_theta(self, theta):
        th = np.ascontiguousarray(theta, np.float64).reshape(self.B, self.ntheta)
        self.theta = th
        check(lib().bihip_pfbank_set_theta(self.h, th.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def set_stream_base(self, base):
        check(lib().bihip_pfbank_set_stream_base(self.h, base))

    def increments(self, k):
        li = np.zeros(self.B)
        check(lib().bihip_pfbank_increments(self.h, k, li.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))
        return li

    def resample_filters(self, anc):
        a = np.ascontiguousarray(anc, np.int32)
        check(lib().bihip_pfbank_resample_filters(self.h, a.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))

    def accept(self, src, acc):
        a = np.ascontiguousarray(acc, np.int32)
        check(lib().bihip_pfbank_accept(self.h, src.h, a.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))

    def set_timing(self, on, transition_only=False):
        mode = (TIMING_TRANSITION if transition_only else TIMING_ALL) if on else 0
        check(lib().bihip_pfbank_set_timing(self.h, mode))

    def kernel_time(self, kind):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        check(lib().bihip_pfbank_kernel_time(self.h, kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class BatchFilter:
    """bihip_batch: B inner bootstrap filters with their own N_j and theta_j."""

    def __init__(self, ctx, model, N, theta, T, dt, seed, resampler=SYSTEMATIC, ess_rel=0.5,
                 t0=0.0, record=False, use_graph=False, block=0, capacity=0, stream_base=0,
                 items=0, oversub=0):
        self.ctx, self.model = ctx, model
        self.N = np.ascontiguousarray(N, dtype=np.int32)
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)
        B = len(self.N)
        ntheta = self.theta.shape[1] if self.theta.ndim == 2 else 0
        self.cfg = BatchConfig(B=B, T=T, resampler=resampler, essRel=ess_rel, t0=t0, dt=dt,
                               seed=seed, record=1 if record else 0,
                               use_graph=1 if use_graph else 0, block=block, capacity=capacity,
                               stream_base=stream_base, items=items, oversub=oversub)
        h = ctypes.c_void_p()
        check(lib().bihip_batch_create(
            ctx.h, ctypes.byref(model), ctypes.byref(self.cfg),
            self.N.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            self.theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ntheta, ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.B, self.T, self.nobs = B, T, nobs(model)

    def close(self):
        if self.h:
            lib().bihip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.T, self.nobs)
        check(lib().bihip_batch_set_observations(self.h, y.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def init(self):
        check(lib().bihip_batch_init(self.h))

    def step(self, n=1):
        check(lib().bihip_batch_step(self.h, n))

    def outputs(self):
        D, I = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        ll = np.zeros(self.B)
        li = np.zeros((self.B, self.T))
        ess = np.zeros((self.B, self.T))
        res = np.zeros((self.B, self.T), np.int32)
        st = np.zeros(self.B, np.int32)
        check(lib().bihip_batch_outputs(self.h, ll.ctypes.data_as(D), li.ctypes.data_as(D),
                                        ess.ctypes.data_as(D), res.ctypes.data_as(I),
                                        st.ctypes.data_as(I)))
        return dict(logLikelihood=ll, logIncrements=li, ess=ess, resampled=res, status=st)

    def ancestors(self, j, k):
        a = np.zeros(int(self.N[j]), np.int32)
        check(lib().bihip_batch_ancestors(self.h, j, k, a.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))
        return a

    def tuning(self):
        blk, grid = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(3)
        check(lib().bihip_batch_tuning(self.h, ctypes.byref(blk), ctypes.byref(grid),
                                       ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return blk.value, grid.value, ms

    def tuning_table(self):
        """(items, oversub) chosen, and every timed configuration as a list of
        ((block, items, oversub), ms)."""
        it, ov, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(64)
        cfg = np.zeros((64, 3), np.int32)
        ms = np.zeros(64)
        check(lib().bihip_batch_tuning_table(
            self.h, ctypes.byref(it), ctypes.byref(ov), ctypes.byref(n),
            cfg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        m = min(n.value, 64)
        return it.value, ov.value, [(tuple(int(v) for v in cfg[i]), float(ms[i])) for i in range(m)]

    def set_timing(self, on):
        check(lib().bihip_batch_set_timing(self.h, 1 if on else 0))

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        check(lib().bihip_batch_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def particles(self):
        return lib().bihip_batch_particles(self.h)

    def set_theta(self, theta):
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)
        check(lib().bihip_batch_set_theta(self.h, self.theta.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def set_sizes(self, N):
        N = np.ascontiguousarray(N, dtype=np.int32)
        check(lib().bihip_batch_set_sizes(self.h, N.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        self.N = N

    def set_stream_base(self, base):
        check(lib().bihip_batch_set_stream_base(self.h, base))

    def increments(self, k):
        li = np.zeros(self.B)
        check(lib().bihip_batch_increments(self.h, k, li.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))
        return li

    def resample_filters(self, anc):
        anc = np.ascontiguousarray(anc, dtype=np.int32)
        check(lib().bihip_batch_resample_filters(self.h, anc.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))
        self.N = self.N[anc]
        if self.theta.ndim == 2:
            self.theta = self.theta[anc]

    def accept(self, src, mask):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        check(lib().bihip_batch_accept(self.h, src.h, mask.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int))))
        if self.theta.ndim == 2:
            self.theta = np.where(mask[:, None] != 0, src.theta, self.theta)


class SMC2:
    """bihip_smc2: SMC^2 (MarginalSIR) over log theta with batched inner filters."""

    def __init__(self, ctx, model, N, T, prior_mean, prior_sd, prop_sd, dt=1.0, seed=1,
                 nmoves=1, resampler=SYSTEMATIC, ess_rel_inner=0.5, ess_rel_outer=0.5, t0=0.0,
                 use_graph=False, block=0, bank=None, mean0=None):
        """prior_mean = prior_sd = None: the prior of theta (the first len(prop_sd)
        params) is the parameter block of the generated model's file.  bank=(B, P):
        inner filters as filter banks of P particles (N unused; mean0 for Lorenz-96)."""
        from_model = prior_mean is None
        nt = len(prop_sd)
        if bank is not None:
            B, Pb = bank
            self.N = np.full(B, Pb, np.int32)
        else:
            self.N = np.ascontiguousarray(N, dtype=np.int32)
        B = len(self.N)
        cfg = SMC2Config(B=B, T=T, ntheta=nt, nmoves=nmoves, resampler=resampler,
                         essRel_inner=ess_rel_inner, essRel_outer=ess_rel_outer, t0=t0, dt=dt,
                         seed=seed, use_graph=1 if use_graph else 0, block=block,
                         prior_from_model=1 if from_model else 0)
        for i in range(nt):
            if not from_model:
                cfg.prior_mean[i] = prior_mean[i]
                cfg.prior_sd[i] = prior_sd[i]
            cfg.prop_sd[i] = prop_sd[i]
        self.cfg = cfg
        h = ctypes.c_void_p()
        if bank is not None:
            self._mean0 = (np.ascontiguousarray(mean0, np.float64) if mean0 is not None
                           else None)
            m0 = (self._mean0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                  if self._mean0 is not None else None)
            check(lib().bihip_smc2_create_bank(ctx.h, ctypes.byref(model), ctypes.byref(cfg),
                                               int(bank[1]), m0, ctypes.byref(h)))
        else:
            check(lib().bihip_smc2_create(ctx.h, ctypes.byref(model), ctypes.byref(cfg),
                                          self.N.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.B, self.T, self.nt, self.nobs = B, T, nt, nobs(model)

    def close(self):
        if self.h:
            lib().bihip_smc2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.T, self.nobs)
        check(lib().bihip_smc2_set_observations(self.h, y.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def run(self):
        check(lib().bihip_smc2_run(self.h))

    def set_surrogate(self, mlp, norm):
        """The learned surrogate flow in every inner filter (bank=... only)."""
        check(lib().bihip_smc2_set_surrogate(self.h, mlp.h if mlp is not None else None,
                                             ctypes.byref(norm) if norm is not None else None))

    def set_guided(self, mlp, norm):
        """The learned proposal in every inner filter (bank=... only)."""
        check(lib().bihip_smc2_set_guided(self.h, mlp.h if mlp is not None else None,
                                          ctypes.byref(norm) if norm is not None else None))

    def outputs(self):
        D, I = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        logZ = ctypes.c_double()
        theta = np.zeros((self.B, self.nt))
        lw = np.zeros(self.B)
        acc = np.zeros(self.T + 1, np.int32)
        ess = np.zeros(self.T)
        res = np.zeros(self.T, np.int32)
        check(lib().bihip_smc2_outputs(self.h, ctypes.byref(logZ), theta.ctypes.data_as(D),
                                       lw.ctypes.data_as(D), acc.ctypes.data_as(I),
                                       ess.ctypes.data_as(D), res.ctypes.data_as(I)))
        return dict(logZ=logZ.value, theta=theta, lw=lw, accepts=acc, ess=ess, resampled=res)


class PMMH:
    """bihip_pmmh: MarginalMH with B independent chains, proposals filtered as one batch
    (or, with bank=(B, P), as one filter bank of B filters of P particles)."""

    def __init__(self, ctx, model, N, T, C, prior_mean, prior_sd, prop_sd, dt=1.0, seed=1,
                 resampler=SYSTEMATIC, ess_rel=0.5, t0=0.0, use_graph=False, block=0,
                 bank=None, mean0=None):
        """prior_mean = prior_sd = None: the model file's parameter block is the prior.
        bank=(B, P): the chains' filters run as a filter bank (N is ignored); Lorenz-96
        filters start around mean0."""
        from_model = prior_mean is None
        nt = len(prop_sd)
        if bank is not None:
            self.N = np.full(bank[0], bank[1], np.int32)
        else:
            self.N = np.ascontiguousarray(N, dtype=np.int32)
        B = len(self.N)
        cfg = SMC2Config(B=B, T=T, ntheta=nt, nmoves=0, resampler=resampler,
                         essRel_inner=ess_rel, essRel_outer=0.0, t0=t0, dt=dt, seed=seed,
                         use_graph=1 if use_graph else 0, block=block,
                         prior_from_model=1 if from_model else 0)
        for k in range(nt):
            if not from_model:
                cfg.prior_mean[k] = prior_mean[k]
                cfg.prior_sd[k] = prior_sd[k]
            cfg.prop_sd[k] = prop_sd[k]
        self.cfg = cfg
        h = ctypes.c_void_p()
        if bank is not None:
            self._mean0 = (np.ascontiguousarray(mean0, np.float64) if mean0 is not None
                           else None)
            m0 = (self._mean0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                  if self._mean0 is not None else None)
            check(lib().bihip_pmmh_create_bank(ctx.h, ctypes.byref(model), ctypes.byref(cfg), C,
                                               int(bank[1]), m0, ctypes.byref(h)))
        else:
            check(lib().bihip_pmmh_create(ctx.h, ctypes.byref(model), ctypes.byref(cfg), C,
                                          self.N.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.B, self.T, self.C, self.nt, self.nobs = B, T, C, nt, nobs(model)

    def close(self):
        if self.h:
            lib().bihip_pmmh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.T, self.nobs)
        check(lib().bihip_pmmh_set_observations(self.h, y.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))

    def run(self):
        check(lib().bihip_pmmh_run(self.h))

    def set_surrogate(self, mlp, norm):
        """The learned surrogate flow in the chains' filters (bank=... only)."""
        check(lib().bihip_pmmh_set_surrogate(self.h, mlp.h if mlp is not None else None,
                                             ctypes.byref(norm) if norm is not None else None))

    def set_guided(self, mlp, norm):
        """The learned proposal in the chains' filters (bank=... only)."""
        check(lib().bihip_pmmh_set_guided(self.h, mlp.h if mlp is not None else None,
                                          ctypes.byref(norm) if norm is not None else None))

    def outputs(self):
        D, I = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        theta = np.zeros((self.C, self.B, self.nt))
        ll = np.zeros((self.C, self.B))
        acc = np.zeros(self.B, np.int32)
        check(lib().bihip_pmmh_outputs(self.h, theta.ctypes.data_as(D), ll.ctypes.data_as(D),
                                       acc.ctypes.data_as(I)))
        return dict(theta=theta, ll=ll, accepted=acc)


class MLP:
    """bihip_mlp: BF16 MFMA multilayer perceptron (GELU z sigmoid(1.702 z) on
    hidden layers, linear output).  Device arrays are torch float32 tensors."""

    def __init__(self, ctx, dims, seed=1):
        self.ctx = ctx
        self.dims = [int(d) for d in dims]
        d = (ctypes.c_int * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        check(lib().bihip_mlp_create(ctx.h, len(self.dims) - 1, d, seed, ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if self.h:
            lib().bihip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nparams(self):
        return lib().bihip_mlp_nparams(self.h)

    def set_loss(self, kind):
        check(lib().bihip_mlp_set_loss(self.h, kind))

    def set_params(self, flat):
        f = np.ascontiguousarray(flat, dtype=np.float32)
        assert f.size == self.nparams()
        check(lib().bihip_mlp_set_params(self.h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_params(self):
        f = np.zeros(self.nparams(), np.float32)
        check(lib().bihip_mlp_get_params(self.h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return f

    def forward(self, X):
        import torch
        assert X.dtype == torch.float32 and X.is_cuda and X.dim() == 2 and X.is_contiguous()
        Y = torch.empty((X.shape[0], self.dims[-1]), dtype=torch.float32, device=X.device)
        check(lib().bihip_mlp_forward(self.h, X.data_ptr(), X.shape[1], X.shape[0],
                                      Y.data_ptr(), Y.shape[1]))
        return Y

    def loss_grad(self, X, T):
        assert X.is_contiguous() and T.is_contiguous()
        loss = ctypes.c_double()
        g = np.zeros(self.nparams(), np.float32)
        check(lib().bihip_mlp_loss_grad(self.h, X.data_ptr(), X.shape[1], T.data_ptr(),
                                        T.shape[1], X.shape[0], ctypes.byref(loss),
                                        g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return loss.value, g

    def train(self, X, T, batch, steps, lr, seed=1):
        assert X.is_contiguous() and T.is_contiguous()
        losses = np.zeros(max(steps, 1), np.float32)
        check(lib().bihip_mlp_train(self.h, X.data_ptr(), X.shape[1], T.data_ptr(), T.shape[1],
                                    X.shape[0], batch, steps, lr, seed,
                                    losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return losses[:steps]

    def surrogate_flow(self, norm, par, X):
        """Out-of-place surrogate flow of X (torch float32, d x P column-major, i.e. shape (d, ld))."""
        import torch
        assert X.dtype == torch.float32 and X.is_contiguous()
        Y = torch.empty_like(X)
        check(lib().bihip_surrogate_flow(self.ctx.h, self.h, ctypes.byref(norm), par, X.data_ptr(),
                                         Y.data_ptr(), X.shape[1], X.shape[1]))
        return Y

    def guided_propose(self, norm, par, Phi, y, seed, step, noise_sd=0.1, obs_sd=1.0, batch=0):
        """In-place proposal over Phi (torch float32, shape (200, ld)); returns the
        log-weight increments (torch float64, P)."""
        import torch
        assert Phi.dtype == torch.float32 and Phi.is_contiguous()
        P = Phi.shape[1]
        yd = torch.as_tensor(np.ascontiguousarray(y, np.float64)).cuda()
        inc = torch.empty(P, dtype=torch.float64, device=Phi.device)
        check(lib().bihip_guided_propose(self.ctx.h, self.h, ctypes.byref(norm), par,
                                         Phi.data_ptr(), P, P, yd.data_ptr(), seed, batch, step,
                                         noise_sd, obs_sd, inc.data_ptr()))
        self.ctx.sync()
        return inc


class PFOutput:
    """bihip_pfout: LibBi's ParticleFilter NetCDF schema written from host arrays
    (ParticleFilterNetCDFBuffer's writeTime / writeState / writeLogWeights /
    writeLogLikelihood / writeClock)."""

    def __init__(self, path, model, nr, np_, fmt=0):
        """fmt: 0 (CDF-2, or CDF-5 when a variable before the last exceeds 4 GiB), 2 or 5."""
        self.h = ctypes.c_void_p()
        self.model, self.nr, self.np = model, nr, np_
        check(lib().bihip_pfout_create_format(os.fsencode(path), ctypes.byref(model), nr, np_,
                                              fmt, ctypes.byref(self.h)))

    def write_time(self, k, t):
        check(lib().bihip_pfout_write_time(self.h, k, float(t)))

    def write_state(self, k, X, ancestors):
        """X: d x ld (ld >= np), float32 or float64 by the model's precision."""
        dt = np.float32 if self.model.precision == F32 else np.float64
        X = np.ascontiguousarray(X, dt)
        a = np.ascontiguousarray(ancestors, np.int32)
        check(lib().bihip_pfout_write_state(self.h, k, X.ctypes.data_as(ctypes.c_void_p),
                                            X.shape[-1],
                                            a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def write_noise(self, k, Z):
        """Z: noise_width x ld (ld >= np) noise variables, by the model's precision."""
        dt = np.float32 if self.model.precision == F32 else np.float64
        Z = np.ascontiguousarray(Z, dt)
        check(lib().bihip_pfout_write_noise(self.h, k, Z.ctypes.data_as(ctypes.c_void_p),
                                            Z.shape[-1]))

    def write_log_weights(self, k, lws):
        lw = np.ascontiguousarray(lws, np.float64)
        check(lib().bihip_pfout_write_log_weights(
            self.h, k, lw.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def write_log_likelihood(self, ll):
        check(lib().bihip_pfout_write_log_likelihood(self.h, float(ll)))

    def write_clock(self, us):
        check(lib().bihip_pfout_write_clock(self.h, float(us)))

    def close(self):
        if self.h:
            check(lib().bihip_pfout_close(self.h))
            self.h = ctypes.c_void_p()
