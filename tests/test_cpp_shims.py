"""The C++ shims (include/bi/amd/) drive a LibBi-style filter client through
the C ABI; the client's results match the oracle's filter bit for bit in the
ancestors, and the device-resident path gives the same results."""
import os
import subprocess

import numpy as np
import pytest

from gpu_helpers import l96_spun_up

HERE = os.path.dirname(os.path.abspath(__file__))
CLIENT = os.path.join(HERE, "cpp", "_build", "shim_filter")


def test_client_built_and_prints_usage():
    assert os.path.exists(CLIENT), "run __graft_entry__.build() first"
    r = subprocess.run([CLIENT], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def _run(tmp_path, model, res, P, dt, seed, ess_rel, y, mean0=None, fused=False, extra=()):
    obs = tmp_path / "obs.bin"
    out = tmp_path / "out.bin"
    y = np.asarray(y, np.float64)
    y.tofile(obs)
    args = [CLIENT, model, str(res), str(P), repr(dt), str(seed), repr(ess_rel), str(obs), str(out)]
    if mean0 is not None:
        m0 = tmp_path / "mean0.bin"
        np.asarray(mean0, np.float64).tofile(m0)
        args.append(str(m0))
    if fused:
        args.append("fused")
    args.extend(extra)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    T = y.shape[0]
    raw = out.read_bytes()
    head = np.frombuffer(raw[: 8 * (T + 1)], np.float64)
    anc = np.frombuffer(raw[8 * (T + 1):], np.int32).reshape(T + 1, P)
    return head[0], head[1:], anc


@pytest.mark.gpu
@pytest.mark.parametrize("res", [0, 1, 2])
@pytest.mark.parametrize("fused", [False, True])
def test_shim_pz_filter_matches_oracle(tmp_path, orc, torch_gpu, res, fused):
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, 100, 1)[:, 0]
    _, y = orc.simulate(m, 100, 40, 1.0, x0)
    P = 2048
    logZ, li, anc = _run(tmp_path, "pz", res, P, 1.0, 42, 0.8, y, fused=fused)
    ref = orc.pf_run(m, y, P, 1.0, 42, resampler=res, ess_rel=0.8, want_ancestors=True)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-10)
    np.testing.assert_allclose(li, ref["logIncrements"], rtol=1e-9, atol=1e-9)
    np.testing.assert_array_equal(anc[0], np.arange(P))
    np.testing.assert_array_equal(anc[1:], ref["ancestors"])


@pytest.mark.gpu
def test_shim_l96_filter_matches_oracle(tmp_path, orc, torch_gpu):
    om = orc.make_model("l96", d=40)
    mean0 = l96_spun_up(orc)
    _, y = orc.simulate(om, 7, 10, 0.05, mean0)
    P = 1024
    logZ, li, anc = _run(tmp_path, "l96", 0, P, 0.05, 9, 0.5, y, mean0=mean0)
    ref = orc.pf_run(om, y, P, 0.05, 9, ess_rel=0.5, mean0=mean0, want_ancestors=True)
    assert logZ == pytest.approx(ref["logZ"], rel=1e-9)
    np.testing.assert_array_equal(anc[1:], ref["ancestors"])


@pytest.mark.gpu
def test_shim_netcdf_outputs_agree(tmp_path, orc, torch_gpu):
    """ParticleFilterNetCDFBuffer behind the shim filter and the device filter's
    recorded history write the same ParticleFilter NetCDF file: times, states,
    ancestors (as the client's own output), log-weights and logZ."""
    from scipy.io import netcdf_file
    m = orc.make_model("pz")
    x0 = orc.init_samples(m, 100, 1)[:, 0]
    _, y = orc.simulate(m, 100, 12, 1.0, x0)
    P, T = 512, 12
    files = {}
    for fused in (False, True):
        nc = tmp_path / f"out_{int(fused)}.nc"
        sub = tmp_path / f"run_{int(fused)}"
        sub.mkdir()
        logZ, li, anc = _run(sub, "pz", 0, P, 1.0, 7, 0.6, y, fused=fused, extra=[f"nc={nc}"])
        f = netcdf_file(str(nc), "r", mmap=False)
        assert f.dimensions == {"nr": T + 1, "np": P}
        v = {k: np.array(x.data) for k, x in f.variables.items()}
        f.close()
        np.testing.assert_array_equal(v["ancestor"], anc)
        assert float(v["loglikelihood"]) == logZ
        np.testing.assert_array_equal(v["time"], np.arange(T + 1, dtype=float))
        files[fused] = v
    a, b = files[False], files[True]
    np.testing.assert_array_equal(a["ancestor"], b["ancestor"])
    for name in ("P", "Z", "logweight"):
        np.testing.assert_allclose(a[name], b[name], rtol=1e-12, atol=1e-12)
    assert float(a["loglikelihood"]) == pytest.approx(float(b["loglikelihood"]), rel=1e-12)


ODE_CLIENT = os.path.join(HERE, "cpp", "_build", "shim_ode")


def test_ode_client_built_and_prints_usage():
    """The generated ode block's call form, bi_ode_set(h, atoler, rtoler) then
    bi::{RK4,DOPRI5}Integrator<Model, action_typelist>::update(t1, t2, s)
    (share/tt/cpp/block/ode.hpp.tt:40-52), compiles against the shims."""
    assert os.path.exists(ODE_CLIENT), "run __graft_entry__.build() first"
    r = subprocess.run([ODE_CLIENT], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("alg,model,t2,h,tol", [("dopri5", "l96", 0.05, 0.05, 1e-5),
                                                ("rk4", "pz", 1.0, 0.1, 1e-5)])
def test_shim_integrator_update_matches_flow(tmp_path, orc, ctx, alg, model, t2, h, tol):
    """The static update(t1, t2, s) of the shims runs the ode block with the
    constants bi_ode_set gave: it equals bihip_flow of the same model and the
    oracle's flow (FP64, 1e-10), and the constants reach the integrator (a
    DOPRI5 update at atol = rtol = tol differs from one at 1e-3)."""
    import ctypes
    from libbi_amd import capi
    from gpu_helpers import dev, host
    P = 777
    rng = np.random.default_rng(3)
    if model == "l96":
        om = orc.make_model("l96", d=40, noise_sd=0.0, atol=tol, rtol=tol, h0=h)
        X0 = l96_spun_up(orc)[:, None] + rng.normal(size=(40, P))
        gm = capi.make_model("l96", d=40, precision=capi.F64, noise_sd=0.0, atol=tol, rtol=tol,
                             h0=h)
    else:
        om = orc.make_model("pz", h0=h, sigma=0.0)
        X0 = np.exp(rng.normal(np.log(2.0), 0.2, size=(2, P)))
        gm = capi.make_model("pz", precision=capi.F64, h0=h)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    X0.astype(np.float64).tofile(inp)

    def client(atol, rtol):
        r = subprocess.run([ODE_CLIENT, alg, model, str(P), "0.0", repr(t2), repr(h), repr(atol),
                            repr(rtol), str(inp), str(out)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        return np.fromfile(out, np.float64).reshape(X0.shape)

    got = client(tol, tol)
    X = dev(X0)
    capi.check(capi.lib().bihip_flow(ctx.h, ctypes.byref(gm), 0.0, t2, P,
                                     ctypes.c_void_p(X.data_ptr()), P, None, None))
    np.testing.assert_array_equal(got, host(X))
    ref = orc.transition(om, 1, 1, 0.0, t2, X0)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10)
    if alg == "dopri5":
        loose = client(1e-3, 1e-3)
        assert np.abs(loose - got).max() > 1e-6

