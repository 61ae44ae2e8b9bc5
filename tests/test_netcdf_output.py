"""NetCDF output in LibBi's ParticleFilter schema (ParticleFilterNetCDFBuffer,
share/src/bi/netcdf/ParticleFilterNetCDFBuffer.cpp:19-38, over
SimulatorNetCDFBuffer::create, share/src/bi/netcdf/SimulatorNetCDFBuffer.cpp:23-83).

The files are read back with scipy's NetCDF classic reader, which shares no
code with the writer: dimensions, variable shapes and types, the global
attributes and every value written.  CPU only (host arrays)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from libbi_amd import capi


def _read(path):
    f = netcdf_file(str(path), "r", mmap=False)
    out = {"dims": dict(f.dimensions), "attrs": dict(f._attributes),
           "vars": {k: (v.dimensions, v.data.dtype, np.array(v.data)) for k, v in f.variables.items()}}
    f.close()
    return out


def _fill(o, model, nr, np_, ld, rng, nw=0):
    d = model.d
    dt = np.float32 if model.precision == capi.F32 else np.float64
    Xs, As, Ls, ts, Zs = [], [], [], [], []
    for k in range(nr):
        X = rng.normal(size=(d, ld)).astype(dt)
        a = rng.integers(0, np_, size=np_).astype(np.int32)
        lw = rng.normal(size=np_)
        t = 0.25 * k
        o.write_time(k, t)
        o.write_state(k, X, a)
        o.write_log_weights(k, lw)
        if nw:
            Z = rng.normal(size=(nw, ld)).astype(dt)
            o.write_noise(k, Z)
            Zs.append(Z[:, :np_])
        Xs.append(X[:, :np_])
        As.append(a)
        Ls.append(lw)
        ts.append(t)
    o.write_log_likelihood(-123.456)
    o.write_clock(789.0)
    out = np.array(Xs), np.array(As), np.array(Ls), np.array(ts)
    return out + (np.array(Zs),) if nw else out


@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_l96_schema_and_values(tmp_path, prec):
    m = capi.make_model("l96", d=6, precision=prec)
    nr, np_, ld = 4, 9, 12
    path = tmp_path / "pf.nc"
    o = capi.PFOutput(path, m, nr, np_)
    X, A, L, t, Z = _fill(o, m, nr, np_, ld, np.random.default_rng(1), nw=6)
    o.close()
    r = _read(path)
    assert r["dims"] == {"nr": nr, "n": 6, "np": np_}
    assert r["attrs"]["libbi_schema"] == b"ParticleFilter"
    assert int(r["attrs"]["libbi_schema_version"]) == 1
    v = r["vars"]
    assert set(v) == {"time", "x", "w", "ancestor", "logweight", "loglikelihood", "clock"}
    # the noise variable w[n] (R_VAR) takes the state's layout
    assert v["w"][0] == ("nr", "n", "np")
    np.testing.assert_array_equal(v["w"][2], Z)
    # state variables take (nr, model dims, np), in the model's precision
    assert v["x"][0] == ("nr", "n", "np")
    assert v["x"][1] == np.dtype(">f4" if prec == capi.F32 else ">f8")
    np.testing.assert_array_equal(v["x"][2], X)
    assert v["ancestor"][0] == ("nr", "np") and v["ancestor"][1] == np.dtype(">i4")
    np.testing.assert_array_equal(v["ancestor"][2], A)
    assert v["logweight"][0] == ("nr", "np")
    np.testing.assert_array_equal(v["logweight"][2], L)
    np.testing.assert_array_equal(v["time"][2], t)
    assert v["loglikelihood"][0] == () and float(v["loglikelihood"][2]) == -123.456
    assert float(v["clock"][2]) == 789.0


def test_pz_variables_are_p_z_and_alpha(tmp_path):
    m = capi.make_model("pz", precision=capi.F64)
    nr, np_ = 3, 5
    path = tmp_path / "pz.nc"
    o = capi.PFOutput(path, m, nr, np_)
    X, A, L, t, Z = _fill(o, m, nr, np_, np_, np.random.default_rng(2), nw=1)
    o.close()
    r = _read(path)
    assert r["dims"] == {"nr": nr, "np": np_}
    v = r["vars"]
    assert v["P"][0] == ("nr", "np") and v["Z"][0] == ("nr", "np")
    assert v["alpha"][0] == ("nr", "np")
    np.testing.assert_array_equal(v["alpha"][2], Z[:, 0, :])
    np.testing.assert_array_equal(v["P"][2], X[:, 0, :])
    np.testing.assert_array_equal(v["Z"][2], X[:, 1, :])


@pytest.mark.parametrize("name,state,noise", [
    ("PZ", {"P": ("nr", "np"), "Z": ("nr", "np")}, {"alpha": ("nr", "np")}),
    ("Lorenz96d40", {"x": ("nr", "n", "np")}, {"w": ("nr", "n", "np")}),
    ("LinearGaussian", {"x": ("nr", "np")}, {"w": ("nr", "np")})])
def test_generated_model_variables_from_its_file(tmp_path, name, state, noise):
    """A generated model's state and noise variables take the names and
    dimensions of its model file; its noise rows follow the declaration order."""
    m = capi.generated_model(name, precision=capi.F64)
    nw = sum(capi.model_var_sizes(m, capi.VAR_NOISE))
    nr, np_, ld = 3, 5, 8
    path = tmp_path / "gen.nc"
    o = capi.PFOutput(path, m, nr, np_)
    X, A, L, t, Z = _fill(o, m, nr, np_, ld, np.random.default_rng(2), nw=nw)
    o.close()
    v = _read(path)["vars"]
    assert set(v) == set(state) | set(noise) | {"time", "ancestor", "logweight",
                                                "loglikelihood", "clock"}
    for k, dims in {**state, **noise}.items():
        assert v[k][0] == dims
    row = 0
    for k in noise:  # one noise variable per file here
        got = v[k][2].reshape(nr, -1, np_)
        np.testing.assert_array_equal(got, Z[:, row:row + got.shape[1], :])
        row += got.shape[1]
    assert row == nw


def test_outputs_written_out_of_order(tmp_path):
    """Outputs are written at their index, in any order (the buffer is random access)."""
    m = capi.make_model("lg", precision=capi.F64)
    nr, np_ = 5, 3
    path = tmp_path / "lg.nc"
    o = capi.PFOutput(path, m, nr, np_)
    for k in (4, 0, 2, 1, 3):
        o.write_time(k, float(k))
        o.write_state(k, np.full((1, np_), float(k)), np.arange(np_) + k)
        o.write_log_weights(k, np.full(np_, -float(k)))
    o.close()
    v = _read(path)["vars"]
    np.testing.assert_array_equal(v["time"][2], np.arange(nr, dtype=float))
    np.testing.assert_array_equal(v["x"][2], np.arange(nr, dtype=float)[:, None].repeat(np_, 1))
    np.testing.assert_array_equal(v["ancestor"][2], np.arange(np_)[None, :] + np.arange(nr)[:, None])


def test_bad_arguments_are_rejected(tmp_path):
    m = capi.make_model("lg")
    o = capi.PFOutput(tmp_path / "x.nc", m, 2, 2)
    with pytest.raises(capi.BihipError):
        o.write_time(2, 0.0)
    with pytest.raises(capi.BihipError):
        o.write_state(0, np.zeros((1, 1)), np.zeros(2))  # ld < np
    o.close()
    with pytest.raises(capi.BihipError):
        capi.PFOutput(tmp_path / "no_such_dir" / "x.nc", m, 2, 2)
    with pytest.raises(capi.BihipError):
        capi.PFOutput(tmp_path / "y.nc", m, 0, 2)
