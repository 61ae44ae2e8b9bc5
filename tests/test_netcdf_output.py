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


def _read_classic(path, with_data=True):
    """A reader of the classic NetCDF grammar written for these tests (CDF-1,
    CDF-2 and CDF-5: counts, lengths, dimension ids and sizes are 32-bit, or
    64-bit in CDF-5; begin offsets 64-bit from CDF-2 on), pinned against
    scipy's reader on CDF-2 files.  Returns the version, dimensions, global
    attributes and variables (dimension names, type code, begin, data)."""
    import struct
    f = open(path, "rb")
    magic = f.read(4)
    assert magic[:3] == b"CDF"
    ver = magic[3]
    i32 = lambda: struct.unpack(">i", f.read(4))[0]
    i64 = lambda: struct.unpack(">q", f.read(8))[0]
    nn = i64 if ver == 5 else i32
    off = i32 if ver == 1 else i64
    size = {2: 1, 4: 4, 5: 4, 6: 8}
    dtype = {4: ">i4", 5: ">f4", 6: ">f8"}

    def name():
        n = nn()
        t = f.read(n)
        f.read(-n % 4)
        return t.decode()

    def atts():
        i32()  # the attribute tag, or zero when absent
        out = {}
        for _ in range(nn()):
            k = name()
            t = i32()
            m = nn()
            raw = f.read(m * size[t])
            f.read(-(m * size[t]) % 4)
            out[k] = raw if t == 2 else np.frombuffer(raw, dtype[t])
        return out

    assert nn() == 0  # numrecs: no record dimension
    i32()
    dims = [(name(), nn()) for _ in range(nn())]
    gatt = atts()
    i32()
    heads = []
    for _ in range(nn()):
        k = name()
        ids = [nn() for _ in range(nn())]
        atts()
        t = i32()
        nn()  # vsize
        heads.append((k, ids, t, off()))
    out = {}
    for k, ids, t, begin in heads:
        shape = tuple(dims[i][1] for i in ids)
        data = None
        if with_data:
            f.seek(begin)
            data = np.frombuffer(f.read(int(np.prod(shape)) * size[t]), dtype[t]).reshape(shape)
        out[k] = (tuple(dims[i][0] for i in ids), t, begin, data)
    f.close()
    return ver, dict(dims), gatt, out


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


def test_cdf5_carries_what_cdf2_does(tmp_path):
    """The test's reader agrees with scipy on a CDF-2 file, and the same
    content forced into CDF-5 (64-bit counts, dimension ids and sizes) reads
    back the same: dimensions, attributes, variables and values."""
    m = capi.make_model("l96", d=6, precision=capi.F32)
    nr, np_, ld = 3, 7, 9
    files = {}
    for fmt in (2, 5):
        path = tmp_path / f"pf{fmt}.nc"
        o = capi.PFOutput(path, m, nr, np_, fmt=fmt)
        _fill(o, m, nr, np_, ld, np.random.default_rng(7), nw=6)
        o.close()
        files[fmt] = _read_classic(path)
    ref = _read(tmp_path / "pf2.nc")
    v2, dims2, att2, vars2 = files[2]
    v5, dims5, att5, vars5 = files[5]
    assert (v2, v5) == (2, 5)
    assert dims2 == ref["dims"] == dims5
    assert att2["libbi_schema"] == ref["attrs"]["libbi_schema"] == att5["libbi_schema"]
    assert set(vars2) == set(ref["vars"]) == set(vars5)
    for k, (dims, dt, data) in ref["vars"].items():
        assert vars2[k][0] == dims == vars5[k][0]
        np.testing.assert_array_equal(vars2[k][3], data)
        np.testing.assert_array_equal(vars5[k][3], data)


def test_variables_beyond_4_gib_take_cdf5(tmp_path):
    """nr = 520 outputs of 2^20 particles: the log-weights (4.4 GB) and the
    noise variable come before the state, so the file is CDF-5 (CDF-2 refuses
    it); rows at both ends read back at their offsets.  The file is sparse."""
    m = capi.make_model("pz", precision=capi.F64)
    nr, np_ = 520, 1 << 20
    with pytest.raises(capi.BihipError, match="4 GiB"):
        capi.PFOutput(tmp_path / "big2.nc", m, nr, np_, fmt=2)
    path = tmp_path / "big.nc"
    o = capi.PFOutput(path, m, nr, np_)
    rng = np.random.default_rng(3)
    rows = {}
    for k in (0, nr - 1):
        X = rng.normal(size=(2, np_))
        a = rng.integers(0, np_, size=np_).astype(np.int32)
        lw = rng.normal(size=np_)
        Z = rng.normal(size=(1, np_))
        o.write_time(k, 0.5 * k)
        o.write_state(k, X, a)
        o.write_log_weights(k, lw)
        o.write_noise(k, Z)
        rows[k] = (X, a, lw, Z)
    o.close()
    ver, dims, _, v = _read_classic(path, with_data=False)
    assert ver == 5 and dims == {"nr": nr, "np": np_}
    assert [k for k in v][-2:] == ["P", "Z"]
    with open(path, "rb") as f:
        def row(name, k, dt, width=1):
            begin = v[name][2]
            f.seek(begin + k * np_ * width * np.dtype(dt).itemsize)
            return np.frombuffer(f.read(np_ * width * np.dtype(dt).itemsize), dt)
        for k, (X, a, lw, Z) in rows.items():
            np.testing.assert_array_equal(row("P", k, ">f8"), X[0])
            np.testing.assert_array_equal(row("Z", k, ">f8"), X[1])
            np.testing.assert_array_equal(row("ancestor", k, ">i4"), a)
            np.testing.assert_array_equal(row("logweight", k, ">f8"), lw)
            np.testing.assert_array_equal(row("alpha", k, ">f8"), Z[0])
            f.seek(v["time"][2] + 8 * k)
            assert np.frombuffer(f.read(8), ">f8")[0] == 0.5 * k
    assert v["logweight"][2] + nr * np_ * 8 <= v["alpha"][2]


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
