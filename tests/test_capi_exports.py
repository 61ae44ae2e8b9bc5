"""The C-ABI library loads and exports every function include/bihip.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

from libbi_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bihip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|long|double|const char\*)\s+(bihip_\w+)\s*\(", text,
                                 re.M)))


def test_header_declares_expected_set():
    names = declared_functions()
    assert len(names) >= 35
    assert set(names) == set(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_loads():
    L = capi.lib()
    assert L.bihip_last_error() is not None
    m = capi.make_model("l96", d=40)
    assert capi.nobs(m) == 20
    assert L.bihip_model_elem_size(ctypes.byref(capi.make_model("pz", precision=capi.F64))) == 8
    assert L.bihip_model_elem_size(ctypes.byref(capi.make_model("pz", precision=capi.F32))) == 4


def test_host_uniform_matches_oracle(orc):
    L = capi.lib()
    for (seed, idx, step, tag, batch) in [(0, 0, 0, 3, 0), (2 ** 64 - 1, 7, 99, 4, 5),
                                          (123456789, 0, 17, 3, 1)]:
        u = L.bihip_uniform_host(seed, idx, step, tag, batch)
        r = orc.philox_block(seed, idx, step, tag, batch, 0)
        assert u == orc.lib().orc_u53(r[0], r[1])


def test_struct_layouts_match_header():
    # bihip_model: 6 ints, 3 doubles, 16 doubles; bihip_pf_config as declared
    assert ctypes.sizeof(capi.Model) == 6 * 4 + 19 * 8
    assert ctypes.sizeof(capi.PFConfig) == 4 * 4 + 3 * 8 + 8 + 4 + 4 + 4 + 4
