"""CPU-side checks of the boundary: libculamp loads and exports every symbol
declared in include/culamp.h, argument structs match the header, settings
JSON parsing matches EOTArgs, and errors are reported without a GPU."""
import ctypes as C
import json
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "culamp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(culamp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import culamp_amd
    from culamp_amd import _lib
    lib = culamp_amd.load()
    syms = header_symbols()
    assert len(syms) >= 29
    for s in syms:
        assert hasattr(lib, s), s
    # every declared symbol has a ctypes signature in the binding
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_abi_version_and_defaults():
    import culamp_amd
    from culamp_amd._lib import culamp_args
    lib = culamp_amd.load()
    assert lib.culamp_abi_version() == 1
    a = culamp_args()
    lib.culamp_args_default(C.byref(a))
    # EOTArgs defaults, src/Utilities.jl:18-31
    assert (a.eta_p, a.eta_mu, a.tau_p, a.tau_mu, a.alpha, a.B, a.epsilon, a.anneal_mult) == \
        (0.0, 0.0, 1.0, 1.0, 0.01, 1.0, 1e-4, 0.95)
    assert (a.itermax, a.inner_iter, a.verbose, a.precision, a.stop) == (10000, 10, 1, 0, 0)
    assert math.isinf(a.tmax)
    d = culamp_amd.EOTArgs()
    c = d.to_c()
    for f, _ in culamp_args._fields_:
        if f != "reserved":
            assert getattr(c, f) == getattr(a, f), f


def test_struct_layout_matches_header():
    from culamp_amd._lib import culamp_args, culamp_result
    assert C.sizeof(culamp_args) == 8 * 11 + 4 * 4
    assert C.sizeof(culamp_result) == 8 * 11 + 4 * 2


def test_no_device_is_an_error_not_a_fallback():
    import culamp_amd
    if culamp_amd.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(culamp_amd.CulampError) as e:
        culamp_amd.Context(0)
    assert e.value.code == 5  # CULAMP_ERR_NO_DEVICE
    assert "no HIP device" in str(e.value)
    with pytest.raises(culamp_amd.CulampError):
        culamp_amd.sinkhorn_color_transfer(np.zeros((2, 3)), np.zeros((2, 3)), np.ones(2) / 2,
                                           np.ones(2) / 2, culamp_amd.EOTArgs(eta_p=0.1))


def test_null_arguments_rejected():
    import culamp_amd
    lib = culamp_amd.load()
    assert lib.culamp_ctx_create(0, None) == 1
    assert b"null" in lib.culamp_last_error()
    assert lib.culamp_ctx_set_launch(None, 1, 1) == 1
    assert lib.culamp_ctx_destroy(None) == 0
    assert lib.culamp_problem_destroy(None) == 0


def test_read_args_json(tmp_path):
    import culamp_amd
    # the keys of the reference's configurations/default_1e-02.json
    cfg = {"eta_p": 1e-2, "eta_mu": 0.0, "tau_p": 1.0, "tau_mu": 1.0, "alpha": 0.01, "B": 1.1,
           "epsilon": 1e-10, "itermax": 1000000, "anneal_mult": 0.9, "verbose": True, "tmax": 600,
           "inner_iter": 10}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(cfg))
    a = culamp_amd.read_args_json(str(p))
    assert a.eta_p == 1e-2 and a.B == 1.1 and a.itermax == 1000000 and a.tmax == 600.0
    assert a.precision == "fp64" and a.stop == "reference"
    # missing keys keep defaults, unknown keys are ignored, new keys parse
    p.write_text(json.dumps({"eta_p": 0, "foo": 3, "precision": "mixed", "stop": "infeas"}))
    a = culamp_amd.read_args_json(str(p))
    assert a.eta_p == 0.0 and a.epsilon == 1e-4 and a.precision == "mixed" and a.stop == "infeas"
    assert a.to_c().precision == 1 and a.to_c().stop == 1
    for bad in ({"itermax": 1.5}, {"verbose": 1}, {"precision": "fp16"}, {"eta_p": "x"}):
        p.write_text(json.dumps(bad))
        with pytest.raises(ValueError):
            culamp_amd.read_args_json(str(p))


def test_julia_float_text():
    from culamp_amd.solvers import julia_float_str
    assert julia_float_str(0.1) == "0.1"
    assert julia_float_str(1.0) == "1.0"
    assert julia_float_str(-2.5e-7) == "-2.5e-7"
    assert julia_float_str(1e-5) == "1.0e-5"
    assert julia_float_str(0.0001) == "0.0001"
    assert julia_float_str(123456.0) == "123456.0"
    assert julia_float_str(float("inf")) == "Inf" and julia_float_str(float("nan")) == "NaN"


def test_oracle_is_not_linked_into_the_product():
    """libculamp must not reference the oracle in any form."""
    import culamp_amd
    from culamp_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"or_lse_rows" not in data and b"liboracle" not in data
    src = os.path.join(ROOT, "culamp.jl_amd")
    for dp, _, fs in os.walk(src):
        for f in fs:
            if f.endswith((".py", ".hip", ".h")):
                t = open(os.path.join(dp, f)).read()
                assert "oracle" not in t.lower(), f
