"""Host-side checks of the C-ABI library that need no GPU: it loads, exports every
entry point include/avbd_hip.h declares, its records have the header's layout,
and its constructor-field builders equal the oracle's (and the reference's)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "avbd_hip.h")


def _lib():
    from avbd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libavbd_hip.so is not built: run __graft_entry__.build()")
    return capi.load()


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(avbd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_record_layouts_match_header():
    from avbd.state import check_layouts
    lib = _lib()
    check_layouts(lib.avbd_sizeof)
    assert lib.avbd_sizeof(99) == -1


def test_default_params_match_reference_defaults():
    from avbd.state import PARAMS
    lib = _lib()
    p = np.zeros((), PARAMS)
    assert lib.avbd_default_params(p.ctypes.data_as(ctypes.c_void_p)) == 0
    assert p["iterations"] == 10 and p["gravity"] == np.float32(-10.0)
    assert p["dt"] == np.float32(1.0 / 60.0) and p["alpha"] == np.float32(0.99)
    assert p["beta_lin"] == np.float32(1e4) and p["beta_ang"] == np.float32(100) and p["gamma"] == np.float32(0.999)


def test_record_builders_equal_oracle(oracle):
    from avbd import capi, scenes
    _lib()
    for sc in (scenes.pyramid_chain(base=5, links=30), scenes.pile(200, seed=4), scenes.stack()):
        a = scenes.build_state(sc, capi.Maker())
        b = scenes.build_state(sc, oracle.OracleMaker())
        for f in a.bodies.dtype.names:
            np.testing.assert_array_equal(a.bodies[f], b.bodies[f], err_msg=f)
        np.testing.assert_array_equal(a.joints, b.joints)
        np.testing.assert_array_equal(a.springs, b.springs)


def test_invalid_arguments_fail_cleanly():
    lib = _lib()
    assert lib.avbd_create(None, None) == -1
    assert b"null" in lib.avbd_last_error()
    assert lib.avbd_step(None, 1) == -1
    assert lib.avbd_default_params(None) == -1
    assert lib.avbd_make_box_body(None, 1.0, 0.5, None, None, None, None, None) == -1


def test_create_without_device_reports_no_device():
    """On a host without a GPU the library refuses to create a handle rather than falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from avbd import capi
    from avbd.state import default_params
    with pytest.raises(capi.AvbdError) as e:
        capi.HipWorld(default_params(), max_bodies=16)
    assert e.value.code == -4
