"""The C-ABI library without a GPU: it loads, exports every declared entry
point, and rejects bad arguments before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import particles_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ptc_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("ptc_fields_create", "ptc_integrate", "ptc_particles_advance", "ptc_last_error"):
        assert must in names


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    lib = C.CDLL(P.library_path)
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_options_roundtrip_and_validation():
    old = P.get_option("layout")
    try:
        P.set_option("layout", "soa")
        assert P.get_option("layout") == "soa"
        P.set_option("layout", "aos")
        assert P.get_option("layout") == "aos"
        with pytest.raises(P.PtcError) as e:
            P.set_option("layout", "bricks?")
        assert e.value.code == 1
        with pytest.raises(P.PtcError):
            P.set_option("block", "100")
        with pytest.raises(P.PtcError):
            P.set_option("no_such_option", "1")
    finally:
        P.set_option("layout", old)


def test_fields_argument_validation_before_device():
    bad = np.zeros((1, 4, 4, 3))
    with pytest.raises(P.PtcError) as e:
        P.Fields(bad, bad)
    assert e.value.code == 1 and "dimensions" in str(e.value)
    with pytest.raises(ValueError):
        P.Fields(np.zeros((4, 4, 4, 3)), np.zeros((4, 4, 5, 3)))


def test_version_string():
    assert "gfx950" in P.version()
