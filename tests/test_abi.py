"""The C ABI: libfwant.so loads (no GPU needed to load it) and exports every
function include/fwant.h declares; the Python binding declares the same set."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fwant.h")
LIB = os.path.join(ROOT, "fwant_amd", "libfwant.so")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fw(?:k|l|ae)?_[a-z0-9_]+)\s*\(", txt)))


def test_header_parses():
    fns = header_functions()
    assert "fw_create" in fns and "fwk_scatter" in fns and len(fns) >= 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfwant.so not built")
def test_library_exports_header():
    lib = C.CDLL(LIB)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_binding_matches_header():
    from fwant_amd import _lib
    assert sorted(_lib.SYMBOLS) == header_functions()


def test_reference_step_count():
    # 8*ceil((nt+1)/8): the driver tests ntime > nt only at the top of each 8-step block
    from fwant_amd import models
    m = models.two_layer_100x100x50()
    assert m.ref_steps() == 1008
    if os.path.exists(LIB):
        from fwant_amd._lib import lib
        L = lib()
        assert [L.fw_reference_steps(n) for n in (0, 7, 8, 1000, 3000)] == [8, 8, 16, 1008, 3008]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfwant.so not built")
def test_hbm_probe_rejects_bad_arguments():
    # argument checks come before any device call
    lib = C.CDLL(LIB)
    lib.fw_hbm_probe.argtypes = [C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    r, c = C.c_float(), C.c_float()
    assert lib.fw_hbm_probe(1024, 1024, 2, C.byref(r), C.byref(c)) != 0  # below 1 MiB
    assert lib.fw_hbm_probe(1 << 26, 0, 2, C.byref(r), C.byref(c)) != 0
    assert lib.fw_hbm_probe(1 << 26, 1024, 2, None, C.byref(c)) != 0
