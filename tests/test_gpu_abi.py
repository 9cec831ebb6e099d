"""C ABI calls that need a handle (a GPU): argument handling at the boundary."""
import ctypes as C

import pytest

from fwant_amd._lib import FwConfig

pytestmark = pytest.mark.gpu


def test_no_sources_with_null_arrays(gpu_lib):
    """n = 0 sources with NULL arrays and a nonzero window count: accepted
    without touching the pointers (a Fortran or C caller passing 'none')."""
    L = gpu_lib
    cfg = FwConfig()
    cfg.ni, cfg.nj, cfg.nk, cfg.stept, cfg.freenode = 16, 16, 16, 0.01, 1
    for n in range(6):
        cfg.abs_velocity[n], cfg.cfs_bmax[n], cfg.cfs_amax[n] = 6000.0, 3.0, 1.0
    h = C.c_void_p()
    assert L.fw_create(C.byref(cfg), C.byref(h)) == 0
    try:
        assert L.fw_set_force(h, 0, 5, 200, None, None, None, None, None, None, None, None) == 0
        assert L.fw_set_moment(h, 0, 3, 200, None, None, None, None, None, None) == 0
        assert L.fw_set_force(h, 1, 1, 200, None, None, None, None, None, None, None, None) != 0
    finally:
        L.fw_destroy(h)
