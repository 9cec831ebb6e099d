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


def test_fw_stage_order_and_equivalence(gpu_lib):
    """fw_stage: stages of a step in order 0..3 (FW_ERR_STATE otherwise, and
    fw_run refuses to start inside a step); with the 8-step cycle's
    patterns it reproduces fw_run bit for bit."""
    import numpy as np
    from fwant_amd import models
    from fwant_amd.solver import Solver
    m = models.small_test(ni=24, nj=22, nk=20, nt=16, seed=35)
    a = Solver(m)
    a.run(16)
    ra = a.receivers()
    b = Solver(m, metric=a.metric)
    L = gpu_lib
    assert L.fw_stage(b.h, 0, 1, 0, 0, 0) != 0  # stage 1 before stage 0
    pat = [0, 6, 7, 1, 2, 4, 5, 3]  # first-stage patterns of the 8-step block
    for t in range(16):
        for st in range(4):
            p = pat[t % 8] if st % 2 == 0 else 7 - pat[t % 8]
            assert L.fw_stage(b.h, t, st, (p >> 2) & 1, (p >> 1) & 1, p & 1) == 0
            if t == 3 and st == 1:
                assert L.fw_run(b.h, 3, 1) != 0  # inside a step
    np.testing.assert_array_equal(b.receivers(), ra)
    a.close()
    b.close()


def test_hbm_probe(gpu_lib):
    """The bandwidth probe returns positive, finite rates (a 256 MiB buffer
    fits the 256 MB Infinity Cache, so the rate may exceed the HBM peak;
    bench.py --stream-probe uses 2 GiB)."""
    import ctypes as C
    from fwant_amd._lib import lib
    r, c = C.c_float(), C.c_float()
    assert lib().fw_hbm_probe(1 << 28, 1024, 3, C.byref(r), C.byref(c)) == 0
    print(f"read {r.value:.0f} GB/s, copy {c.value:.0f} GB/s")
    assert 100 < r.value < 1e5 and 100 < c.value < 1e5
