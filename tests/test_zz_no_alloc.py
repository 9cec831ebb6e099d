"""The allocation-free scatter2kernel against the allocating one.

fwk_scatter2kernel_noalloc stages V and U in a fixed device array instead
of a per-call allocation; its results must equal fwk_scatter2kernel's bit
for bit, on two calls in a row (the second with a smaller input, so stale
staged values from the first call would show), through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, name, npt, knt, nt1, nt2, dt, M0, V, U, dSa, dSb, outs):
    from fwant_amd._lib import fptr
    rc = getattr(L, name)(npt, knt, nt1, nt2, dt, M0, fptr(V), fptr(U), dSa.ptr, dSb.ptr,
                          *[o.ptr for o in outs], None)
    assert rc == 0, L.fwk_last_error().decode()
    return [o.to_host() for o in outs]


def test_noalloc_matches_allocating_entry(gpu_lib):
    from fwant_amd._lib import DeviceArray
    L = gpu_lib
    rng = np.random.default_rng(3)
    for npt, knt, nt1, nt2 in ((5000, 700, 17, 650), (1200, 90, 3, 80)):
        print(f"npt={npt} knt={knt} window=({nt1},{nt2})", flush=True)
        V = rng.standard_normal(knt).astype(np.float32)
        U = rng.standard_normal(knt).astype(np.float32)
        Sa = rng.standard_normal((knt, npt)).astype(np.float32)
        Sb = rng.standard_normal((knt, npt)).astype(np.float32)
        dSa, dSb = DeviceArray.from_host(Sa), DeviceArray.from_host(Sb)
        ref = [DeviceArray((npt,)) for _ in range(4)]
        got = [DeviceArray((npt,)) for _ in range(4)]
        r = _call(L, "fwk_scatter2kernel", npt, knt, nt1, nt2, 0.1, 1e15, V, U, dSa, dSb, ref)
        g = _call(L, "fwk_scatter2kernel_noalloc", npt, knt, nt1, nt2, 0.1, 1e15, V, U, dSa, dSb, got)
        for a, b in zip(g, r):
            assert np.isfinite(b).all() and np.abs(b).max() > 0
            np.testing.assert_array_equal(a, b)
        for d in [dSa, dSb] + ref + got:
            d.free()


def test_noalloc_rejects_long_series(gpu_lib):
    from fwant_amd._lib import DeviceArray, fptr
    L = gpu_lib
    knt, npt = 16385, 4
    V = np.ones(knt, np.float32)
    d = DeviceArray((npt,))
    rc = L.fwk_scatter2kernel_noalloc(npt, knt, 1, 10, 0.1, 1.0, fptr(V), fptr(V), d.ptr, d.ptr, d.ptr, d.ptr,
                                      d.ptr, d.ptr, None)
    assert rc != 0
    assert "16384" in L.fwk_last_error().decode()
    d.free()
