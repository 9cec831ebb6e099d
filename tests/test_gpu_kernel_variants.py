"""GPU parity of the kernel variants against the oracle: lambda/mu and
density scattered series (1e-5 rel-L2 vs the FP32 oracle, FP64
accumulation on the GPU), eoshift (exact), re-parameterisation (FP32
pointwise, 1e-6)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def P(lib, a):
    return a.ctypes.data_as(C.POINTER(lib.R))


@pytest.mark.parametrize("knt", [61, 337])  # 337: the LDS-tiled scatter kernel
def test_lamu_and_density_scatter(gpu_lib, knt):
    from fwant_amd import kernel
    from fwant_amd._lib import DeviceArray
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(11)
    npt, dt = 700, 0.3
    R = rng.standard_normal((6, knt, npt)).astype(np.float32)
    S = rng.standard_normal((6, knt, npt)).astype(np.float32)
    lam = rng.uniform(1, 3, npt).astype(np.float32)
    mu = rng.uniform(0.5, 2, npt).astype(np.float32)
    rho = rng.uniform(2000, 3000, npt).astype(np.float32)
    d = DeviceArray.from_host
    A, B = kernel.sgt2scatter(d(R), d(S), d(lam), d(mu), dt, param="lamu")
    V = kernel.vel2scatter(d(R[:3]), d(S[:3]), d(rho), dt)
    lib = Lib(32)
    Rl = [np.ascontiguousarray(R[c]) for c in range(6)]
    Sl = [np.ascontiguousarray(S[c]) for c in range(6)]
    oA, oB, oV = (np.zeros((knt, npt), np.float32) for _ in range(3))
    RP = (C.POINTER(lib.R) * 6)(*[P(lib, r) for r in Rl])
    SP = (C.POINTER(lib.R) * 6)(*[P(lib, q) for q in Sl])
    lib.scatter1_lamu(npt, knt, dt, RP, SP, P(lib, lam), P(lib, mu), P(lib, oA), P(lib, oB))
    lib.vscatter1(npt, knt, dt, RP, SP, P(lib, rho), P(lib, oV))
    for name, g, o in (("lambda", A, oA), ("mu", B, oB), ("rho", V, oV)):
        e = rel(g.to_host(), o)
        print(f"{name} series rel-L2 vs FP32 oracle {e:.2e}")
        assert e < 1e-5


def test_eoshift_and_reparam(gpu_lib):
    from fwant_amd import kernel
    from fwant_amd._lib import DeviceArray
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(12)
    knt, npt = 40, 300
    E = rng.standard_normal((6, knt, npt)).astype(np.float32)
    lib = Lib(32)
    for shift in (9, -9):
        dE = DeviceArray.from_host(E)
        got = kernel.eoshift(dE, shift).to_host()
        for c in (0, 5):
            want = np.zeros((knt, npt), np.float32)
            lib.eoshift(npt, knt, shift, P(lib, np.ascontiguousarray(E[c])), P(lib, want))
            np.testing.assert_array_equal(got[c], want)
    n = 5000
    lam = rng.uniform(2e10, 6e10, n).astype(np.float32)
    mu = rng.uniform(1e10, 4e10, n).astype(np.float32)
    K = [rng.standard_normal(n).astype(np.float32) for _ in range(4)]
    for mode, name in enumerate(("lamu2kamu", "lamu2vpvs", "vpvs2kamu")):
        dK = [DeviceArray.from_host(k) for k in K]
        kernel.reparam(name, DeviceArray.from_host(lam), DeviceArray.from_host(mu), *dK)
        got = [k.to_host() for k in dK]
        ap, bp, aq, bq = K[0].copy(), K[2].copy(), K[1].copy(), K[3].copy()
        lib.reparam(mode, n, P(lib, lam), P(lib, mu), P(lib, ap), P(lib, bp))
        lib.reparam(mode, n, P(lib, lam), P(lib, mu), P(lib, aq), P(lib, bq))
        for g, w in zip(got, (ap, aq, bp, bq)):
            assert rel(g, w) < 1e-6, name
