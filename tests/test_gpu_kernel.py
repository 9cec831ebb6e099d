"""GPU parity of the sensitivity-kernel time integral against the oracle.

Tolerances: the GPU accumulates the convolution and trapezoid sums in FP64
(products in FP32); it must stay within 1e-5 relative L2 of the FP32 oracle
and within 1e-6 of the FP64 oracle run on the same FP32 inputs.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def dev(gpu_lib):
    from fwant_amd._lib import DeviceArray
    return DeviceArray.from_host


def ocall(lib, name, *args):
    return getattr(lib, name)(*args)


def P(lib, a):
    return a.ctypes.data_as(C.POINTER(lib.R))


def test_stress2strain(dev):
    from fwant_amd import kernel
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(0)
    npt, knt = 777, 33
    lam = rng.uniform(1e10, 5e10, npt).astype(np.float32)
    mu = rng.uniform(1e10, 4e10, npt).astype(np.float32)
    E = (rng.standard_normal((6, knt, npt)) * 1e6).astype(np.float32)
    dE = dev(E)
    kernel.stress2strain(dE, dev(lam), dev(mu))
    got = dE.to_host()
    lib = Lib(32)
    ref = [E[c].copy() for c in range(6)]
    lib.stress2strain(npt, knt, *[P(lib, r) for r in ref], P(lib, lam), P(lib, mu))
    for c in range(6):
        assert rel(got[c], ref[c]) < 1e-6


# knt >= 320 runs the LDS-tiled kernel (64-point by 32-output tiles; 1000
# points leave a partial point tile, 333 and 411 partial output tiles)
@pytest.mark.parametrize("knt", [5, 64, 203, 333, 411])
def test_scatter(dev, knt):
    from fwant_amd import kernel
    from oracle.pyoracle import Lib
    print(f"scatter knt={knt}", flush=True)
    rng = np.random.default_rng(knt)
    npt, dt = 1000, 0.4
    R = rng.standard_normal((6, knt, npt)).astype(np.float32)
    S = rng.standard_normal((6, knt, npt)).astype(np.float32)
    lam = rng.uniform(1.0, 3.0, npt).astype(np.float32)
    mu = rng.uniform(0.5, 2.0, npt).astype(np.float32)
    A, B = kernel.sgt2scatter(dev(R), dev(S),
                              dev(lam), dev(mu), dt)
    A, B = A.to_host(), B.to_host()
    for prec, tol in ((32, 1e-5), (64, 1e-6)):
        lib = Lib(prec)
        d = lib.dtype
        Rl = [np.ascontiguousarray(R[c], d) for c in range(6)]
        Sl = [np.ascontiguousarray(S[c], d) for c in range(6)]
        ra = np.zeros((knt, npt), d)
        rb = np.zeros((knt, npt), d)
        RP = (C.POINTER(lib.R) * 6)(*[P(lib, r) for r in Rl])
        SP = (C.POINTER(lib.R) * 6)(*[P(lib, s) for s in Sl])
        lib.scatter1(npt, knt, dt, RP, SP, P(lib, lam.astype(d)), P(lib, mu.astype(d)), P(lib, ra), P(lib, rb))
        ea, eb = rel(A, ra), rel(B, rb)
        print(f"prec {prec}: alpha {ea:.2e} beta {eb:.2e}")
        assert ea < tol and eb < tol


@pytest.mark.parametrize("mode", [1, 2])
def test_filter(dev, mode):
    from scipy import signal
    from fwant_amd import kernel
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(mode)
    npt, knt = 513, 150
    b, a = signal.butter(4, 0.15)
    x = rng.standard_normal((knt, npt)).astype(np.float32)
    y = kernel._filter(mode, dev(x), b, a).to_host()
    lib = Lib(64)
    xd = x.astype(np.float64)
    yd = np.zeros_like(xd)
    D = C.POINTER(C.c_double)
    lib.series_filter(mode, npt, knt, len(b), np.ascontiguousarray(b).ctypes.data_as(D),
                      np.ascontiguousarray(a).ctypes.data_as(D), P(lib, xd), P(lib, yd))
    assert rel(y, yd) < 1e-6
    # host synthetic filter equals the same padded recurrence
    hs = kernel.filter_series_host(x[:, 7], b, a, mode)
    assert rel(hs, yd[:, 7]) < 1e-6


def test_scatter2kernel(dev):
    from fwant_amd import kernel
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(4)
    npt, knt, dt, M0 = 2000, 120, 0.2, 1e15
    V = rng.standard_normal(knt).astype(np.float32)
    U = rng.standard_normal(knt).astype(np.float32)
    Sa = rng.standard_normal((knt, npt)).astype(np.float32)
    Sb = rng.standard_normal((knt, npt)).astype(np.float32)
    out = kernel.scatter2kernel(V, U, dev(Sa), dev(Sb), 17, 95, dt, M0)
    got = [o.to_host() for o in out]
    for prec, tol in ((32, 1e-5), (64, 1e-6)):
        lib = Lib(prec)
        d = lib.dtype
        ref = [np.zeros(npt, d) for _ in range(4)]
        lib.scatter2kernel(npt, knt, 17, 95, dt, M0, P(lib, V.astype(d)), P(lib, U.astype(d)),
                           P(lib, Sa.astype(d)), P(lib, Sb.astype(d)), *[P(lib, r) for r in ref])
        for g, r in zip(got, ref):
            assert rel(g, r) < tol


def test_rotate(dev):
    from fwant_amd import kernel
    rng = np.random.default_rng(5)
    X, Y, Z = (rng.standard_normal((4, 300)).astype(np.float32) for _ in range(3))
    Rm = rng.standard_normal((3, 3)).astype(np.float32)
    dX, dY, dZ = (dev(v) for v in (X, Y, Z))
    kernel.scatter_rotate(dX, dY, dZ, Rm)
    want = np.einsum("rc,r...->c...", Rm.astype(np.float64), np.stack([X, Y, Z]).astype(np.float64))
    got = np.stack([dX.to_host(), dY.to_host(), dZ.to_host()])
    assert rel(got, want) < 1e-6


def test_pipeline_on_solver_sgt(dev):
    """Station SGT (force at a receiver) and event wavefield from the GPU
    solver, captured as 6-component stress snapshots, through the whole
    kernel pipeline on the GPU vs the FP32 oracle pipeline on the same SGT."""
    from scipy import signal
    from fwant_amd import kernel, models
    from fwant_amd.solver import Solver
    from oracle.pyoracle import Lib
    snap = dict(subs=(8, 8, 10), subc=(12, 10, 6), subt=(2, 2, 2), tinv=2, fields=("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"),
                max_records=40)
    sgts = []
    for seed, fz in ((31, 1e15), (32, -2e15)):
        m = models.small_test(nt=80, seed=seed, moment=False)
        m.force.comps[2] = fz
        s = Solver(m, snapshots=[snap])
        s.run(80)
        sgts.append(s.snapshot(0))  # (knt, 6, npt) in field order Txx Tyy Tzz Txy Txz Tyz
        s.close()
    knt, _, npt = sgts[0].shape
    lam_full = m.lam[7:7 + 24:2, 7:7 + 20:2, 9:9 + 12:2].ravel(order="F").astype(np.float32)
    mu_full = m.mu[7:7 + 24:2, 7:7 + 20:2, 9:9 + 12:2].ravel(order="F").astype(np.float32)
    R = np.ascontiguousarray(sgts[0].transpose(1, 0, 2))
    S = np.ascontiguousarray(sgts[1].transpose(1, 0, 2))
    dt = 2 * m.stept
    t = dt * np.arange(1, knt + 1)
    V = np.sin(2 * np.pi * 0.3 * t).astype(np.float32)
    U = np.cos(2 * np.pi * 0.3 * t).astype(np.float32)
    b, a = signal.butter(2, 0.3)
    dR, dS = dev(R), dev(S)
    dl, dm = dev(lam_full), dev(mu_full)
    got = [o.to_host() for o in kernel.pair_kernel(None, None, dR, dS, dl, dm, dt, V, U, 5, 36, 1e15, 2,
                                                       filt=("P2", b, a))]
    # oracle pipeline
    lib = Lib(32)
    Rl = [np.ascontiguousarray(R[c]) for c in range(6)]
    Sl = [np.ascontiguousarray(S[c]) for c in range(6)]
    lib.stress2strain(npt, knt, *[P(lib, r) for r in Rl], P(lib, lam_full), P(lib, mu_full))
    lib.stress2strain(npt, knt, *[P(lib, r) for r in Sl], P(lib, lam_full), P(lib, mu_full))
    A = np.zeros((knt, npt), np.float32)
    B = np.zeros((knt, npt), np.float32)
    lib.scatter1(npt, knt, dt, (C.POINTER(lib.R) * 6)(*[P(lib, r) for r in Rl]),
                 (C.POINTER(lib.R) * 6)(*[P(lib, s) for s in Sl]), P(lib, lam_full), P(lib, mu_full), P(lib, A), P(lib, B))
    D = C.POINTER(C.c_double)
    bb, aa = np.ascontiguousarray(b), np.ascontiguousarray(a)
    Af, Bf = np.zeros_like(A), np.zeros_like(B)
    lib.series_filter(2, npt, knt, len(b), bb.ctypes.data_as(D), aa.ctypes.data_as(D), P(lib, A), P(lib, Af))
    lib.series_filter(2, npt, knt, len(b), bb.ctypes.data_as(D), aa.ctypes.data_as(D), P(lib, B), P(lib, Bf))
    Vf = np.zeros(knt, np.float32)
    Uf = np.zeros(knt, np.float32)
    lib.series_filter(2, 1, knt, len(b), bb.ctypes.data_as(D), aa.ctypes.data_as(D), P(lib, V.copy()), P(lib, Vf))
    lib.series_filter(2, 1, knt, len(b), bb.ctypes.data_as(D), aa.ctypes.data_as(D), P(lib, U.copy()), P(lib, Uf))
    ref = [np.zeros(npt, np.float32) for _ in range(4)]
    lib.scatter2kernel(npt, knt, 5, 36, dt, 1e15, P(lib, Vf), P(lib, Uf), P(lib, Af), P(lib, Bf), *[P(lib, r) for r in ref])
    for name, g, r in zip(("Kap", "Kaq", "Kbp", "Kbq"), got, ref):
        e = rel(g, r)
        print(f"{name}: rel-L2 {e:.2e}")
        assert e < 1e-5
