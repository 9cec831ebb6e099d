"""Pins of the oracle's kernel time integral (code/srcF.kernel/SI_ker_sta.F90):
  - stress2strain (:1185-1211) inverts isotropic Hooke's law (round trip)
  - sgt2scatter3 (:1290-1331) equals the causal convolutions computed with
    numpy.convolve in FP64
  - filter / filtfilt (:1433-1557) equal scipy.signal.lfilter with the
    reference's own padding (zero front pad + odd end pad for P1; odd pads
    at both ends, NFACT = 3*NOD, for P2)
  - scatter2kernel (:1375-1431) is exact for linear integrands (trapezoid)
"""
import ctypes as C

import numpy as np
import pytest
from scipy import signal

from oracle.pyoracle import Lib


def P(lib, a):
    return a.ctypes.data_as(C.POINTER(lib.R))


def test_stress2strain_roundtrip():
    lib = Lib(64)
    rng = np.random.default_rng(0)
    npt, knt = 50, 7
    lam = rng.uniform(1e10, 5e10, npt)
    mu = rng.uniform(1e10, 4e10, npt)
    S = [rng.standard_normal((knt, npt)) * 1e6 for _ in range(6)]
    E = [s.copy() for s in S]
    lib.stress2strain(npt, knt, *[P(lib, e) for e in E], P(lib, lam), P(lib, mu))
    tr = E[0] + E[1] + E[2]
    back = [lam * tr + 2 * mu * E[q] for q in range(3)] + [2 * mu * E[q] for q in range(3, 6)]
    for q in range(6):
        np.testing.assert_allclose(back[q], S[q], rtol=1e-12, atol=1e-6)


@pytest.mark.parametrize("prec", [32, 64])
def test_scatter_is_causal_convolution(prec):
    lib = Lib(prec)
    rng = np.random.default_rng(1)
    npt, knt, dt = 9, 37, 0.25
    R = [rng.standard_normal((knt, npt)).astype(lib.dtype) for _ in range(6)]
    S = [rng.standard_normal((knt, npt)).astype(lib.dtype) for _ in range(6)]
    lam = rng.uniform(1.0, 3.0, npt).astype(lib.dtype)
    mu = rng.uniform(0.5, 2.0, npt).astype(lib.dtype)
    A = np.zeros((knt, npt), lib.dtype)
    B = np.zeros((knt, npt), lib.dtype)
    RP = (C.POINTER(lib.R) * 6)(*[P(lib, r) for r in R])
    SP = (C.POINTER(lib.R) * 6)(*[P(lib, s) for s in S])
    lib.scatter1(npt, knt, dt, RP, SP, P(lib, lam), P(lib, mu), P(lib, A), P(lib, B))
    w = [1, 1, 1, 2, 2, 2]
    for p in range(npt):
        trR = sum(R[c][:, p].astype(float) for c in range(3))
        trS = sum(S[c][:, p].astype(float) for c in range(3))
        # x(n) = dt * sum_{m=1}^{n-1} r(m) s(n-m)  ->  full convolution index n-2
        x1 = dt * np.concatenate([[0.0], np.convolve(trR, trS)[: knt - 1]])
        x2 = dt * np.concatenate([[0.0], sum(w[c] * np.convolve(R[c][:, p].astype(float), S[c][:, p].astype(float))
                                             for c in range(6))[: knt - 1]])
        ra = x1 * 2 * (lam[p] + 2 * mu[p])
        rb = 2 * (x2 - x1) * 2 * mu[p]
        tol = 2e-5 if prec == 32 else 1e-12
        np.testing.assert_allclose(A[:, p], ra, rtol=tol, atol=tol * np.abs(ra).max())
        np.testing.assert_allclose(B[:, p], rb, rtol=tol, atol=tol * np.abs(rb).max())


def butter(nod, wn=0.2):
    b, a = signal.butter(nod, wn)
    return np.ascontiguousarray(b, float), np.ascontiguousarray(a, float)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("nod", [2, 4])
def test_filters_match_scipy(mode, nod):
    lib = Lib(64)
    rng = np.random.default_rng(2 + nod)
    npt, knt = 5, 120
    b, a = butter(nod)
    nod1 = nod + 1
    NFACT = 3 * nod
    x = rng.standard_normal((knt, npt))
    y = np.zeros_like(x)
    D = C.POINTER(C.c_double)
    lib.series_filter(mode, npt, knt, nod1, b.ctypes.data_as(D), a.ctypes.data_as(D), P(lib, x), P(lib, y))
    for p in range(npt):
        s = x[:, p]
        end = 2 * s[-1] - s[-1 - np.arange(1, NFACT + 1)]
        if mode == 1:
            front = np.zeros(NFACT)
        else:
            front = (2 * s[0] - s[1:NFACT + 1])[::-1]
        v = np.concatenate([front, s, end])
        if mode == 1:
            out = signal.lfilter(b, a, v)
        else:
            out = signal.lfilter(b, a, signal.lfilter(b, a, v)[::-1])[::-1]
        np.testing.assert_allclose(y[:, p], out[NFACT:NFACT + knt], rtol=1e-10, atol=1e-12)


def test_scatter2kernel_linear_exact():
    lib = Lib(64)
    npt, knt, dt, M0 = 4, 50, 0.1, 1e15
    t = dt * np.arange(1, knt + 1)
    V = np.ascontiguousarray(np.ones(knt))
    U = np.ascontiguousarray(2.0 * np.ones(knt))
    slope = np.array([1.0, -2.0, 0.5, 3.0])
    Sa = np.ascontiguousarray(np.outer(t, slope))
    Sb = np.ascontiguousarray(np.outer(t, slope) + 1.0)
    out = [np.zeros(npt) for _ in range(4)]
    nt1, nt2 = 11, 41
    lib.scatter2kernel(npt, knt, nt1, nt2, dt, M0, P(lib, V), P(lib, U), P(lib, Sa), P(lib, Sb), *[P(lib, o) for o in out])
    t1, t2 = t[nt1 - 1], t[nt2 - 1]
    L = t2 - t1
    ints = slope * (t2 ** 2 - t1 ** 2) / 2
    np.testing.assert_allclose(out[0], ints / (L * M0), rtol=1e-12)          # V = 1
    np.testing.assert_allclose(out[1], 2 * ints / (4 * L * M0), rtol=1e-12)  # U = 2
    np.testing.assert_allclose(out[2], (ints + L) / (L * M0), rtol=1e-12)
    np.testing.assert_allclose(out[3], 2 * (ints + L) / (4 * L * M0), rtol=1e-12)


def test_rotate_is_matrix_product():
    lib = Lib(64)
    rng = np.random.default_rng(3)
    n = 40
    X, Y, Z = (rng.standard_normal(n) for _ in range(3))
    Rm = rng.standard_normal((3, 3))
    rot = np.ascontiguousarray(Rm.ravel(order="F"))
    x, y, z = X.copy(), Y.copy(), Z.copy()
    lib.scatter_rotate(n, 1, P(lib, x), P(lib, y), P(lib, z), P(lib, rot))
    v = np.stack([X, Y, Z])  # out_c = sum_r v_r rot(r, c)
    want = Rm.T @ v
    np.testing.assert_allclose(np.stack([x, y, z]), want, rtol=1e-13)
