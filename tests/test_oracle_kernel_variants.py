"""Pins of the oracle's kernel variants (code/srcF.kernel):
  - lambda/mu scattered series (SI_ker_sta_lamu.F90:1290-1336) and the
    density series from velocities (SI_ker_sta_rho.F90:1478-1552) equal
    numpy FP64 causal convolutions;
  - eoshift equals Fortran's end-off shift (zero fill) on both directions;
  - the three re-parameterisations (SI_ker_lamu2kamu / lamu2vpvs /
    vpvs2kamu) conserve the first-order traveltime change: for any
    perturbation, sum K_old * dln m_old = sum K_new * dln m_new, with
    kappa = lambda + 2mu/3, Vp^2 = (lambda + 2mu)/rho, Vs^2 = mu/rho at
    fixed rho.  A wrong factor or sign in any line breaks the identity.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.pyoracle import Lib


def P(lib, a):
    return a.ctypes.data_as(C.POINTER(lib.R))


def conv(r, s, dt, knt):
    return dt * np.concatenate([[0.0], np.convolve(r, s)[: knt - 1]])


def test_lamu_scatter_is_convolution():
    lib = Lib(64)
    rng = np.random.default_rng(5)
    npt, knt, dt = 7, 29, 0.3
    R = [rng.standard_normal((knt, npt)) for _ in range(6)]
    S = [rng.standard_normal((knt, npt)) for _ in range(6)]
    lam, mu = rng.uniform(1, 3, npt), rng.uniform(0.5, 2, npt)
    A, B = np.zeros((knt, npt)), np.zeros((knt, npt))
    lib.scatter1_lamu(npt, knt, dt, (C.POINTER(lib.R) * 6)(*[P(lib, r) for r in R]),
                      (C.POINTER(lib.R) * 6)(*[P(lib, q) for q in S]), P(lib, lam), P(lib, mu), P(lib, A), P(lib, B))
    w = [1, 1, 1, 2, 2, 2]
    for p in range(npt):
        x1 = conv(sum(R[c][:, p] for c in range(3)), sum(S[c][:, p] for c in range(3)), dt, knt)
        x2 = sum(w[c] * conv(R[c][:, p], S[c][:, p], dt, knt) for c in range(6))
        np.testing.assert_allclose(A[:, p], x1 * lam[p], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(B[:, p], 2 * x2 * mu[p], rtol=1e-10, atol=1e-12)


def test_density_scatter_is_convolution():
    lib = Lib(64)
    rng = np.random.default_rng(6)
    npt, knt, dt = 5, 33, 0.2
    R = [rng.standard_normal((knt, npt)) for _ in range(3)]
    S = [rng.standard_normal((knt, npt)) for _ in range(3)]
    rho = rng.uniform(2000, 3000, npt)
    out = np.zeros((knt, npt))
    lib.vscatter1(npt, knt, dt, (C.POINTER(lib.R) * 3)(*[P(lib, r) for r in R]),
                  (C.POINTER(lib.R) * 3)(*[P(lib, q) for q in S]), P(lib, rho), P(lib, out))
    for p in range(npt):
        x1 = conv(sum(R[c][:, p] for c in range(3)), sum(S[c][:, p] for c in range(3)), dt, knt)
        np.testing.assert_allclose(out[:, p], x1 * rho[p], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("shift", [9, -9, 0, 40])
def test_eoshift_fortran_semantics(shift):
    lib = Lib(64)
    knt, npt = 30, 4
    x = np.arange(knt * npt, dtype=np.float64).reshape(knt, npt) + 1
    y = np.zeros_like(x)
    lib.eoshift(npt, knt, shift, P(lib, x), P(lib, y))
    want = np.zeros_like(x)
    if shift >= 0:
        want[: max(knt - shift, 0)] = x[shift:]
    else:
        want[-shift:] = x[: knt + shift]
    np.testing.assert_array_equal(y, want)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_reparam_conserves_first_order_change(mode):
    lib = Lib(64)
    rng = np.random.default_rng(7 + mode)
    n = 200
    lam, mu = rng.uniform(2e10, 6e10, n), rng.uniform(1e10, 4e10, n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    a2, b2 = a.copy(), b.copy()
    lib.reparam(mode, n, P(lib, lam), P(lib, mu), P(lib, a2), P(lib, b2))
    dl, dm = rng.standard_normal(n) * 1e-3, rng.standard_normal(n) * 1e-3  # dln lambda, dln mu
    if mode == 2:  # input in (Vp, Vs): express the same perturbation as dln Vp, dln Vs
        d_old = (a, b, (lam * dl + 2 * mu * dm) / (2 * (lam + 2 * mu)), dm / 2)
    else:
        d_old = (a, b, dl, dm)
    old = d_old[0] * d_old[2] + d_old[1] * d_old[3]
    if mode == 1:  # output (Vp, Vs)
        new = a2 * (lam * dl + 2 * mu * dm) / (2 * (lam + 2 * mu)) + b2 * dm / 2
    else:          # output (kappa, mu)
        kap = lam + 2 * mu / 3
        new = a2 * (lam * dl + 2 * mu * dm / 3) / kap + b2 * dm
    np.testing.assert_allclose(new, old, rtol=1e-9, atol=1e-12)
