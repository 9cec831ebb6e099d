"""Pins of the oracle's finite-difference operators, RK4 and metric.

Each check is against what the mathematics or the reference's coefficient
table fixes, not against a restatement of the oracle's own code:
  - DRP/opt MacCormack taps of code/srcF/mod_macdrp.h:19 (F on -1..+3; B the
    mirrored, negated stencil on -3..+1, mod_macdrp.h:69-85): polynomial
    moments of the operator equal the moments of the coefficient table
  - 2-4 / 2-2 operators (mod_macdrp.h:230-268) are exact on linear data
  - RK4 in 4-register form (mod_macdrp.F90:291-393) reproduces the RK4
    stability polynomial on y' = lambda y
  - metric of a uniform grid (seis3d_metric.F90:82-112) is 1/(0.99998 h)
"""
import ctypes as C

import numpy as np
import pytest

from oracle.pyoracle import Lib, build_metric

# reference coefficient table, code/srcF/mod_macdrp.h:19
MCA = {-1: -0.30874, 0: -0.6326, 1: 1.233, 2: -0.3334, 3: 0.04168}


def op(lib, kind, fwd, v, centre):
    a = np.ascontiguousarray(v, lib.dtype)
    p = a.ctypes.data_as(C.POINTER(lib.R))
    addr = C.cast(p, C.c_void_p).value + centre * a.itemsize
    return lib.op_apply(kind, fwd, C.cast(addr, C.POINTER(lib.R)))


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("power", [0, 1, 2, 3])
def test_drp_moments(prec, power):
    lib = Lib(prec)
    k = np.arange(-3, 4)
    v = k.astype(float) ** power
    momF = sum(c * o ** power for o, c in MCA.items())
    momB = sum(-c * (-o) ** power for o, c in MCA.items())
    tol = 1e-6 if prec == 32 else 1e-13
    assert op(lib, 0, 1, v, 3) == pytest.approx(momF, abs=tol * 10)
    assert op(lib, 0, 0, v, 3) == pytest.approx(momB, abs=tol * 10)


def test_drp_known_moments():
    # sum of taps is -6e-5 (F) / +6e-5 (B), not zero; first moment 0.99998 for both
    lib = Lib(64)
    one = np.ones(7)
    lin = np.arange(-3, 4, dtype=float)
    assert op(lib, 0, 1, one, 3) == pytest.approx(-6.0e-5, abs=1e-15)
    assert op(lib, 0, 0, one, 3) == pytest.approx(6.0e-5, abs=1e-15)
    assert op(lib, 0, 1, lin, 3) == pytest.approx(0.99998, abs=1e-14)
    assert op(lib, 0, 0, lin, 3) == pytest.approx(0.99998, abs=1e-14)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("fwd", [0, 1])
def test_low_order_exact_on_linear(kind, fwd):
    lib = Lib(64)
    for a0, a1 in [(0.0, 1.0), (2.5, -3.0), (1e3, 0.25)]:
        v = a0 + a1 * np.arange(-3, 4, dtype=float)
        assert op(lib, kind, fwd, v, 3) == pytest.approx(a1, rel=1e-13, abs=1e-12)


def test_low_order_known_quadratic_error():
    # 2-4 forward on k^2: (8 - 4)/6 = 2/3 (one-sided, second-order); 2-2 forward: 1
    lib = Lib(64)
    q = np.arange(-3, 4, dtype=float) ** 2
    assert op(lib, 1, 1, q, 3) == pytest.approx(2.0 / 3.0, rel=1e-14)
    assert op(lib, 2, 1, q, 3) == pytest.approx(1.0, rel=1e-14)
    assert op(lib, 1, 0, q, 3) == pytest.approx(-2.0 / 3.0, rel=1e-14)


@pytest.mark.parametrize("lam_dt", [-0.5, -1.3, 0.2, -2.7])
def test_rk4_stability_polynomial(lam_dt):
    lib = Lib(64)
    a4 = np.zeros(4)
    b4 = np.zeros(4)
    P = C.POINTER(C.c_double)
    lib.rk_coef(a4.ctypes.data_as(P), b4.ctypes.data_as(P))
    dt = 0.1
    lam = lam_dt / dt
    y0 = np.array([1.0, -2.0, 0.5])
    W = y0.copy()
    m = y0.copy()
    t = np.zeros(3)
    for st in range(4):
        h = lam * W
        kind = 0 if st == 0 else (2 if st == 3 else 1)
        lib.rk_stage(kind, a4[st], b4[st], dt, 3, W.ctypes.data_as(P), m.ctypes.data_as(P),
                     t.ctypes.data_as(P), h.ctypes.data_as(P))
    z = lam_dt
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(W, R * y0, rtol=1e-14)


def test_rk_stage_pattern_cycle():
    """8-step F/B cycle of seis3d_wave_8op_p3.F90:82-591: block patterns
    BBB FFB FFF BBF BFB FBB FBF BFF, each applied as P, ~P, P, ~P."""
    lib = Lib(32)
    blocks = ["BBB", "FFB", "FFF", "BBF", "BFB", "FBB", "FBF", "BFF"]
    p = (C.c_int * 3)()
    for nt in range(24):
        for st in range(4):
            lib.stage_pattern(nt, st, p)
            want = blocks[nt % 8]
            if st % 2:
                want = "".join("F" if c == "B" else "B" for c in want)
            assert "".join("F" if v else "B" for v in p) == want


def test_metric_uniform_grid():
    from fwant_amd import models
    m = models.small_test(ni=20, nj=18, nk=16, nt=1, h=2000.0, colat0=1.1, r_top=5.0e6)
    g = build_metric(m)
    dth = m.x[1] - m.x[0]
    dr = m.z[1] - m.z[0]
    np.testing.assert_allclose(g["xi_x"], 1.0 / (0.99998 * dth), rtol=1e-9)
    np.testing.assert_allclose(g["eta_y"], 1.0 / (0.99998 * dth), rtol=1e-9)
    np.testing.assert_allclose(g["zeta_z"], 1.0 / (0.99998 * dr), rtol=1e-9)
    np.testing.assert_allclose(g["xsin"], np.sin(m.x), rtol=1e-15)
    np.testing.assert_allclose(g["xcot"], np.cos(m.x) / np.sin(m.x), rtol=1e-13)
