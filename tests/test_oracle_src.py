"""Pins of the oracle's source-time functions and source smoothing
(code/srcF/mod_src.F90:419-622) by closed forms and calculus:
  - Ricker peak value at t = t0 is -sqrt(pi)/4 and src_svf is d/dt src_stf
  - Gaussian STF integrates to one; its svf is its time derivative
  - bell integrates to one and bell_int is its running integral
  - triangle (svf) integrates to one
  - the 7^3 smoothing weights sum to one, peak at the shift, and are
    separable Gaussians of radius 1.5 cells
"""
import ctypes as C

import numpy as np
import pytest

from oracle.pyoracle import Lib

RICKER, GAUSS, BELL, BELL_INT, TRI = 200, 220, 230, 100, 110


def stf(lib, which, sid, t, t0, f0, ntime=5, dt=0.01):
    return lib.stf(which, sid, t, t0, f0, ntime, dt)


def test_ricker_peak_and_derivative():
    lib = Lib(64)
    t0, fc = 2.4, 0.5
    assert stf(lib, 0, RICKER, t0, t0, fc) == pytest.approx(-np.sqrt(np.pi) / 4, rel=1e-14)
    for t in np.linspace(0.3, 5.0, 17):
        h = 1e-5
        num = (stf(lib, 0, RICKER, t + h, t0, fc) - stf(lib, 0, RICKER, t - h, t0, fc)) / (2 * h)
        assert stf(lib, 1, RICKER, t, t0, fc) == pytest.approx(num, rel=1e-7, abs=1e-9)
    assert stf(lib, 0, RICKER, -0.1, t0, fc) == 0.0
    assert stf(lib, 1, RICKER, 0.0, t0, fc) == 0.0


def test_ricker_zero_mean():
    lib = Lib(64)
    t = np.linspace(0.0, 20.0, 20001)
    v = np.array([stf(lib, 0, RICKER, x, 5.0, 0.5) for x in t])
    assert abs(np.trapezoid(v, t)) < 1e-8


def test_gauss_unit_area_and_derivative():
    lib = Lib(64)
    a, t0 = 0.4, 2.0
    t = np.linspace(0.0, 4.0, 40001)
    v = np.array([stf(lib, 0, GAUSS, x, t0, a) for x in t])
    assert np.trapezoid(v, t) == pytest.approx(1.0, rel=1e-6)
    for x in (1.2, 2.0, 2.7):
        h = 1e-6
        num = (stf(lib, 0, GAUSS, x + h, t0, a) - stf(lib, 0, GAUSS, x - h, t0, a)) / (2 * h)
        assert stf(lib, 1, GAUSS, x, t0, a) == pytest.approx(num, rel=1e-6, abs=1e-9)


def test_bell_and_integral():
    lib = Lib(64)
    r = 1.5
    t = np.linspace(0.0, r, 30001)
    v = np.array([stf(lib, 0, BELL, x, 0.0, r) for x in t])
    assert np.trapezoid(v, t) == pytest.approx(1.0, rel=1e-7)
    for x in (0.2, 0.75, 1.3):
        run = np.trapezoid(v[t <= x], t[t <= x])
        assert stf(lib, 0, BELL_INT, x, 0.0, r) == pytest.approx(run, abs=1e-6)
    assert stf(lib, 0, BELL_INT, 2 * r, 0.0, r) == 1.0


def test_triangle_unit_area():
    lib = Lib(64)
    r = 2.0
    t = np.linspace(0.0, r, 20001)
    v = np.array([stf(lib, 1, TRI, x, 0.0, r) for x in t])
    assert np.trapezoid(v, t) == pytest.approx(1.0, rel=1e-6)


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.45)])
def test_smoothing_weights(shift):
    lib = Lib(64)
    d = np.zeros(343)
    lib.norm_delt(d.ctypes.data_as(C.POINTER(C.c_double)), *shift)
    w = d.reshape(7, 7, 7)  # [k][j][i]
    assert w.sum() == pytest.approx(1.0, rel=1e-14)
    k = np.arange(-3, 4)
    g = [np.exp(-((k - s) ** 2) / 1.5 ** 2) for s in shift]
    ref = g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    np.testing.assert_allclose(w, ref / ref.sum(), rtol=1e-13)
    # centroid moves with the shift
    assert (w.sum(axis=(0, 1)) * k).sum() == pytest.approx(shift[0], abs=0.05)
