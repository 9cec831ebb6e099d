"""Whole-solver pins of the oracle against physics the scheme must reproduce
(the reference ships no outputs to compare with):

  - reciprocity of the elastic Green's function, G_pq(B, A) = G_qp(A, B):
    in a heterogeneous medium with a free surface and CFS-PML, the velocity
    component q at B due to a force along p at A equals component p at A due
    to the same force along q at B.  The body force is spread over 7^3 cells
    with normalised Gaussian weights w (mod_src.F90:228-232, 419-437), so the
    reciprocal receiver is the same w-weighted sum of the velocity.  A
    coupling term with a wrong sign or index breaks the symmetry.
  - Rayleigh waves along the free surface of a Poisson half-space
    (Vp = sqrt(3) Vs) travel at c_R = 0.9194 Vs (root of the Rayleigh
    equation); the free-surface imaging and traction condition
    (code/srcF/mod_macdrp.F90:3414-3436, :942-950) decide this speed.
"""
import ctypes as C

import numpy as np
import pytest

from fwant_amd import models
from oracle.pyoracle import Lib, OracleSolver


def point_force(i, j, k, comp, f0=0.4, amp=1e15):
    s = models.Sources(n=1, ntw=1, stf_id=models.STF_RICKER)
    s.indx = np.array([[i + 1, j + 1, k + 1]], np.int32)
    s.shift = np.zeros((1, 3), np.float32)
    s.t0 = np.zeros(1, np.float32)
    s.stf_time = np.array([1.2 / f0], np.float32)
    s.stf_freq = np.array([f0], np.float32)
    c = np.zeros((3, 1), np.float32)
    c[comp, 0] = amp
    s.comps = c
    return s


RECIP = dict(ni=44, nj=40, nk=34, seed=21, moment=False, h=200.0, colat0=1.0, r_top=20000.0, dt=0.005)


def test_reciprocity_of_point_forces():
    """On a curved grid (r = 13.4..20 km, colatitude ~57 deg, wavelength
    ~0.2 r), where the 1/r and cot(theta) terms matter."""
    steps = 600
    nk2 = 3 + RECIP["nk"] - 1
    A = (3 + 14, 3 + 15, nk2 - 5)
    B = (3 + 29, 3 + 24, nk2 - 9)
    w = np.zeros(343)
    Lib(64).norm_delt(w.ctypes.data_as(C.POINTER(C.c_double)), 0.0, 0.0, 0.0)
    off = np.array([(i, j, k) for k in range(-3, 4) for j in range(-3, 4) for i in range(-3, 4)])
    pairs = [(0, 2), (1, 0), (2, 2)]
    for p, q in pairs:
        out = []
        for src, rec, comp_s, comp_r in ((A, B, p, q), (B, A, q, p)):
            m = models.small_test(nt=steps, **RECIP)
            m.moment = models.Sources(comps=np.zeros((6, 0), np.float32))
            m.force = point_force(*src, comp_s, f0=1.0)
            m.rec_indx = (np.array(rec)[None, :] + off + 1).astype(np.int32)
            o = OracleSolver(m, 32)
            out.append(o.run(steps)[:, :, comp_r].astype(np.float64) @ w)
            o.close()
        e = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])
        print(f"force {p} at A -> V{q} at B vs force {q} at B -> V{p} at A: rel diff {e:.2e}")
        assert e < 0.03, (p, q, e)


def test_rayleigh_phase_speed():
    ni, nj, nk, h = 120, 40, 40, 1000.0
    steps, dt, f0 = 420, 0.08, 0.25
    vs, vp, rho = 3000.0, 3000.0 * np.sqrt(3.0), 2700.0
    x, y, z = models.spherical_grid(ni, nj, nk, h)
    shape = (ni + 6, nj + 6, nk + 6)
    lam, mu, r = models.lame(np.full(shape, vp), np.full(shape, vs), np.full(shape, rho))
    m = models.Model("halfspace", ni, nj, nk, dt, steps, True, x, y, z,
                     np.asfortranarray(lam), np.asfortranarray(mu), np.asfortranarray(r),
                     abs_number=(10, 10, 10, 10, 10, 0), abs_velocity=(vp,) * 6,
                     cfs_bmax=(1.0,) * 6, cfs_amax=(np.pi * f0,) * 6)
    i0, j0 = 3 + 15, 3 + nj // 2
    m.force = models.surface_force(ni, nj, nk, i0, j0, fz=1e16, f0=f0)
    d1, d2 = 25, 75  # receiver offsets along x in cells (km)
    m.rec_indx = np.array([[i0 + d + 1, j0 + 1, m.nk2 + 1] for d in (d1, d2)], np.int32)
    o = OracleSolver(m, 32)
    rec = o.run(steps).astype(np.float64)
    o.close()
    cR = 0.9194 * vs
    t = (np.arange(steps) + 1) * dt
    t0 = 1.2 / f0

    def window(tr, d):
        ta = t0 + d * h / cR
        return tr * np.exp(-((t - ta) / 3.0) ** 8)

    a = window(rec[:, 0, 2], d1)
    b = window(rec[:, 1, 2], d2)
    # lag of b behind a from the cross-correlation peak with parabolic refinement
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    frac = 0.5 * (cc[k - 1] - cc[k + 1]) / (cc[k - 1] - 2 * cc[k] + cc[k + 1])
    lag = (k - (steps - 1) + frac) * dt
    c = (d2 - d1) * h / lag
    print(f"Rayleigh phase speed {c:.1f} m/s, theory {cR:.1f} m/s (Vs {vs:.0f})")
    assert c == pytest.approx(cR, rel=0.005)
