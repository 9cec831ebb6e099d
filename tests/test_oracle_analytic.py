"""Whole-solver pins of the oracle against closed-form solutions (CPU only).

The reference ships no outputs, so the coupled step of the oracle (the 8
F/B patterns, RK4 stages with their source times, the smoothed point
sources and the spherical geometry, all interleaved as
code/srcF/seis3d_wave_8op_p3.F90:97-157 runs them) is compared with
Stokes' solution for a point force and for a point moment tensor in a
homogeneous, isotropic full space (Aki & Richards, Quantitative
Seismology, 2nd ed., eqs. 4.23 and 4.29; tests/analytic.py), near-,
intermediate- and far-field terms included.  The source of the scheme is
the reference's 7^3 Gaussian spread (mod_src.F90:228-232, 419-437) of a
Ricker force history (mod_src.F90:509-530) divided by cell volume and
density (src_force, mod_src.F90:275-345) or cell volume (src_stress,
mod_src.F90:207-273); the analytic response is the same weighted sum of
point responses, evaluated at the grid nodes' Cartesian positions on the
sphere and rotated into the local (theta, phi, r) components.  Receivers
record velocity, so the force history enters as its time derivative.

What the pins fix, and the controls that show they would catch a mistake:
  - the source time of every RK stage (ntime + tinc) dt: the fitted arrival
    shift is within 0.1 dt; a quarter-step error in tinc moves it by 0.25 dt
    (control);
  - the force amplitude through the smoothing weights, cell volume and
    density: within 1%; a doubled body force fails (control);
  - the moment-tensor path with its sign convention (hT -= M rate w / V);
  - the propagation of the coupled scheme: relative L2 misfit below 1.5e-2
    at 6-9 km (measured: force 1.3e-3 .. 3.2e-3, moment 1.9e-3 .. 1.0e-2,
    shifts within 0.025 dt, amplitudes within 0.25%; the scheme's own
    dispersion and the CFS-PML's reflections set the floor).
"""
import ctypes as C

import numpy as np
import pytest

import analytic as an
from fwant_amd import models
from oracle.pyoracle import Lib, OracleSolver, mutation

VP, VS, RHO, H = 6000.0, 3464.0, 2700.0, 1000.0
N, DT, STEPS, F0 = 48, 0.08, 240, 0.15
TC = 1.2 / F0
OFFS = [(9, 0, 0), (0, -9, 0), (0, 0, 9), (0, 0, -8), (6, 6, 0), (-5, 5, 5), (4, -6, -5)]


def homogeneous():
    x, y, z = models.spherical_grid(N, N, N, H)
    shape = (N + 6, N + 6, N + 6)
    lam, mu, rho = models.lame(np.full(shape, VP), np.full(shape, VS), np.full(shape, RHO))
    return models.Model("homogeneous", N, N, N, DT, STEPS, False, x, y, z, np.asfortranarray(lam),
                        np.asfortranarray(mu), np.asfortranarray(rho), abs_number=(10,) * 6,
                        abs_velocity=(VP,) * 6, cfs_bmax=(3.0,) * 6, cfs_amax=(np.pi * F0,) * 6)


def source(comps):
    s = models.Sources(n=1, ntw=1, stf_id=models.STF_RICKER)
    c = 3 + N // 2
    s.indx = np.array([[c + 1, c + 1, c + 1]], np.int32)  # 1-based, ghost offset included
    s.shift = np.zeros((1, 3), np.float32)
    s.t0 = np.zeros(1, np.float32)
    s.stf_time = np.array([TC], np.float32)
    s.stf_freq = np.array([F0], np.float32)
    s.comps = np.asarray(comps, np.float32).reshape(-1, 1)
    return s


def cartesian(m, i, j, k):
    """Position and local basis (columns theta, phi, r) of grid node (i, j, k)."""
    th, ph, r = m.x[i], m.y[j], m.z[k]
    er = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    et = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
    ep = np.array([-np.sin(ph), np.cos(ph), 0.0])
    return r * er, np.stack([et, ep, er], 1)


def simulate(comps, moment=False, mut=0):
    m = homogeneous()
    c = 3 + N // 2
    if moment:
        m.moment = source(comps)
    else:
        m.force = source(comps)
    recs = [(c + a, c + b, c + d) for a, b, d in OFFS]
    m.rec_indx = (np.array(recs) + 1).astype(np.int32)
    o = OracleSolver(m, 32, nthreads=8)
    with mutation(mut):
        rec = o.run(STEPS).astype(np.float64)
    o.close()
    return m, recs, rec[:, :, :3]


def stokes(m, recs, comps, moment):
    """Analytic local-component velocities [steps, nrec, 3] for the smoothed source."""
    c = 3 + N // 2
    w = np.zeros(343)
    Lib(64).norm_delt(w.ctypes.data_as(C.POINTER(C.c_double)), 0.0, 0.0, 0.0)
    off = [(i, j, k) for k in range(-3, 4) for j in range(-3, 4) for i in range(-3, 4)]
    t = (np.arange(STEPS) + 1) * DT
    ps, Rs = cartesian(m, c, c, c)
    g1 = lambda tt: an.ricker(tt, F0, TC, 1)
    g2 = lambda tt: an.ricker(tt, F0, TC, 2)
    if moment:
        M6 = np.asarray(comps, np.float64)
        Ml = np.array([[M6[0], M6[3], M6[4]], [M6[3], M6[1], M6[5]], [M6[4], M6[5], M6[2]]])
        Mc = Rs @ Ml @ Rs.T
    else:
        Fc = Rs @ np.asarray(comps, np.float64)
    out = np.zeros((STEPS, len(recs), 3))
    for q, rc in enumerate(recs):
        pr, Rr = cartesian(m, *rc)
        v = np.zeros((STEPS, 3))
        for wq, o3 in zip(w, off):
            pq, _ = cartesian(m, c + o3[0], c + o3[1], c + o3[2])
            if moment:
                v += wq * an.stokes_moment(pr - pq, Mc, t, VP, VS, RHO, g1, g2)
            else:
                v += wq * an.stokes_force(pr - pq, Fc, t, VP, VS, RHO, g1)
        out[:, q] = v @ Rr
    return out


def fit(sim, ref):
    """rel-L2 misfit, arrival shift (in steps, parabolic cross-correlation
    peak) and amplitude ratio of one receiver's three components."""
    err = np.linalg.norm(sim - ref) / np.linalg.norm(ref)
    cc = np.array([np.sum(np.roll(sim, -s, 0)[10:-10] * ref[10:-10]) for s in range(-4, 5)])
    k = int(np.argmax(cc))
    k = min(max(k, 1), len(cc) - 2)
    a, b, c = cc[k - 1], cc[k], cc[k + 1]
    shift = (k - 4) + 0.5 * (a - c) / (a - 2 * b + c)
    return err, shift, np.linalg.norm(sim) / np.linalg.norm(ref)


FORCE = (0.3e15, -0.5e15, 1.0e15)
MOMENT = (1.0e16, -0.4e16, -0.6e16, 0.5e16, -0.3e16, 0.8e16)


@pytest.fixture(scope="module")
def force_case():
    m, recs, rec = simulate(FORCE)
    return m, recs, rec, stokes(m, recs, FORCE, False)


def test_point_force_matches_stokes(force_case):
    m, recs, rec, ref = force_case
    for q in range(len(recs)):
        err, shift, amp = fit(rec[:, q], ref[:, q])
        assert err < 1.5e-2, (OFFS[q], err)
        assert abs(shift) < 0.1, (OFFS[q], shift)
        assert abs(amp - 1.0) < 1e-2, (OFFS[q], amp)


def test_moment_tensor_matches_stokes():
    m, recs, rec = simulate(MOMENT, moment=True)
    ref = stokes(m, recs, MOMENT, True)
    for q in range(len(recs)):
        err, shift, amp = fit(rec[:, q], ref[:, q])
        assert err < 1.5e-2, (OFFS[q], err)
        assert abs(shift) < 0.1, (OFFS[q], shift)
        assert abs(amp - 1.0) < 1e-2, (OFFS[q], amp)


@pytest.mark.parametrize("mut,what", [(5, "source time a quarter step late at every stage"),
                                       (7, "body force doubled")])
def test_planted_mistakes_fail_the_stokes_pin(force_case, mut, what):
    """Controls: each planted mistake in the oracle breaks the pin above."""
    m, recs, _, ref = force_case
    _, _, rec = simulate(FORCE, mut=mut)
    bad = []
    for q in range(len(recs)):
        err, shift, amp = fit(rec[:, q], ref[:, q])
        bad.append(err >= 1.5e-2 or abs(shift) >= 0.1 or abs(amp - 1.0) >= 1e-2)
    assert all(bad), what
