"""What one stage's source terms inject (CPU only, FP64 oracle).

The reference spreads a point force over 7^3 cells with normalised Gaussian
weights and divides by cell volume and density (src_force,
code/srcF/mod_src.F90:275-345); near a free surface it renormalises the
weights over the planes at and below the surface and leaves the surface
plane out (:296-299, :325-331); a force on the surface becomes a traction
F/S on one cell and its rate (src_surface, :373-414); a moment tensor is
spread the same way, divided by cell volume (src_stress, :207-273).  These
tests check what the injected terms add up to against geometry computed
independently of the scheme (the exact volume of a spherical cell and the
exact area of a cell on the sphere, which the reference's centred
approximations match to (h/r)^2 ~ 1e-8) and the Ricker closed form
(tests/analytic.py):
  - a buried force: sum of rho V h_V equals F R(t) (1e-6);
  - a force one cell below the surface: nothing on the surface plane, the
    spread below it is the buried spread renormalised (1e-7), so the total
    is F R(t) times the renormalisation;
  - a surface force: traction times cell area equals F R(t), its rate
    F R'(t) (1e-6); the spread below as above;
  - a moment tensor: sum of V h_T equals -M R'(t) (1e-6).
Controls: the body force doubled, the surface area halved, the surface
plane kept, the renormalisation left out (mistakes planted in the oracle)
each break one of them.
"""
import numpy as np
import pytest

import analytic as an
from fwant_amd import models
from oracle.pyoracle import OracleSolver, mutation

NI, NJ, NK, H = 20, 18, 16, 1000.0
F0, TC, DT = 0.5, 2.4, 0.04
NTIME, TINC = 50, 0.5


def model(force=None, moment=None):
    m = models.small_test(ni=NI, nj=NJ, nk=NK, nt=10, h=H, pml=0, dt=DT)
    m.force = force if force is not None else models.Sources(n=0, ntw=1)
    m.moment = moment if moment is not None else models.Sources(n=0, ntw=1)
    return m


def point(k, comps):
    s = models.Sources(n=1, ntw=1, stf_id=models.STF_RICKER)
    s.indx = np.array([[3 + NI // 2 + 1, 3 + NJ // 2 + 1, k + 1]], np.int32)  # 1-based
    s.shift = np.zeros((1, 3), np.float32)
    s.t0 = np.zeros(1, np.float32)
    s.stf_time = np.array([TC], np.float32)
    s.stf_freq = np.array([F0], np.float32)
    s.comps = np.asarray(comps, np.float64).reshape(-1, 1)
    return s


def edges(v):
    """Cell faces of a node coordinate vector (midpoints, ends mirrored)."""
    v = np.asarray(v, np.float64)
    e = np.empty(len(v) + 1)
    e[1:-1] = 0.5 * (v[1:] + v[:-1])
    e[0], e[-1] = v[0] - 0.5 * (v[1] - v[0]), v[-1] + 0.5 * (v[-1] - v[-2])
    return e


def exact_volume(m):
    """Exact volume of every cell [theta-, theta+] x [phi-, phi+] x [r-, r+]."""
    tx, ty, tz = edges(m.x), edges(m.y), edges(m.z)
    dcos = np.cos(tx[:-1]) - np.cos(tx[1:])
    dphi = np.diff(ty)
    dr3 = (tz[1:] ** 3 - tz[:-1] ** 3) / 3.0
    return dcos[:, None, None] * dphi[None, :, None] * dr3[None, None, :]


def exact_area(m, i, j, k):
    tx, ty = edges(m.x), edges(m.y)
    return m.z[k] ** 2 * (np.cos(tx[i]) - np.cos(tx[i + 1])) * (ty[j + 1] - ty[j])


def stage(m, mut=0):
    o = OracleSolver(m, 64, nthreads=2)
    with mutation(mut):
        h, surf = o.src_terms(NTIME, TINC)
    o.close()
    return h, surf


T = (NTIME + TINC) * DT
TC32 = float(np.float32(TC))  # the source tables hold float32 times
R, R1 = an.ricker(T, F0, TC32, 0), an.ricker(T, F0, TC32, 1)
FORCE = (2.0e15, -1.0e15, 3.0e15)


def injected_force(m, h):
    rho = np.asarray(m.rho, np.float64)
    V = exact_volume(m)
    return np.array([(h[f] * rho * V).sum() for f in range(3)])


def test_buried_force_injects_F_times_stf():
    m = model(force=point(3 + NK // 2 - 2, FORCE))
    h, surf = stage(m)
    np.testing.assert_allclose(injected_force(m, h), np.array(FORCE) * R, rtol=1e-6)
    assert not surf.any() and not h[3:].any()


def renormalised(m, h_buried, shift_k):
    """The buried spread moved up by shift_k planes, cut at the surface and renormalised."""
    rho, V = np.asarray(m.rho, np.float64), exact_volume(m)
    w = h_buried * rho * V  # per-cell share of F R(t) (component 2)
    w = np.roll(w, shift_k, axis=2)
    k2 = m.nk2
    keep = w[:, :, :k2 + 1].sum()
    w[:, :, k2:] = 0.0  # the surface plane and the planes above it get nothing
    return w / keep


def test_force_below_surface_is_cut_and_renormalised():
    kb = 3 + NK // 2 - 2
    mb = model(force=point(kb, (0.0, 0.0, 1.0e15)))
    hb, _ = stage(mb)
    m1 = model(force=point(mb.nk2 - 1, (0.0, 0.0, 1.0e15)))
    h1, _ = stage(m1)
    rho, V = np.asarray(m1.rho, np.float64), exact_volume(m1)
    assert not h1[:, :, :, m1.nk2:].any()
    got = h1[2] * rho * V / (1.0e15 * R)
    want = renormalised(mb, hb[2] / (1.0e15 * R), (m1.nk2 - 1) - kb)
    np.testing.assert_allclose(got, want, rtol=1e-7, atol=1e-12)
    # renormalised over the planes at and below the surface, then the surface
    # plane left out: the total is the share of the planes below it
    assert 0.5 < got.sum() < 1.0


def test_surface_force_traction_and_rate():
    m = model(force=point(3 + NK - 1, FORCE))
    h, surf = stage(m)
    i, j = 3 + NI // 2, 3 + NJ // 2
    S = exact_area(m, i, j, m.nk2)
    np.testing.assert_allclose(surf[:3, i, j] * S, np.array(FORCE) * R, rtol=1e-6)
    np.testing.assert_allclose(surf[3:, i, j] * S, np.array(FORCE) * R1, rtol=1e-6)
    surf[:, i, j] = 0.0
    assert not surf.any()
    assert not h[:, :, :, m.nk2:].any()
    # below the surface: the spread of the same force, cut and renormalised
    share = injected_force(m, h) / (np.array(FORCE) * R)
    np.testing.assert_allclose(share, share[0], rtol=1e-9)
    assert 0.3 < share[0] < 0.9


def test_moment_injects_minus_M_rate():
    M6 = (1.0e16, -2.0e16, 0.5e16, 0.7e16, -0.3e16, 1.1e16)
    m = model(moment=point(3 + NK // 2 - 2, M6))
    h, surf = stage(m)
    V = exact_volume(m)
    got = np.array([(h[3 + q] * V).sum() for q in range(6)])
    np.testing.assert_allclose(got, -np.array(M6) * R1, rtol=1e-6)
    assert not h[:3].any()


def test_planted_mistakes_fail():
    mb = model(force=point(3 + NK // 2 - 2, FORCE))
    h, _ = stage(mb, mut=7)  # body force doubled
    assert not np.allclose(injected_force(mb, h), np.array(FORCE) * R, rtol=1e-6)
    ms = model(force=point(3 + NK - 1, FORCE))
    i, j = 3 + NI // 2, 3 + NJ // 2
    _, surf = stage(ms, mut=9)  # surface cell area halved
    assert not np.allclose(surf[:3, i, j] * exact_area(ms, i, j, ms.nk2), np.array(FORCE) * R, rtol=1e-6)
    h, _ = stage(ms, mut=6)  # body force also on the surface plane
    assert h[:, :, :, ms.nk2].any()
    m1 = model(force=point(ms.nk2 - 1, (0.0, 0.0, 1.0e15)))
    h1, _ = stage(m1, mut=8)  # no renormalisation near the surface
    assert injected_force(m1, h1)[2] / (1.0e15 * R) < 0.99
