"""Pin of the oracle's operator sweep (code/srcF/mod_macdrp.F90:790-962)
against the continuous velocity-stress equations in spherical coordinates
(x = colatitude theta, y = longitude phi, z = radius r; Aki & Richards,
spherical strain and equation of motion):

  rho dv_t/dt = (1/r) dS_tt/dt + 1/(r sin) dS_tp/dp + dS_tr/dr + [3 S_tr + (S_tt - S_pp) cot] / r
  rho dv_p/dt = (1/r) dS_tp/dt + 1/(r sin) dS_pp/dp + dS_pr/dr + [2 S_tp cot + 3 S_pr] / r
  rho dv_r/dt = (1/r) dS_tr/dt + 1/(r sin) dS_pr/dp + dS_rr/dr + [2 S_rr - S_tt - S_pp + S_tr cot] / r
  e_tt = (dv_t/dt + v_r)/r,  e_pp = (v_t cot + dv_p/dp / sin + v_r)/r,  e_rr = dv_r/dr
  e_tp = (dv_p/dt + dv_t/dp / sin - v_p cot)/(2r)
  e_tr = (dv_r/dt / r + dv_t/dr - v_t/r)/2,  e_pr = (dv_r/dp /(r sin) + dv_p/dr - v_p/r)/2
  dS/dt = lambda tr(e) I + 2 mu e

Smooth analytic fields (wavelength ~ 25 cells) on a small, strongly curved
grid (r = 5..10 km) make the geometric terms comparable to the derivative
terms, so a dropped or mis-signed term, a swapped index or a wrong F/B
direction moves the result far outside the 2% tolerance that covers the
one-sided DRP truncation error.
"""
import numpy as np
import pytest
import sympy as sp

from fwant_amd import models
from oracle.pyoracle import OracleSolver


def analytic_rhs():
    th, ph, r = sp.symbols("th ph r")
    k = 2 * sp.pi / 5000.0  # 5 km wavelength in r; angles scaled to similar arc length
    a = 2 * sp.pi / 0.55
    F = {
        "Vx": sp.sin(a * th + 0.3) * sp.cos(0.7 * a * ph) * sp.cos(k * r),
        "Vy": sp.cos(0.8 * a * th) * sp.sin(a * ph + 0.2) * sp.sin(k * r + 0.4),
        "Vz": sp.sin(0.6 * a * th + 0.1) * sp.sin(0.9 * a * ph) * sp.cos(1.3 * k * r),
        "Txx": 1e6 * sp.cos(a * th) * sp.cos(a * ph + 0.5) * sp.sin(k * r + 0.3),
        "Tyy": 1e6 * sp.sin(0.7 * a * th + 0.2) * sp.cos(0.6 * a * ph) * sp.cos(0.9 * k * r),
        "Tzz": 1e6 * sp.cos(0.9 * a * th + 0.4) * sp.sin(a * ph) * sp.sin(1.1 * k * r),
        "Txy": 1e6 * sp.sin(a * th) * sp.sin(0.8 * a * ph + 0.1) * sp.cos(k * r + 0.2),
        "Txz": 1e6 * sp.cos(0.6 * a * th) * sp.cos(a * ph + 0.3) * sp.sin(0.8 * k * r + 0.1),
        "Tyz": 1e6 * sp.sin(0.9 * a * th + 0.5) * sp.cos(0.7 * a * ph) * sp.cos(1.2 * k * r + 0.6),
    }
    lam, mu, rho = 3.0e10, 2.5e10, 2700.0
    s, c = sp.sin(th), sp.cot(th)
    D = sp.diff
    vt, vp, vr = F["Vx"], F["Vy"], F["Vz"]
    Stt, Spp, Srr, Stp, Str, Spr = F["Txx"], F["Tyy"], F["Tzz"], F["Txy"], F["Txz"], F["Tyz"]
    h = {}
    h["Vx"] = (D(Stt, th) / r + D(Stp, ph) / (r * s) + D(Str, r) + (3 * Str + (Stt - Spp) * c) / r) / rho
    h["Vy"] = (D(Stp, th) / r + D(Spp, ph) / (r * s) + D(Spr, r) + (2 * Stp * c + 3 * Spr) / r) / rho
    h["Vz"] = (D(Str, th) / r + D(Spr, ph) / (r * s) + D(Srr, r) + (2 * Srr - Stt - Spp + Str * c) / r) / rho
    ett = (D(vt, th) + vr) / r
    epp = (vt * c + D(vp, ph) / s + vr) / r
    err = D(vr, r)
    etp = (D(vp, th) + D(vt, ph) / s - vp * c) / (2 * r)
    etr = (D(vr, th) / r + D(vt, r) - vt / r) / 2
    epr = (D(vr, ph) / (r * s) + D(vp, r) - vp / r) / 2
    tr = ett + epp + err
    h["Txx"] = lam * tr + 2 * mu * ett
    h["Tyy"] = lam * tr + 2 * mu * epp
    h["Tzz"] = lam * tr + 2 * mu * err
    h["Txy"] = 2 * mu * etp
    h["Txz"] = 2 * mu * etr
    h["Tyz"] = 2 * mu * epr
    f = {n: sp.lambdify((th, ph, r), e, "numpy") for n, e in F.items()}
    g = {n: sp.lambdify((th, ph, r), e, "numpy") for n, e in h.items()}
    return f, g, (lam, mu, rho)


@pytest.fixture(scope="module")
def setup():
    f, g, (lam, mu, rho) = analytic_rhs()
    m = models.small_test(ni=28, nj=26, nk=26, nt=1, h=200.0, colat0=1.0, r_top=10000.0, pml=0,
                          freenode=False, moment=False)
    m.abs_number = (0,) * 6
    m.force = models.Sources()
    m.rec_indx = np.zeros((0, 3), np.int32)
    shape = (m.nx, m.ny, m.nz)
    m.lam = np.full(shape, lam, np.float32, order="F")
    m.mu = np.full(shape, mu, np.float32, order="F")
    m.rho = np.full(shape, rho, np.float32, order="F")
    TH, PH, R = np.meshgrid(m.x, m.y, m.z, indexing="ij")
    o = OracleSolver(m, 64, nthreads=2)
    from fwant_amd._lib import FIELDS
    for name in FIELDS:
        o.set_wavefield(name, f[name](TH, PH, R) * np.ones(shape))
    want = {n: g[n](TH, PH, R) * np.ones(shape) for n in FIELDS}
    return m, o, want


@pytest.mark.parametrize("pat", range(8))
def test_rhs_matches_spherical_equations(setup, pat):
    m, o, want = setup
    from fwant_amd._lib import FIELDS
    xf, yf, zf = (pat >> 2) & 1, (pat >> 1) & 1, pat & 1
    h = o.rhs(xf, yf, zf)
    sl = (slice(3, 3 + m.ni), slice(3, 3 + m.nj), slice(3, 3 + m.nk))
    for q, name in enumerate(FIELDS):
        got = h[q][sl]
        ref = want[name][sl]
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert e < 0.02, f"{name} pattern {pat}: rel err {e:.3e}"


def test_rhs_detects_a_dropped_geometric_term(setup):
    """The tolerance is tight enough to see a missing 3 S_tr / r in hVx."""
    m, o, want = setup
    TH, PH, R = np.meshgrid(m.x, m.y, m.z, indexing="ij")
    h = o.rhs(1, 1, 1)
    sl = (slice(3, 3 + m.ni), slice(3, 3 + m.nj), slice(3, 3 + m.nk))
    Txz = o.wavefield("Txz")
    wrong = want["Vx"] - 3.0 * Txz / R / 2700.0
    e_right = np.linalg.norm(h[0][sl] - want["Vx"][sl]) / np.linalg.norm(want["Vx"][sl])
    e_wrong = np.linalg.norm(h[0][sl] - wrong[sl]) / np.linalg.norm(wrong[sl])
    assert e_right < 0.02 < e_wrong
