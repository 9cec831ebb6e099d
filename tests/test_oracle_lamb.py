"""Pin of the oracle's free-surface response to a surface force against
Lamb's problem (a vertical point force on a homogeneous half-space; the
frequency-wavenumber solution of tests/analytic.py, which reproduces
Stokes' solution for a buried force).

The reference's surface force is not a point load of F alone: with its
build flags (SrcSmooth, SrcSurface) src_surface applies the whole F as
the surface traction F/S at the source node (code/srcF/mod_src.F90:
400-409), and src_force applies the 7^3 Gaussian spread renormalised over
the planes at and below the surface (:322-324) as body force on every
plane but the surface one (:325-331) -- the planes 1-3 cells below carry
45.3% of F more.  The analytic response is therefore Lamb's solution for
F at the surface plus the half-space responses of the buried forces with
the reference's weights at their nodes' distances.  On a 0.15 Hz Ricker
(19 cells per Rayleigh wavelength) the oracle's vertical surface velocity
at 10 and 16 km matches that sum within 0.12 relative L2 and 5% in
amplitude (measured 6.5e-2 .. 8.8e-2 and 0.997 .. 1.033 with 8-20 km
receivers; 0.40-0.44 at 0.25 Hz, so the misfit converges with
resolution); Lamb's solution for F alone is off by 40%+ in amplitude,
and a planted mistake (the surface traction doubled) fails the pin.
"""
import ctypes as C

import numpy as np

import analytic as an
from fwant_amd import models
from oracle.pyoracle import Lib, OracleSolver, mutation

VP, VS, RHO, H = 6000.0, 3464.0, 2700.0, 1000.0
NI, NK, F0, DT, STEPS = 80, 40, 0.15, 0.08, 300
TC = 1.2 / F0
FZ = 1e16
DISTS = (10, 16)


def halfspace_run(mut=0):
    x, y, z = models.spherical_grid(NI, NI, NK, H)
    shape = (NI + 6, NI + 6, NK + 6)
    lam, mu, rho = models.lame(np.full(shape, VP), np.full(shape, VS), np.full(shape, RHO))
    m = models.Model("halfspace", NI, NI, NK, DT, STEPS, True, x, y, z, np.asfortranarray(lam),
                     np.asfortranarray(mu), np.asfortranarray(rho), abs_number=(10, 10, 10, 10, 10, 0),
                     abs_velocity=(VP,) * 6, cfs_bmax=(3.0,) * 6, cfs_amax=(np.pi * F0,) * 6)
    ic = 3 + NI // 2
    m.force = models.surface_force(NI, NI, NK, ic, ic, fz=FZ, f0=F0, t0=TC)
    m.rec_indx = np.array([[ic + d + 1, ic + 1, m.nk2 + 1] for d in DISTS], np.int32)
    o = OracleSolver(m, 32)
    with mutation(mut):
        rec = o.run(STEPS).astype(np.float64)
    o.close()
    return rec[:, :, 2]  # Vz (radial, up)


def lamb_responses(t):
    """(F alone at the surface, the buried planes' share): vertical velocity
    at the surface for an upward force F (= the downward-force response in
    depth-positive u_z, by symmetry)."""
    w = np.zeros(343)
    Lib(64).norm_delt(w.ctypes.data_as(C.POINTER(C.c_double)), 0.0, 0.0, 0.0)
    w = w.reshape(7, 7, 7)[:4]  # planes 3, 2, 1, 0 cells below the surface
    w = w / w.sum()
    g = lambda tt: FZ * an.ricker(tt, F0, TC, 1)  # noqa: E731  force rate -> velocity
    fmax = 3.0 * F0
    kw = dict(T=t[-1] + 20.0, fmax=fmax, kmax=2 * np.pi * fmax / VS * 2.5, nk=3000)
    rs = np.array([d * H for d in DISTS], float)
    surf, _ = an.fk_time_series(lambda k, ww: an.fk_vertical_force(k, ww, 0.0, 0.0, VP, VS, RHO), rs, t, g,
                                static=an.surface_static(VP, VS, RHO), **kw)
    body = np.zeros_like(surf)
    for kk in (1, 2, 3):
        W = w[3 - kk]
        rr, ws = [], []
        for n, d in enumerate(DISTS):
            for lj in range(-3, 4):
                for li in range(-3, 4):
                    rr.append(np.hypot((d - li) * H, lj * H))
                    ws.append((n, W[lj + 3, li + 3]))
        uk, _ = an.fk_time_series(lambda k, ww, dd=kk * H: an.fk_vertical_force(k, ww, dd, 0.0, VP, VS, RHO),
                                  np.array(rr), t, g, **kw)
        for q, (n, wt) in enumerate(ws):
            body[n] += wt * uk[q]
    return surf, body, w[:3].sum()


def fit(v, ref):
    return np.linalg.norm(v - ref) / np.linalg.norm(ref), float(np.dot(v, ref) / np.dot(ref, ref))


def test_surface_force_matches_lamb_with_the_buried_share():
    t = (np.arange(STEPS) + 1) * DT
    surf, body, share = lamb_responses(t)
    assert abs(share - 0.453) < 0.002
    v = halfspace_run()
    for n, d in enumerate(DISTS):
        e, a = fit(v[:, n], surf[n] + body[n])
        e0, a0 = fit(v[:, n], surf[n])
        print(f"{d} km: Lamb + buried share rel-L2 {e:.3e}, amplitude {a:.4f}; Lamb alone {e0:.3e}, {a0:.4f}")
        assert e < 0.12 and abs(a - 1.0) < 0.05
        assert a0 > 1.3
    bad = halfspace_run(mut=9)  # surface traction doubled
    e, a = fit(bad[:, 0], surf[0] + body[0])
    print(f"planted mistake (surface traction doubled): rel-L2 {e:.3e}, amplitude {a:.4f}")
    assert e > 0.3 and a > 1.3
