"""Pins of the oracle's mode C: the spherical velocity-stress scheme on a
terrain-following grid under topography (theta = x(i), phi = y(j),
r = r(i, j, k)).  The reference's grid is a spherical shell with a 1-D
metric, so mode C has no reference output; it is pinned by mathematics:

  - flat grid: with r = z(k) (slopes p = q = 0) a mode C run is the
    reference's spherical scheme (op_sweep, free_timg, sources) up to
    rounding, with CFS-PML, a free surface and all source kinds;
  - metric: slopes and zeta_r of an analytic surface;
  - interior equations: the RHS of smooth analytic fields on a grid under
    steep topography (slopes up to 1) against the continuous spherical
    equations evaluated at the nodes' physical positions, as accurate as
    the spherical scheme on the flat grid; planted mistakes (the slope terms
    of the velocity update left out, the sign of p aZ in E13) fail it;
  - free surface: the stress rates at the surface satisfy the traction-rate
    condition sigma_t . N = S with N = (p, q, 1) (a surface force on), the
    imaging sets tau = sigma . N to the source traction at the surface and
    images it oddly above; a flipped slope in the surface solve fails;
  - Rayleigh waves along a surface sloping at 0.15 travel at the Rayleigh
    speed along the surface;
  - a long run under steep hills with CFS-PML decays once the sources stop.
"""
import numpy as np
import pytest

from fwant_amd import models
from fwant_amd._lib import FIELDS
from oracle.pyoracle import OracleSolver, build_metric3d, mutation

from test_oracle_rhs import analytic_rhs


def test_flat_grid_is_the_spherical_scheme():
    m = models.small_test(ni=40, nj=36, nk=28, nt=40)
    mf = models.with_topography(m, np.zeros((m.nx, m.ny)))
    a, b = OracleSolver(m, 64), OracleSolver(mf, 64)
    ra, rb = a.run(40), b.run(40)
    assert np.linalg.norm(ra - rb) <= 1e-13 * np.linalg.norm(ra)
    for f in FIELDS:
        wa, wb = a.wavefield(f), b.wavefield(f)
        assert np.abs(wa - wb).max() <= 1e-13 * np.abs(wa).max(), f
    a.close()
    b.close()


def _analytic_surface(m, amp):
    """heights amp sin(.) cos(.) over the grid, and their theta / phi derivatives"""
    X, Y = np.meshgrid(m.x, m.y, indexing="ij")
    ax = 2 * np.pi * 1.3 / (m.x[-1] - m.x[0])
    ay = 2 * np.pi * 1.1 / (m.y[-1] - m.y[0])
    u, v = ax * (X - m.x[0]) + 0.4, ay * (Y - m.y[0])
    return amp * np.sin(u) * np.cos(v), amp * ax * np.cos(u) * np.cos(v), -amp * ay * np.sin(u) * np.sin(v)


def test_metric_of_an_analytic_surface():
    m = models.small_test(ni=60, nj=56, nk=40, nt=1, h=200.0, colat0=1.0, r_top=10000.0, pml=0)
    topo, t_th, t_ph = _analytic_surface(m, 400.0)
    k_flat = 3
    m.r3 = models.terrain_radius(m.z, topo, k_flat, m.nk2)
    g = build_metric3d(m)
    k = np.arange(m.nz, dtype=np.float64)
    s = np.clip((k - k_flat) / float(m.nk2 - k_flat), 0.0, 1.0)
    w = s ** 3 * (10 - 15 * s + 6 * s * s)
    dw = 30 * s * s * (1 - s) ** 2 / float(m.nk2 - k_flat)  # dw/dk
    r = m.r3
    p = -t_th[:, :, None] * w[None, None, :] / r
    q = -t_ph[:, :, None] * w[None, None, :] / (r * np.sin(m.x)[:, None, None])
    zz = 1.0 / ((m.z[1] - m.z[0]) + topo[:, :, None] * dw[None, None, :])
    sl = (slice(3, 3 + m.ni), slice(3, 3 + m.nj), slice(6, 3 + m.nk - 3))
    for name, want in (("p", p), ("q", q), ("zz", zz)):
        e = np.abs(g[name][sl] - want[sl]).max() / np.abs(want[sl]).max()
        print(f"{name}: max rel err {e:.2e}")
        assert e < 1e-4, name
    assert np.abs(g["p"]).max() > 0.05 and np.abs(g["q"]).max() > 0.05
    # flat where the heights vanish: p = q = 0 exactly, zeta_r the 1-D metric
    m.r3 = models.terrain_radius(m.z, 0.0 * topo, k_flat, m.nk2)
    g0 = build_metric3d(m)
    from oracle.pyoracle import build_metric
    assert np.all(g0["p"] == 0.0) and np.all(g0["q"] == 0.0)
    assert np.all(g0["zz"] == build_metric(m)["zeta_z"][None, None, :])


@pytest.fixture(scope="module")
def analytic_setup():
    f, g, (lam, mu, rho) = analytic_rhs()
    out = {}
    for amp in (0.0, 600.0):
        m = models.small_test(ni=28, nj=26, nk=26, nt=1, h=200.0, colat0=1.0, r_top=10000.0, pml=0,
                              freenode=False, moment=False)
        m.abs_number = (0,) * 6
        m.force = models.Sources()
        m.rec_indx = np.zeros((0, 3), np.int32)
        shape = (m.nx, m.ny, m.nz)
        m.lam = np.full(shape, lam, np.float32, order="F")
        m.mu = np.full(shape, mu, np.float32, order="F")
        m.rho = np.full(shape, rho, np.float32, order="F")
        topo = _analytic_surface(m, amp)[0]
        m.r3 = models.terrain_radius(m.z, topo, 0, m.nk2)
        TH, PH, _ = np.meshgrid(m.x, m.y, m.z, indexing="ij")
        R = m.r3
        o = OracleSolver(m, 64, nthreads=2)
        for name in FIELDS:
            o.set_wavefield(name, f[name](TH, PH, R) * np.ones(shape))
        want = {n: g[n](TH, PH, R) * np.ones(shape) for n in FIELDS}
        out[amp] = (m, o, want)
    return out


def _rhs_err(m, o, want, pat):
    xf, yf, zf = (pat >> 2) & 1, (pat >> 1) & 1, pat & 1
    h = o.rhs(xf, yf, zf)
    sl = (slice(3, 3 + m.ni), slice(3, 3 + m.nj), slice(3, 3 + m.nk))
    return max(np.linalg.norm(h[q][sl] - want[n][sl]) / np.linalg.norm(want[n][sl]) for q, n in enumerate(FIELDS))


@pytest.mark.parametrize("pat", range(8))
def test_rhs_under_topography_matches_spherical_equations(analytic_setup, pat):
    """Slopes up to ~1: the RHS is as close to the continuous equations as
    the spherical scheme's on the flat grid (DRP truncation, ~0.4%)."""
    e_flat = _rhs_err(*analytic_setup[0.0], pat)
    e_hill = _rhs_err(*analytic_setup[600.0], pat)
    print(f"pattern {pat}: flat {e_flat:.3e}, hills {e_hill:.3e}")
    assert e_hill < 1.3 * e_flat and e_hill < 0.01


@pytest.mark.parametrize("mut", [11, 12])
def test_rhs_under_topography_detects_planted_mistakes(analytic_setup, mut):
    m, o, want = analytic_setup[600.0]
    with mutation(mut):
        e = _rhs_err(m, o, want, 7)
    print(f"mutation {mut}: rel err {e:.3e}")
    assert e > 0.1


def _surface_state(seed=3):
    m = models.hills_test(ni=56, nj=52, nt=8, seed=seed)
    o = OracleSolver(m, 64, nthreads=4)
    from fwant_amd.models import _correlated_field
    shape = (m.nx, m.ny, m.nz)
    for q, f in enumerate(FIELDS):
        o.set_wavefield(f, _correlated_field(shape, 4.0, seed + q, coarse=1) * (1.0 if q < 3 else 1e6))
    return m, o


def _traction_residual(m, o):
    met = build_metric3d(m)
    k = m.nk2
    h, surf = o.src_terms(3, 0.5)  # fills the surface traction planes of the force sources
    hh = o.rhs(1, 0, 1)
    p, q = met["p"][:, :, k], met["q"][:, :, k]
    Txx, Tyy, Tzz, Txy, Txz, Tyz = (hh[FIELDS.index(n)][:, :, k] for n in ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"))
    S = surf[3:6]
    res = np.stack([p * Txx + q * Txy + Txz - S[0], p * Txy + q * Tyy + Tyz - S[1], p * Txz + q * Tyz + Tzz - S[2]])
    sl = (slice(None), slice(3, 3 + m.ni), slice(3, 3 + m.nj))
    scale = max(np.abs(Txx[sl[1:]]).max(), np.abs(S).max())
    return np.abs(res[sl]).max() / scale, np.abs(S).max() / scale, np.abs(p[sl[1:]]).max()


def test_surface_stress_rates_meet_the_traction_condition():
    m, o = _surface_state()
    r, s_rel, pmax = _traction_residual(m, o)
    print(f"traction-rate residual {r:.2e} (source {s_rel:.2e}, max slope {pmax:.2f})")
    assert pmax > 0.1 and s_rel > 1e-6
    assert r < 1e-10
    with mutation(10):
        m2, o2 = _surface_state()
        r2, _, _ = _traction_residual(m2, o2)
    print(f"planted mistake: residual {r2:.2e}")
    assert r2 > 1e-3


def test_surface_imaging_of_the_traction():
    m, o = _surface_state(5)
    met = build_metric3d(m)
    _, surf = o.src_terms(2, 0.0)
    o.rhs(0, 0, 0)  # images the state in place
    W = {f: o.wavefield(f) for f in FIELDS}
    k = m.nk2

    def tau(kk):
        p, q = met["p"][:, :, kk], met["q"][:, :, kk]
        return np.stack([W["Txz"][:, :, kk] + p * W["Txx"][:, :, kk] + q * W["Txy"][:, :, kk],
                         W["Tyz"][:, :, kk] + p * W["Txy"][:, :, kk] + q * W["Tyy"][:, :, kk],
                         W["Tzz"][:, :, kk] + p * W["Txz"][:, :, kk] + q * W["Tyz"][:, :, kk]])
    sl = (slice(None), slice(3, 3 + m.ni), slice(3, 3 + m.nj))
    T = surf[0:3]
    scale = np.abs(W["Txx"][:, :, k]).max()
    assert np.abs((tau(k) - T)[sl]).max() < 1e-10 * scale
    for mm in (1, 2, 3):
        img = np.stack([W["Txz"][:, :, k + mm], W["Tyz"][:, :, k + mm], W["Tzz"][:, :, k + mm]])
        assert np.abs((img - (2 * T - tau(k - mm)))[sl]).max() < 1e-10 * scale


def _rayleigh_run(slope, steps=560, dt=0.06, f0=0.25, mut=0, nthreads=None):
    ni, nj, nk, h = 120, 40, 48, 1000.0
    vs, vp, rho = 3000.0, 3000.0 * np.sqrt(3.0), 2700.0
    x, y, z = models.spherical_grid(ni, nj, nk, h)
    shape = (ni + 6, nj + 6, nk + 6)
    lam, mu, r = models.lame(np.full(shape, vp), np.full(shape, vs), np.full(shape, rho))
    m = models.Model("slope", ni, nj, nk, dt, steps, True, x, y, z, np.asfortranarray(lam), np.asfortranarray(mu),
                     np.asfortranarray(r), abs_number=(10, 10, 10, 10, 10, 0), abs_velocity=(vp,) * 6,
                     cfs_bmax=(1.0,) * 6, cfs_amax=(np.pi * f0,) * 6)
    i0, j0 = 3 + 28, 3 + nj // 2
    m.force = models.surface_force(ni, nj, nk, i0, j0, fz=1e16, f0=f0)
    d1, d2 = 20, 65
    m.rec_indx = np.array([[i0 + d + 1, j0 + 1, m.nk2 + 1] for d in (d1, d2)], np.int32)
    X, Y = np.meshgrid(x, y, indexing="ij")
    m = models.with_topography(m, slope * (X - x[(ni + 6) // 2]) * models.R_EARTH + 0.0 * Y)
    o = OracleSolver(m, 32, nthreads=nthreads)
    with mutation(mut):
        rec = o.run(steps).astype(np.float64)
    o.close()
    # distances along the surface, source -> receivers
    rs = m.r3[:, j0, m.nk2]

    def arc(a, b):
        return float(np.sum(np.hypot(np.diff(rs[a:b + 1]), 0.5 * (rs[a:b] + rs[a + 1:b + 1]) * np.diff(x[a:b + 1]))))
    s1, s2 = arc(i0, i0 + d1), arc(i0, i0 + d2)
    cR = 0.9194 * vs
    t = (np.arange(steps) + 1) * dt
    t0 = 1.2 / f0
    a = rec[:, 0, 2] * np.exp(-((t - (t0 + s1 / cR)) / 3.0) ** 8)
    b = rec[:, 1, 2] * np.exp(-((t - (t0 + s2 / cR)) / 3.0) ** 8)
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    frac = 0.5 * (cc[k - 1] - cc[k + 1]) / (cc[k - 1] - 2 * cc[k] + cc[k + 1])
    lag = (k - (steps - 1) + frac) * dt
    return (s2 - s1) / lag / cR, np.abs(rec).max()


def test_rayleigh_speed_along_a_sloping_surface():
    """A surface sloping at 0.15 along the path (the surface 1.1% longer
    than its map distance; the slope tapered to zero across the path over
    the CFS-PML): the surface-distance speed is c_R within 1% (measured
    0.996; the flat surface in the same geometry 0.989).  With the slope's
    sign flipped in the surface solve the speed was 0.979 and the run grew
    without bound after ~500 steps; that mistake is caught directly by
    test_surface_stress_rates_meet_the_traction_condition."""
    c, amp = _rayleigh_run(0.15)
    print(f"Rayleigh speed along the slope: {c:.4f} c_R (peak {amp:.2e})")
    assert c == pytest.approx(1.0, abs=0.01)


def test_long_run_under_hills_decays():
    m = models.hills_test(nt=900)
    o = OracleSolver(m, 32)
    rec = o.run(900)
    o.close()
    amp = np.abs(rec[:, :, :3]).max(axis=(1, 2))
    peak = amp.max()
    print(f"peak {peak:.3e}, last 100 steps {amp[-100:].max():.3e}")
    assert np.isfinite(amp).all()
    assert amp[-100:].max() < 0.2 * peak


def _pml_flat_violations(m):
    """cells of a CFS-PML slab whose +-3 box holds a node off r = z(k) (the
    rule fw_init enforces for mode C)"""
    bad = m.r3 != m.z[None, None, :]
    from scipy import ndimage
    bad = ndimage.maximum_filter(bad, size=7, mode="constant")  # the +-3 box
    nab = m.abs_number
    lo, hi = 3, (3 + m.ni - 1, 3 + m.nj - 1, 3 + m.nk - 1)
    n = 0
    for f in range(6):
        N = nab[f]
        if N <= 0:
            continue
        sl = [slice(lo, hi[0] + 1), slice(lo, hi[1] + 1), slice(lo, hi[2] + 1)]
        ax = f // 2
        sl[ax] = slice(lo, lo + N + 1) if f % 2 == 0 else slice(hi[ax] - N, hi[ax] + 1)
        n += int(bad[tuple(sl)].sum())
    return n


def test_topography_is_flat_over_the_pml_shell():
    for m in (models.hills_test(nt=1), models.hills_test(ni=84, nj=40, nk=28, nt=1),
              models.hills_400x400x200(nt=1), models.hills_256x256x128(nt=1)):
        assert _pml_flat_violations(m) == 0, m.name
        assert np.max(m.r3[:, :, m.nk2] - m.z[m.nk2]) > 500.0
