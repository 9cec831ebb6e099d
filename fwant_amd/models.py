"""Synthetic workloads in the reference's input layout.

A ``Model`` carries what the reference driver reads from SeisFD3D.conf and
the per-rank NetCDF inputs (coord_/metric_/media_/source_/station_ files,
code/srcF/mod_grid.F90, mod_media.F90, mod_src.F90, mod_io.F90): the
ghosted spherical grid (x = colatitude, y = longitude, z = radius), lambda,
mu, rho over the ghosted volume (column-major, i fastest), force / moment
source tables, receivers (1-based local indices) and the PML settings.

Presets (sizes and values of the named workloads):
  two_layer_100x100x50()   100x100x50, h = 1 km, 2-layer crust/mantle, 12-pt
                           CFS-PML on 5 sides, surface vertical force, Ricker
                           0.5 Hz, dt = 0.08 s, 1000 steps, 16 receivers
  gradient_400x400x200()   400x400x200, h = 0.5 km, Vs gradient 3.2->4.6 km/s
                           with +-5% correlated perturbation, Vp = 1.73 Vs,
                           Gardner density, Ricker 1 Hz force
  small_test(...)          a reduced heterogeneous case for parity tests
  hills_400x400x200()      gradient_400x400x200 on a terrain-following grid
                           under 20 seeded Gaussian hills (mode C)
  hills_256x256x128()      the batched-sources model: 256x256x128 under hills
  hills_test(...)          small_test under steep hills (mode C parity)
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

R_EARTH = 6371000.0
LENFD = 3
STF_RICKER = 200


@dataclass
class Sources:
    n: int = 0
    ntw: int = 1
    stf_id: int = STF_RICKER
    indx: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))   # 1-based local
    shift: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    t0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    stf_time: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    stf_freq: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    comps: np.ndarray = field(default_factory=lambda: np.zeros((3, 0), np.float32))  # [ncomp][ntw*n]


@dataclass
class Model:
    name: str
    ni: int
    nj: int
    nk: int
    stept: float
    nt: int
    freenode: bool
    x: np.ndarray        # float64, nx
    y: np.ndarray        # float64, ny
    z: np.ndarray        # float64, nz
    lam: np.ndarray      # float32 (nx, ny, nz) Fortran order
    mu: np.ndarray
    rho: np.ndarray
    abs_number: tuple = (0, 0, 0, 0, 0, 0)
    abs_velocity: tuple = (0.0,) * 6
    cfs_bmax: tuple = (1.0,) * 6
    cfs_amax: tuple = (0.0,) * 6
    force: Sources = field(default_factory=Sources)
    moment: Sources = field(default_factory=lambda: Sources(comps=np.zeros((6, 0), np.float32)))
    rec_indx: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))  # 1-based
    rec_tinv: int = 1
    # mode C: radius of every node of the ghosted grid, float64 (nx, ny, nz)
    # Fortran order (a terrain-following shell); None: the spherical grid of z
    r3: Optional[np.ndarray] = None

    @property
    def nx(self):
        return self.ni + 2 * LENFD

    @property
    def ny(self):
        return self.nj + 2 * LENFD

    @property
    def nz(self):
        return self.nk + 2 * LENFD

    @property
    def nk2(self):  # C index of the free surface
        return LENFD + self.nk - 1

    def ref_steps(self):
        """Steps the reference driver runs: 8*ceil((nt+1)/8) (seis3d_wave_8op_p3.F90:92-95)."""
        return 8 * ((self.nt + 1 + 7) // 8)


@dataclass
class Batch:
    """Members of a batched solve: the member of each force / moment source
    and the number of steps each member runs."""
    force_member: np.ndarray
    moment_member: np.ndarray
    nsteps: np.ndarray


def virtual_sources(model, stations, nsteps, fz=1.0e16, f0=None):
    """Ambient-noise virtual sources: one vertical surface force per station
    (C indices (i, j) on the free surface), member b running nsteps[b] steps.
    Returns a copy of `model` whose force table holds every station's force
    (the model's own sources are dropped) and the matching Batch."""
    import copy
    m = copy.copy(model)
    f0 = model.force.stf_freq[0] if (f0 is None and model.force.n) else (f0 or 0.5)
    n = len(stations)
    s = Sources(n=n, ntw=1, stf_id=STF_RICKER)
    s.indx = np.array([[i + 1, j + 1, m.nk2 + 1] for i, j in stations], np.int32)
    s.shift = np.zeros((n, 3), np.float32)
    s.t0 = np.zeros(n, np.float32)
    s.stf_time = np.array([1.2 / f0], np.float32)
    s.stf_freq = np.array([f0], np.float32)
    s.comps = np.zeros((3, n), np.float32)
    s.comps[2] = fz
    m.force = s
    m.moment = Sources(comps=np.zeros((6, 0), np.float32))
    b = Batch(np.arange(n, dtype=np.int32), np.zeros(0, np.int32), np.asarray(nsteps, np.int32))
    return m, b


def spherical_grid(ni, nj, nk, h, colat0=np.pi / 2, lon0=0.0, r_top=R_EARTH):
    """Uniform spherical grid: dtheta = dphi = h / r_top, dr = h; the top
    interior plane (k = nk2) sits at r_top.  Returns ghosted float64 x, y, z."""
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    d = h / r_top
    ic, jc = LENFD + (ni - 1) / 2.0, LENFD + (nj - 1) / 2.0
    x = colat0 + (np.arange(nx) - ic) * d
    y = lon0 + (np.arange(ny) - jc) * d
    nk2 = LENFD + nk - 1
    z = r_top - (nk2 - np.arange(nz)) * h
    return x, y, z


def lame(vp, vs, rho):
    mu = rho * vs * vs
    lam = rho * vp * vp - 2.0 * mu
    return lam.astype(np.float32), mu.astype(np.float32), rho.astype(np.float32)


def surface_force(model_ni, model_nj, nk, i, j, fz=1.0e16, fx=0.0, fy=0.0, f0=0.5, t0=None):
    """One vertical (or general) force at C indices (i, j, nk2) with a Ricker STF."""
    t0 = 1.2 / f0 if t0 is None else t0
    nk2 = LENFD + nk - 1
    s = Sources(n=1, ntw=1, stf_id=STF_RICKER)
    s.indx = np.array([[i + 1, j + 1, nk2 + 1]], np.int32)
    s.shift = np.zeros((1, 3), np.float32)
    s.t0 = np.zeros(1, np.float32)
    s.stf_time = np.array([t0], np.float32)
    s.stf_freq = np.array([f0], np.float32)
    s.comps = np.array([[fx], [fy], [fz]], np.float32)
    return s


def two_layer_100x100x50(nt=1000):
    """100x100x50 grid, h = 1 km, flat free surface, crust 30 km (Vp 6.0, Vs 3.46 km/s,
    rho 2.7 g/cc) over mantle (Vp 8.0, Vs 4.5, rho 3.3), 12-pt CFS-PML on 5 sides,
    vertical surface point force, Ricker f0 = 0.5 Hz, dt = 0.08 s, 16 surface receivers
    from numpy.random.default_rng(1)."""
    ni, nj, nk, h = 100, 100, 50, 1000.0
    x, y, z = spherical_grid(ni, nj, nk, h)
    depth = R_EARTH - z
    crust = depth < 30000.0 - 1e-6
    vp1 = np.where(crust, 6000.0, 8000.0)
    vs1 = np.where(crust, 3460.0, 4500.0)
    rh1 = np.where(crust, 2700.0, 3300.0)
    shape = (ni + 6, nj + 6, nk + 6)
    vp = np.broadcast_to(vp1[None, None, :], shape)
    vs = np.broadcast_to(vs1[None, None, :], shape)
    rho = np.broadcast_to(rh1[None, None, :], shape)
    lam, mu, rho = lame(vp, vs, rho)
    f0 = 0.5
    m = Model("two_layer_100x100x50", ni, nj, nk, 0.08, nt, True, x, y, z,
              np.asfortranarray(lam), np.asfortranarray(mu), np.asfortranarray(rho),
              abs_number=(12, 12, 12, 12, 12, 0),
              abs_velocity=(8000.0,) * 6, cfs_bmax=(3.0,) * 6, cfs_amax=(np.pi * f0,) * 6)
    m.force = surface_force(ni, nj, nk, LENFD + ni // 2, LENFD + nj // 2, f0=f0)
    rng = np.random.default_rng(1)
    lo, hi = LENFD + 14, LENFD + ni - 1 - 14
    ij = rng.integers(lo, hi + 1, size=(16, 2))
    m.rec_indx = np.stack([ij[:, 0] + 1, ij[:, 1] + 1, np.full(16, m.nk2 + 1)], 1).astype(np.int32)
    m.rec_tinv = 1
    return m


def _correlated_field(shape, corr_cells, seed, coarse=4):
    """Gaussian-correlated random field with unit max amplitude: white noise
    filtered with a Gaussian of the given correlation length (in cells) by
    FFT on a grid coarsened by `coarse`, then linearly interpolated."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    cs = tuple(max(4, s // coarse) for s in shape)
    w = rng.standard_normal(cs).astype(np.float32)
    sig = corr_cells / coarse
    f = ndimage.gaussian_filter(w, sigma=sig, mode="wrap")
    f = ndimage.zoom(f, [s / c for s, c in zip(shape, cs)], order=1)
    f = f[: shape[0], : shape[1], : shape[2]]
    return f / np.max(np.abs(f))


def gradient_400x400x200(nt=3000, ni=400, nj=400, nk=200, h=500.0, seed=3):
    """400x400x200 grid, h = 0.5 km (flat free surface: the spherical reference
    scheme has no topography).  Vs = 3.2 -> 4.6 km/s linear in depth times
    (1 + 0.05 p), p a Gaussian-correlated field (20 km) scaled to |p| <= 1;
    Vp = 1.73 Vs; rho = 1741 Vp[km/s]^0.25 kg/m^3 (Gardner); Ricker 1 Hz force;
    12-pt CFS-PML on 5 sides; dt = 0.04 s."""
    x, y, z = spherical_grid(ni, nj, nk, h)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    depth = np.clip(R_EARTH - z, 0.0, nk * h)
    vs1 = 3200.0 + (4600.0 - 3200.0) * depth / (nk * h)
    p = _correlated_field((nx, ny, nz), 20000.0 / h, seed)
    vs = (vs1[None, None, :] * (1.0 + 0.05 * p)).astype(np.float32)
    vp = (1.73 * vs).astype(np.float32)
    rho = (1741.0 * (vp / 1000.0) ** 0.25).astype(np.float32)
    lam, mu, rho = lame(vp.astype(np.float64), vs.astype(np.float64), rho.astype(np.float64))
    f0 = 1.0
    m = Model("gradient_400x400x200", ni, nj, nk, 0.04, nt, True, x, y, z,
              np.asfortranarray(lam), np.asfortranarray(mu), np.asfortranarray(rho),
              abs_number=(12, 12, 12, 12, 12, 0),
              abs_velocity=(8400.0,) * 6, cfs_bmax=(3.0,) * 6, cfs_amax=(np.pi * f0,) * 6)
    m.force = surface_force(ni, nj, nk, LENFD + ni // 2, LENFD + nj // 2, f0=f0)
    rng = np.random.default_rng(5)
    mg = min(20, ni // 4, nj // 4)
    ij = rng.integers(LENFD + mg, LENFD + min(ni, nj) - mg, size=(32, 2))
    m.rec_indx = np.stack([ij[:, 0] + 1, ij[:, 1] + 1, np.full(32, m.nk2 + 1)], 1).astype(np.int32)
    m.rec_tinv = 1
    return m


def small_test(ni=40, nj=36, nk=32, nt=120, h=1000.0, seed=7, freenode=True, pml=8,
               moment=True, colat0=np.pi / 2, r_top=R_EARTH, dt=0.05, water=False, fscale=1.0):
    """Reduced heterogeneous case exercising every path: random smooth media,
    PML on 5 (or 6) sides, a smoothed surface force with two STF windows,
    an interior moment source, surface and buried receivers.  fscale
    multiplies every source frequency (and divides the source times)."""
    rng = np.random.default_rng(seed)
    x, y, z = spherical_grid(ni, nj, nk, h, colat0=colat0, r_top=r_top)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    p = _correlated_field((nx, ny, nz), 6.0, seed, coarse=1)
    vs = 3400.0 * (1.0 + 0.08 * p)
    vp = 1.75 * vs
    rho = 2600.0 * (1.0 + 0.05 * p)
    if water:
        vs[: nx // 3, : ny // 3, : nz // 2] = 0.0
    lam, mu, rho = lame(vp, vs, rho)
    nab = (pml, pml, pml, pml, pml, 0 if freenode else pml)
    m = Model("small_test", ni, nj, nk, dt, nt, freenode, x, y, z,
              np.asfortranarray(lam), np.asfortranarray(mu), np.asfortranarray(rho),
              abs_number=nab, abs_velocity=(6500.0,) * 6, cfs_bmax=(2.5,) * 6,
              cfs_amax=(np.pi * 0.6 * fscale,) * 6)
    nk2 = LENFD + nk - 1
    f = Sources(n=2, ntw=2, stf_id=STF_RICKER)
    f.indx = np.array([[LENFD + ni // 2 + 1, LENFD + nj // 2 + 1, nk2 + 1],
                       [LENFD + ni // 3 + 1, LENFD + 2 * nj // 3 + 1, nk2 - 1 + 1]], np.int32)
    f.shift = np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 0.1]], np.float32)
    f.t0 = np.array([0.0, 0.3 / fscale], np.float32)
    f.stf_time = np.array([1.6 / fscale, 2.2 / fscale], np.float32)
    f.stf_freq = np.array([0.7 * fscale, 0.5 * fscale], np.float32)
    f.comps = (rng.standard_normal((3, 4)) * 1e15).astype(np.float32)
    f.comps[2] += 2e15
    m.force = f
    if moment:
        mo = Sources(n=1, ntw=1, stf_id=STF_RICKER)
        mo.indx = np.array([[LENFD + 2 * ni // 3 + 1, LENFD + nj // 3 + 1, LENFD + nk // 2 + 1]], np.int32)
        mo.shift = np.array([[0.1, 0.2, -0.3]], np.float32)
        mo.t0 = np.array([0.2 / fscale], np.float32)
        mo.stf_time = np.array([1.5 / fscale], np.float32)
        mo.stf_freq = np.array([0.6 * fscale], np.float32)
        mo.comps = (rng.standard_normal((6, 1)) * 1e15).astype(np.float32)
        m.moment = mo
    ri = rng.integers(LENFD, LENFD + ni, size=8)
    rj = rng.integers(LENFD, LENFD + nj, size=8)
    rk = np.array([nk2] * 4 + list(rng.integers(LENFD + 2, nk2 - 2, size=4)))
    m.rec_indx = np.stack([ri + 1, rj + 1, rk + 1], 1).astype(np.int32)
    m.rec_tinv = 1
    return m


def curved_test(ni=40, nj=36, nk=24, nt=96, seed=41, colat0=1.0, r_top=20000.0, h=200.0, **kw):
    """small_test on a strongly curved grid away from the equator: colatitude
    centred on colat0 (cot = 0.64, 1/sin = 1.19 at 1 rad), top radius r_top
    (20 km: h / r = 1e-2 per cell), spacing h.  The time step keeps the
    Courant number of small_test (0.32 on the smallest arc spacing, at the
    bottom of the grid), and the sources are scaled to the spacing.  On this
    grid the cot(theta), 1/sin(theta) and 1/r terms of the equations are of
    the order of the derivative terms."""
    x, _, z = spherical_grid(ni, nj, nk, h, colat0=colat0, r_top=r_top)
    dth = h / r_top
    dmin = min(h, z[LENFD] * dth * float(np.min(np.sin(x))))
    vmax = 1.75 * 3400.0 * 1.08
    dt = float(np.float32(0.32 * dmin / vmax))
    return small_test(ni=ni, nj=nj, nk=nk, nt=nt, h=h, seed=seed, colat0=colat0, r_top=r_top, dt=dt,
                      fscale=1000.0 / h * 0.8, **kw)


# ---------------------------------------------------------------------------
# Mode C: terrain-following grids (the reference's grid is a spherical shell;
# these workloads add topography).  A node (i, j, k) sits at colatitude x(i),
# longitude y(j) and radius
#     r = z(k) + topo(i, j) w(k),
# w rising smoothly (s^3 (10 - 15 s + 6 s^2), s the height fraction) from 0
# at plane k_flat to 1 at the top interior plane nk2 and staying 1 above.
# The stage kernels need the grid flat (p = q = 0, r = z) on every CFS-PML
# slab cell: topography is tapered to zero over the x / y slabs and 3 more
# columns, and w is 0 up to 3 planes above the bottom slab.
# ---------------------------------------------------------------------------
def terrain_radius(z, topo, k_flat, nk2):
    """r(i, j, k) = z(k) + topo(i, j) w(k) (float64, Fortran order)."""
    k = np.arange(len(z), dtype=np.float64)
    s = np.clip((k - k_flat) / float(nk2 - k_flat), 0.0, 1.0)
    w = s ** 3 * (10.0 - 15.0 * s + 6.0 * s * s)
    r = z[None, None, :] + topo[:, :, None] * w[None, None, :]
    return np.asfortranarray(r)


def _edge_taper(n, flat, ramp):
    """1-D window: 0 on `flat` entries at each end, a raised-cosine ramp of
    `ramp` entries, 1 inside."""
    t = np.ones(n)
    i = np.arange(n, dtype=np.float64)
    d = np.minimum(i - (flat - 1), (n - flat) - i)  # distance from the flat band
    t = np.where(d <= 0, 0.0, np.where(d >= ramp, 1.0, 0.5 - 0.5 * np.cos(np.pi * d / ramp)))
    return t


def gaussian_hills(x, y, r_top, n, seed, amp=(0.0, 2000.0), sigma=(5000.0, 20000.0)):
    """Sum of n Gaussian hills (numpy default_rng(seed): centres uniform over
    the grid, amplitude and width uniform in the given ranges, metres) as a
    height over the (x, y) nodes; distances are arcs at radius r_top."""
    rng = np.random.default_rng(seed)
    xc = rng.uniform(x[0], x[-1], n)
    yc = rng.uniform(y[0], y[-1], n)
    a = rng.uniform(amp[0], amp[1], n)
    sg = rng.uniform(sigma[0], sigma[1], n)
    X, Y = np.meshgrid(x, y, indexing="ij")
    topo = np.zeros(X.shape)
    for q in range(n):
        d2 = ((X - xc[q]) * r_top) ** 2 + ((Y - yc[q]) * r_top * np.sin(xc[q])) ** 2
        topo += a[q] * np.exp(-0.5 * d2 / (sg[q] * sg[q]))
    return topo


def with_topography(m, topo, ramp=8):
    """Mode C copy of model m: the grid under the heights topo (nx, ny),
    tapered to zero over the CFS-PML shell and 3 columns more, and w zero up
    to 3 planes above the bottom slab.  The media stay as they are in index
    space (layers follow the terrain).  The sources and receivers keep their
    nodes."""
    import copy
    n = copy.copy(m)
    nab = m.abs_number
    # a slab of N points spans N + 1 nodes from the first (last) interior
    # node; it and 3 nodes beyond stay flat
    fx = LENFD + max(nab[0], nab[1]) + 4
    fy = LENFD + max(nab[2], nab[3]) + 4
    t = _edge_taper(m.nx, fx, ramp)[:, None] * _edge_taper(m.ny, fy, ramp)[None, :]
    k_flat = LENFD + nab[4] + 3
    n.r3 = terrain_radius(np.asarray(m.z, np.float64), topo * t, k_flat, m.nk2)
    n.name = m.name + "+topo"
    return n


def hills_400x400x200(nt=3000, ni=400, nj=400, nk=200, h=500.0):
    """gradient_400x400x200 under 20 Gaussian hills (numpy default_rng(2):
    amplitude 0-2 km, width sigma 5-20 km), tapered to zero over the 12-point
    CFS-PML shell (mode C)."""
    m = gradient_400x400x200(nt=nt, ni=ni, nj=nj, nk=nk, h=h)
    topo = gaussian_hills(m.x, m.y, R_EARTH, 20, seed=2)
    n = with_topography(m, topo)
    n.name = f"hills_{ni}x{nj}x{nk}"
    return n


def hills_256x256x128(nt=4000):
    """The batched-sources configuration's curvilinear model: the gradient
    model at 256x256x128 (h = 0.5 km, seed 4) under 20 Gaussian hills
    (default_rng(2), 0-2 km, sigma 5-20 km), tapered over the CFS-PML shell."""
    m = gradient_400x400x200(nt=nt, ni=256, nj=256, nk=128, h=500.0, seed=4)
    n = with_topography(m, gaussian_hills(m.x, m.y, R_EARTH, 20, seed=2))
    n.name = "hills_256x256x128"
    return n


def hills_test(ni=40, nj=36, nk=28, nt=96, seed=7, amp=3000.0, sigma=6000.0, nhill=3, **kw):
    """small_test under nhill steep Gaussian hills (mode C parity): heights
    up to amp, width sigma, slopes up to ~0.3 on the default 1 km grid."""
    m = small_test(ni=ni, nj=nj, nk=nk, nt=nt, seed=seed, **kw)
    topo = gaussian_hills(m.x, m.y, R_EARTH, nhill, seed=seed + 100, amp=(0.5 * amp, amp),
                          sigma=(0.7 * sigma, sigma))
    n = with_topography(m, topo, ramp=4)
    n.name = "hills_test"
    return n
