"""The preprocessors that write a run's input files from the reference's
recipe files (one rank, suffix mpi000000).

    python -m fwant_amd.prep SeisFD3D.conf [grid] [metric] [media] [source] [station]

  grid_coords      seis3d_grid.F90:89-133 (GRID_CONF: distance2meter,
                   steph, theta0_phi0_rmax; a step below SEIS_ZERO means an
                   explicit list after the line <x / <y / <z; x, y in
                   degrees; three ghost points each side extrapolated from
                   the first / last interval)
  media            seis3d_media.F90:854-1262 (effmedia_eval): per interior
                   cell max(2N,1) sub-cell samples per axis at the offsets
                   m*h/(2(N+1)), m = +-1..+-N (N = half_sample_point; the
                   node itself when N = 0); each sample yields an upper and
                   a lower value (the same away from an interface); rho is
                   the arithmetic mean, mu and lambda the harmonic means;
                   a sample with mu <= SEIS_ZERO on either side is water:
                   mu = 0 when half the samples or more are water, else the
                   harmonic mean over the solid samples.  Backgrounds
                   cart1d (layer1d_read :331-420, evaluation :915-973),
                   layered (layered_read :535-644, evaluation :1053-1083)
                   and volume (volume_init/read :706-783, evaluation
                   :1123-1133), perturbation volume (:1163-1176); depth is
                   measured from the top radius z(nk2).  Then media_extend
                   (:1264-1299) copies the edge values into the ghosts.
  cfl_check        media_stept_check (:1544-1606): the smallest
                   1.3/Vp * (distance from a node to the 8 planes through
                   its axis neighbours), on Cartesian points
                   (r sin(theta) cos(phi), r sin(theta) sin(phi), -r cos(theta))
  read_src_conf /  seis3d_source.F90:324-398 (SOURCE_CONF) and :57-162
  locate_sources   (nearest node; z above the grid and above
                   src_hyper_height -> the free surface; shift = offset over
                   the spacing on the side of the offset), angle2moment
                   :400-428 (Aki & Richards, then Mxy and Myz change sign)
  read_stations /  io_pt_read (mod_io.F90:112-152: recv_NNN, line_NNN) and
  locate_stations  receiver_locate (seis3d_station.F90:64-105: nearest node;
                   above topo_hyper_height -> k = nk2; outside the interior
                   -> dropped)

Not built: the interface, composite and verpoly media types and Qs
(WITHQS builds).  Arithmetic is FP64; the files hold FP32 as the
reference's do.
"""
import os
import sys
from dataclasses import dataclass

import numpy as np
from scipy.io import netcdf_file

from . import io as fio
from . import models

LENFD = 3
ZERO = 1.0e-20  # SEIS_ZERO, mod_constants.F90:29


def _conf_lines(path):
    with open(path) as f:
        return f.read().splitlines()


def _after(lines, key, path):
    """Lines after the first whose first word is `key` (string_conf's
    positioning before the list-directed reads that follow)."""
    for q, ln in enumerate(lines):
        if fio._words(ln)[:1] == [key]:
            return lines[q + 1:]
    raise KeyError(f"{path}: there is no {key} on column 1")


def _numbers(lines, count, path):
    """`count` numbers read list-directed from the lines (several per line
    allowed)."""
    out = []
    for ln in lines:
        for w in fio._words(ln.replace(",", " ")):
            out.append(float(w.replace("d", "e").replace("D", "e")))
            if len(out) == count:
                return out
    raise ValueError(f"{path}: {count} values wanted, {len(out)} found")


def _rows(lines, count, width, path):
    """`count` records of `width` numbers each, one record starting per line."""
    out = []
    it = iter(lines)
    for _ in range(count):
        rec = []
        while len(rec) < width:
            try:
                ln = next(it)
            except StopIteration:
                raise ValueError(f"{path}: {count} records of {width} values wanted") from None
            rec += [float(w.replace("d", "e").replace("D", "e")) for w in fio._words(ln.replace(",", " "))]
        out.append(rec[:width])
    return np.array(out, np.float64)


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
def grid_coords(path, ni, nj, nk):
    """(x, y, z) ghosted axes of lengths ni+6, nj+6, nk+6 from GRID_CONF:
    x colatitude and y longitude in radians, z radius in metres, z[nk2] =
    rmax (the free surface)."""
    lines = _conf_lines(path)
    conf = fio.Conf("\n".join(lines))
    d2m = conf.get("distance2meter", 2, float)
    dh = conf.row("steph", 2, 3)
    o = conf.row("theta0_phi0_rmax", 2, 3)
    out = []
    for d, (n, key) in enumerate(zip((ni, nj, nk), ("<x", "<y", "<z"))):
        g = np.zeros(n + 2 * LENFD)
        p = np.arange(n + 2 * LENFD, dtype=np.float64)
        if dh[d] >= ZERO:
            g[:] = ((p - LENFD) if d < 2 else (p + 1 - (n + 2 * LENFD) + LENFD)) * dh[d] + o[d]
        else:
            g[LENFD:LENFD + n] = _numbers(_after(lines, key, path), n, path)
        g = g * (np.pi / 180.0 if d < 2 else d2m)
        a, b = LENFD, LENFD + n - 1
        for q in range(1, LENFD + 1):
            g[a - q] = g[a - q + 1] + g[a] - g[a + 1]
            g[b + q] = g[b + q - 1] + g[b] - g[b - 1]
        out.append(g)
    return tuple(out)


# ---------------------------------------------------------------------------
# media models
# ---------------------------------------------------------------------------
def _check_media(vp, vs, dp, path):
    if np.any(dp <= 0) or np.any(vp <= 0) or np.any(vs < 0):
        raise ValueError(f"{path}: media parameter is negative")
    if np.any(vp ** 2 <= 2.0 * vs ** 2):
        raise ValueError(f"{path}: Vp**2 <= 2 Vs**2")


class Cart1D:
    """A 1-D profile of (depth, rho, Vp, Vs) nodes with linear variation
    between nodes; two nodes at the same depth make a discontinuity."""

    def __init__(self, d, dp, vp, vs):
        self.d, self.dp, self.vp, self.vs = (np.asarray(a, np.float64) for a in (d, dp, vp, vs))

    @classmethod
    def read(cls, path):
        lines = _conf_lines(path)
        conf = fio.Conf("\n".join(lines))
        kmax = conf.get("radial_sampling", 2, int)
        if kmax <= 1:
            raise ValueError(f"{path}: radial_sampling must be larger than 1")
        d2m = conf.get("distance2meter", 2, float)
        v2ms = conf.get("velocity2m/s", 2, float)
        d2k = conf.get("density2kg/m^3", 2, float)
        r = _rows(_after(lines, "<anchor_data>", path), kmax, 4, path)
        z, dp, vp, vs = r[:, 0], r[:, 1] * d2k, r[:, 2] * v2ms, r[:, 3] * v2ms
        _check_media(vp, vs, dp, path)
        meaning = conf.get("radial_meaning", 2)
        if meaning == "axis":
            d = max(z[0], z[-1]) - z
        elif meaning == "thickness":
            d = np.concatenate([[0.0], np.cumsum(z[:-1])])
        elif meaning == "depth":
            d = z.copy()
        else:
            raise ValueError("radial_meaning can only take axis, depth or thickness")
        if d[0] > d[-1]:
            d, vp, vs, dp = d[::-1], vp[::-1], vs[::-1], dp[::-1]
        return cls(d * d2m, dp, vp, vs)

    def pair(self, d0):
        """(vp, vs, dp) each (2, n): the upper and lower values at depths d0."""
        d0 = np.atleast_1d(np.asarray(d0, np.float64))
        d, K = self.d, len(self.d)
        i1 = np.zeros(d0.shape, np.int64)
        i2 = np.zeros(d0.shape, np.int64)
        w1 = np.ones(d0.shape)
        lin = np.zeros(d0.shape, bool)  # linear between nodes i1, i2 with weight w1 on i1
        same = lambda a, b: np.abs(d[a] - d[b]) < ZERO  # noqa: E731
        top = (np.abs(d[0] - d0) < ZERO) | (d0 < d[0])
        bot_hit = ~top & (np.abs(d[K - 1] - d0) < ZERO)
        below = ~top & ~bot_hit & (d[K - 1] < d0)
        inner = ~(top | bot_hit | below)
        i1[top], i2[top] = 0, 0
        jump = same(K - 2, K - 1)
        i1[bot_hit], i2[bot_hit] = (K - 2 if jump else K - 1), K - 1
        i1[below], i2[below] = K - 1, K - 1
        if inner.any():
            q = np.nonzero(inner)[0]
            x = d0[q]
            # the node hit by the bisection (first node within SEIS_ZERO the
            # search meets) or the bracketing interval
            k1 = np.zeros(len(q), np.int64)
            k2 = np.full(len(q), K - 1)
            hit = np.zeros(len(q), bool)
            while True:
                act = ~hit & (k2 - k1 > 1)
                if not act.any():
                    break
                k0 = (k2 - k1) // 2 + k1
                eq = act & (np.abs(d[k0] - x) < ZERO)
                hit |= eq
                k1 = np.where(eq, k0, k1)
                k2 = np.where(eq, k0, k2)
                up = act & ~eq & (d[k0] > x)
                k2 = np.where(up, k0, k2)
                k1 = np.where(act & ~eq & ~up, k0, k1)
            a, b = k1.copy(), k2.copy()
            if hit.any():
                h = np.nonzero(hit)[0]
                k = k1[h]
                prev = (k >= 1) & (np.abs(d[np.maximum(k - 1, 0)] - d[k]) < ZERO)
                nxt = ~prev & (k + 1 < K) & (np.abs(d[np.minimum(k + 1, K - 1)] - d[k]) < ZERO)
                a[h] = np.where(prev, k - 1, k)
                b[h] = np.where(nxt, k + 1, k)
            i1[q], i2[q] = a, b
            L1 = np.ones(len(q))
            L1[~hit] = (d[k2[~hit]] - x[~hit]) / (d[k2[~hit]] - d[k1[~hit]])
            w1[q] = L1
            lin[q] = ~hit
        out = []
        for v in (self.vp, self.vs, self.dp):
            up = np.where(lin, v[i1] * w1 + v[i2] * (1.0 - w1), v[i1])
            lo = np.where(lin, up, v[i2])
            out.append(np.stack([up, lo]))
        return out


def _locate(x0, x):
    """indx_locate_1d (seis3d_media.F90:1314-1340): bracketing indices (0-based)
    and the clamped coordinate."""
    x0 = np.asarray(x0, np.float64)
    n = len(x)
    i2 = np.clip(np.searchsorted(x, x0, side="right"), 1, n - 1)
    i1 = i2 - 1
    x1 = np.clip(x0, x[0], x[-1])
    return i1, i2, x1


def _lin2(x, y, f, x0, y0):
    """interp_2d with 2x2 points (mod_math.F90:128-161): bilinear over the
    bracketing cell, clamped at the ends; f (ni, nj[, ...])."""
    i1, i2, x1 = _locate(x0, x)
    j1, j2, y1 = _locate(y0, y)
    ax = (x1 - x[i1]) / (x[i2] - x[i1])
    ay = (y1 - y[j1]) / (y[j2] - y[j1])
    ex = (slice(None),) + (None,) * (f.ndim - 2)
    ax, ay = ax[ex], ay[ex]
    return ((1 - ax) * (1 - ay)) * f[i1, j1] + (ax * (1 - ay)) * f[i2, j1] + \
        ((1 - ax) * ay) * f[i1, j2] + (ax * ay) * f[i2, j2]


class Layered:
    """Layers with thickness h(x, y) given at horizontal nodes; each layer
    has top and bottom values with linear variation in depth."""

    def __init__(self, x, y, h, vp, vs, dp):
        self.x, self.y = np.asarray(x, np.float64), np.asarray(y, np.float64)
        self.h = np.asarray(h, np.float64)  # (ni, nj, nl)
        self.vp, self.vs, self.dp = (np.asarray(a, np.float64) for a in (vp, vs, dp))  # (2, nl)

    @classmethod
    def read(cls, path):
        lines = _conf_lines(path)
        conf = fio.Conf("\n".join(lines))
        kmax = conf.get("number_of_layer", 2, int)
        if kmax < 1:
            raise ValueError(f"{path}: at least 1 layer")
        d2m = conf.get("distance2meter", 2, float)
        v2ms = conf.get("velocity2m/s", 2, float)
        d2k = conf.get("density2kg/m^3", 2, float)
        imax = conf.get("horizontal_sampling", 2, int)
        jmax = conf.get("horizontal_sampling", 3, int)
        m = _rows(_after(lines, "<anchor_media>", path), kmax, 6, path)
        vp, vs, dp = m[:, 0:2].T * v2ms, m[:, 2:4].T * v2ms, m[:, 4:6].T * d2k
        _check_media(vp, vs, dp, path)
        meaning = conf.get("layer_meaning", 2)
        r = _rows(_after(lines, "<anchor_layer>", path), imax * jmax, 2 + kmax, path)
        x = r[:imax, 0]
        y = r[::imax, 1]
        v = r[:, 2:].reshape(jmax, imax, kmax).transpose(1, 0, 2)  # (i, j, layer), j outer in the file
        if meaning == "depth":
            h = np.concatenate([v[:, :, :1], np.diff(v, axis=2)], axis=2)
        elif meaning == "thickness":
            h = v
        else:
            raise ValueError("layer_meaning can only take depth or thickness")
        if np.any(h[:, :, :-1] < 0.0):
            raise ValueError(f"{path}: thickness of layer should not be less than 0")
        if np.any(h[:, :, -1] < ZERO):
            raise ValueError(f"{path}: thickness of lowest layer should be larger than 0")
        return cls(x * np.pi / 180.0, y * np.pi / 180.0, h * d2m, vp, vs, dp)

    def pair(self, x0, y0, d0, tol):
        """(vp, vs, dp) each (2, n) at points (x0, y0, depth d0); within
        `tol` of an interface the upper value is the bottom of the layer
        above and the lower value the top of the next layer of nonzero
        thickness."""
        x0, y0, d0 = (np.atleast_1d(np.asarray(a, np.float64)) for a in np.broadcast_arrays(x0, y0, d0))
        tol = np.broadcast_to(np.asarray(tol, np.float64), d0.shape)
        n, nl = d0.shape[0], self.h.shape[2]
        h = _lin2(self.x, self.y, self.h, x0, y0)  # (n, nl)
        vals = [np.zeros((2, n)) for _ in range(3)]
        done = np.zeros(n, bool)
        d = np.zeros(n)
        for q in range(nl):
            live = ~done & (h[:, q] > ZERO)
            d = np.where(live, d + h[:, q], d)
            at = live & (np.abs(d - d0) <= tol)
            if at.any():
                if q == nl - 1:
                    raise ValueError("please increase the thickness of the lowest layer")
                nxt = np.full(n, -1)
                for m in range(nl - 1, q, -1):
                    nxt = np.where(h[:, m] > ZERO, m, nxt)
                if np.any(nxt[at] < 0):
                    raise ValueError("please increase the thickness of the lowest layer")
                for v, src in zip(vals, (self.vp, self.vs, self.dp)):
                    v[0, at] = src[1, q]
                    v[1, at] = src[0, nxt[at]]
            inside = live & ~at & (d > d0)
            if inside.any():
                L1 = (d[inside] - d0[inside]) / h[inside, q]
                for v, src in zip(vals, (self.vp, self.vs, self.dp)):
                    val = src[0, q] * L1 + src[1, q] * (1.0 - L1)
                    v[0, inside] = val
                    v[1, inside] = val
            done |= at | inside
        if not done.all():
            raise ValueError("please increase the thickness of the lowest layer")
        return vals


class Volume:
    """Vp, Vs, rho (or their relative perturbations) on a (theta, phi,
    depth below `sealevel`) lattice, trilinear with clamped ends."""

    def __init__(self, x, y, rz, rz0, vp, vs, dp):
        self.x, self.y, self.rz = (np.asarray(a, np.float64) for a in (x, y, rz))
        self.rz0 = float(rz0)
        self.vp, self.vs, self.dp = (np.asarray(a, np.float64) for a in (vp, vs, dp))  # (ni, nj, nk)

    @classmethod
    def read(cls, path):
        with netcdf_file(path, "r", mmap=False) as nc:
            g = {k: np.array(nc.variables[k][:], np.float64) for k in ("theta", "phi", "depth2sealevel")}
            v = {k: fio._fortran3(np.array(nc.variables[k][:], np.float64)) for k in ("Vp", "Vs", "rho")}
            rz0 = float(np.asarray(nc.sealevel).ravel()[0])
        return cls(g["theta"] * np.pi / 180.0, g["phi"] * np.pi / 180.0, g["depth2sealevel"], rz0,
                   v["Vp"], v["Vs"], v["rho"])

    def values(self, x0, y0, z0):
        x0, y0, z0 = (np.atleast_1d(np.asarray(a, np.float64)) for a in np.broadcast_arrays(x0, y0, z0))
        i1, i2, x1 = _locate(x0, self.x)
        j1, j2, y1 = _locate(y0, self.y)
        k1, k2, z1 = _locate(self.rz0 - z0, self.rz)
        ax = (x1 - self.x[i1]) / (self.x[i2] - self.x[i1])
        ay = (y1 - self.y[j1]) / (self.y[j2] - self.y[j1])
        az = (z1 - self.rz[k1]) / (self.rz[k2] - self.rz[k1])
        out = []
        for f in (self.vp, self.vs, self.dp):
            r = 0.0
            for (ii, wx) in ((i1, 1 - ax), (i2, ax)):
                for (jj, wy) in ((j1, 1 - ay), (j2, ay)):
                    for (kk, wz) in ((k1, 1 - az), (k2, az)):
                        r = r + wx * wy * wz * f[ii, jj, kk]
            out.append(r)
        return out


@dataclass
class MediaConf:
    half: tuple
    background: object
    perturbed: object = None


def read_media_conf(path):
    """MEDIA_CONF (init_media, seis3d_media.F90:261-329)."""
    conf = fio.read_conf(path)
    base = os.path.dirname(os.path.abspath(path))
    rel = lambda p: p if os.path.isabs(p) else os.path.join(base, p)  # noqa: E731
    half = tuple(conf.row("half_sample_point", 2, 3, int))
    kind = conf.get("background_type")
    fmt = conf.get("background_format")
    fn = rel(conf.get("background_filename"))
    if kind == "cart1d":
        if fmt != "ascii":
            raise ValueError("only ascii input accepted for cart1d type")
        bg = Cart1D.read(fn)
    elif kind == "layered":
        if fmt != "ascii":
            raise ValueError("only ascii input accepted for layered type")
        bg = Layered.read(fn)
    elif kind == "volume":
        if fmt != "nc":
            raise ValueError("only nc input accepted for volume type")
        bg = Volume.read(fn)
    else:
        raise ValueError(f"background_type {kind!r} is not built here (cart1d, layered, volume)")
    pert = None
    pk = conf.get("perturbed_type", 2, str, default="none")
    if pk == "volume":
        if conf.get("perturbed_format") != "nc":
            raise ValueError("only nc input accepted for volume type")
        pert = Volume.read(rel(conf.get("perturbed_filename")))
    elif pk != "none":
        raise ValueError(f"perturbed_type {pk!r} is not built here (volume, none)")
    return MediaConf(half, bg, pert)


def _offsets(c, N):
    """Sample coordinates around each node of axis c (interior nodes):
    (n_interior, max(2N,1)), and the upper half-spacing of each node."""
    c = np.asarray(c, np.float64)
    i = np.arange(LENFD, len(c) - LENFD)
    if N == 0:
        s = c[i][:, None]
    else:
        d1 = (c[i] - c[i - 1]) / 2.0 / (N + 1)
        d2 = (c[i + 1] - c[i]) / 2.0 / (N + 1)
        m = np.arange(1, N + 1)
        s = np.concatenate([c[i][:, None] - m[::-1][None, :] * d1[:, None],
                            c[i][:, None] + m[None, :] * d2[:, None]], axis=1)
    dz2 = (c[i + 1] - c[i]) / 2.0 / (N + 1)
    return s, dz2


def effective_media(mc, x, y, z):
    """lambda, mu, rho (FP32, ghosted, Fortran order) of the sub-cell
    averaged medium on the grid (x, y, z ghosted axes)."""
    Nx, Ny, Nz = mc.half
    nx, ny, nz = len(x), len(y), len(z)
    ni, nj = nx - 2 * LENFD, ny - 2 * LENFD
    xs, _ = _offsets(x, Nx)
    ys, _ = _offsets(y, Ny)
    zs, dz2 = _offsets(z, Nz)
    ztopo = z[nz - LENFD - 1]
    nsamp = xs.shape[1] * ys.shape[1] * zs.shape[1]
    lam = np.zeros((nx, ny, nz), np.float32, order="F")
    mu = np.zeros_like(lam)
    rho = np.zeros_like(lam)
    for kq in range(zs.shape[0]):
        rho0 = np.zeros((ni, nj))
        mu0 = np.zeros((ni, nj))
        lam0 = np.zeros((ni, nj))
        nwat = np.zeros((ni, nj), np.int64)
        for z0 in zs[kq]:
            d0 = max(ztopo - z0, 0.0)
            for mj in range(ys.shape[1]):
                for mi in range(xs.shape[1]):
                    X = np.broadcast_to(xs[:, mi][:, None], (ni, nj))
                    Y = np.broadcast_to(ys[:, mj][None, :], (ni, nj))
                    if isinstance(mc.background, Cart1D):
                        vp, vs, dp = (np.broadcast_to(a[:, :, None], (2, ni, nj)) for a in
                                      (v[:, :1] for v in mc.background.pair([d0])))
                    elif isinstance(mc.background, Layered):
                        vp, vs, dp = (v.reshape(2, ni, nj) for v in
                                      mc.background.pair(X.ravel(), Y.ravel(), np.full(ni * nj, d0), dz2[kq] / 5.0))
                    else:
                        vp, vs, dp = (np.broadcast_to(v.reshape(1, ni, nj), (2, ni, nj)) for v in
                                      mc.background.values(X.ravel(), Y.ravel(), np.full(ni * nj, z0)))
                    if mc.perturbed is not None:
                        p = mc.perturbed.values(X.ravel(), Y.ravel(), np.full(ni * nj, z0))
                        vp, vs, dp = (a * (1.0 + b.reshape(1, ni, nj)) for a, b in zip((vp, vs, dp), p))
                    m = dp * vs * vs
                    la = vp * vp * dp - 2.0 * m
                    rho0 += 0.5 * (dp[0] + dp[1])
                    wat = (m[0] <= ZERO) | (m[1] <= ZERO)
                    nwat += wat
                    with np.errstate(divide="ignore"):
                        mu0 += np.where(wat, 0.0, 0.5 * (1.0 / m[0] + 1.0 / m[1]))
                    lam0 += 0.5 * (1.0 / la[0] + 1.0 / la[1])
        k = kq + LENFD
        rho[LENFD:nx - LENFD, LENFD:ny - LENFD, k] = rho0 / nsamp
        with np.errstate(divide="ignore", invalid="ignore"):
            mu[LENFD:nx - LENFD, LENFD:ny - LENFD, k] = np.where((nwat > 0) & (nwat >= nsamp // 2), 0.0,
                                                                 (nsamp - nwat) / mu0)
        lam[LENFD:nx - LENFD, LENFD:ny - LENFD, k] = nsamp / lam0
    for a in (rho, mu, lam):
        extend_equal(a)
    return lam, mu, rho


def extend_equal(w):
    """media_extend (seis3d_media.F90:1264-1299): the edge values copied into
    the three ghost layers, x then y then z."""
    nx, ny, nz = w.shape
    a, b = LENFD, nx - LENFD - 1
    w[:a] = w[a]
    w[b + 1:] = w[b]
    a, b = LENFD, ny - LENFD - 1
    w[:, :a] = w[:, a:a + 1]
    w[:, b + 1:] = w[:, b:b + 1]
    a, b = LENFD, nz - LENFD - 1
    w[:, :, :a] = w[:, :, a:a + 1]
    w[:, :, b + 1:] = w[:, :, b:b + 1]


@dataclass
class CFL:
    dtmax: float
    indx: tuple  # 1-based ghosted (i, j, k)
    vp: float
    length: float


def cfl_check(x, y, z, lam, mu, rho):
    """media_stept_check: the largest stable step 1.3/Vp * L, L the smallest
    distance from a node to the 8 planes through one neighbour along each
    axis (either side)."""
    x, y, z = (np.asarray(a, np.float64) for a in (x, y, z))
    st = np.sin(x)[:, None]
    ct = np.cos(x)[:, None]
    cp, sp = np.cos(y)[None, :], np.sin(y)[None, :]

    def plane(k):
        return np.stack([z[k] * st * cp, z[k] * st * sp, -z[k] * ct * np.ones_like(cp)])  # (3, nx, ny)

    nx, ny, nz = len(x), len(y), len(z)
    I = slice(LENFD, nx - LENFD)
    J = slice(LENFD, ny - LENFD)
    best = CFL(np.inf, (0, 0, 0), 0.0, 0.0)
    sh = {-1: lambda a, ax: np.roll(a, 1, axis=ax), 1: lambda a, ax: np.roll(a, -1, axis=ax)}
    for k in range(LENFD, nz - LENFD):
        P = plane(k)
        L = np.full((nx, ny), np.inf)
        for sz in (-1, 1):
            C = plane(k + sz)
            for sy in (-1, 1):
                B = sh[sy](P, 2)
                for sx in (-1, 1):
                    A = sh[sx](P, 1)
                    n = np.cross(B - A, C - A, axis=0)
                    L = np.minimum(L, np.abs(np.sum(n * (P - A), axis=0)) / np.sqrt(np.sum(n * n, axis=0)))
        vp = np.sqrt((lam[:, :, k].astype(np.float64) + 2.0 * mu[:, :, k]) / rho[:, :, k])
        dt = (1.3 / vp * L)[I, J]
        q = np.unravel_index(np.argmin(dt.T), dt.T.shape)  # j outer, i inner
        j, i = q
        if dt[i, j] < best.dtmax:
            best = CFL(float(dt[i, j]), (int(i) + LENFD + 1, int(j) + LENFD + 1, k + 1), float(vp[I, J][i, j]),
                       float(L[I, J][i, j]))
    return best


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------
STF_IDS = {v: k for k, v in fio.STF_NAMES.items()}


def angle2moment(strike, dip, rake):
    """(Mxx, Myy, Mzz, Mxy, Mxz, Myz) of a unit double couple from strike,
    dip, rake in degrees (Aki & Richards' north-east-down components, then
    to the grid's (theta, phi, r) axes: Mxy and Myz change sign)."""
    s, d, r = (np.radians(np.asarray(a, np.float64)) for a in (strike, dip, rake))
    M11 = -(np.sin(d) * np.cos(r) * np.sin(2 * s) + np.sin(2 * d) * np.sin(r) * np.sin(s) ** 2)
    M22 = np.sin(d) * np.cos(r) * np.sin(2 * s) - np.sin(2 * d) * np.sin(r) * np.cos(s) ** 2
    M33 = -(M11 + M22)
    M12 = np.sin(d) * np.cos(r) * np.cos(2 * s) + 0.5 * np.sin(2 * d) * np.sin(r) * np.sin(2 * s)
    M13 = -(np.cos(d) * np.cos(r) * np.cos(s) + np.cos(2 * d) * np.sin(r) * np.sin(s))
    M23 = -(np.cos(d) * np.cos(r) * np.sin(s) - np.cos(2 * d) * np.sin(r) * np.cos(s))
    return M11, M22, M33, -M12, M13, -M23


@dataclass
class SourceSet:
    src: models.Sources
    axis: np.ndarray  # (n, 3): colatitude, longitude (rad), radius (m)


def read_src_conf(path):
    """(force, moment, src_hyper_height) from SOURCE_CONF (read_src_para,
    seis3d_source.F90:324-398): per source and time window a line
    'x y z t0 f0 Fx Fy Fz' (x, y in degrees; the components times f0), or
    'x y z t0 m0 Mxx Myy Mzz Mxy Mxz Myz' / 'x y z t0 m0 strike dip rake'."""
    lines = _conf_lines(path)
    conf = fio.Conf("\n".join(lines))
    hyper = conf.get("src_hyper_height", 2, float)
    out = []
    for kind, ncomp in (("force", 3), ("moment", 6)):
        n = conf.get(f"number_of_{kind}_source", 2, int)
        if n < 1:
            s = models.Sources(n=0, ntw=1)
            if kind == "moment":
                s.comps = np.zeros((6, 0), np.float32)
            out.append(SourceSet(s, np.zeros((0, 3))))
            continue
        ntw = conf.get(f"{kind}_stf_window", 2, int)
        stype = conf.get(f"{kind}_stf_type")
        if stype not in STF_IDS:
            raise ValueError(f"Have you told me how to generate the STF of {stype}")
        s = models.Sources(n=n, ntw=ntw, stf_id=STF_IDS[stype])
        s.stf_time = np.array(conf.row(f"{kind}_stf_timefactor", 2, ntw), np.float32)
        s.stf_freq = np.array(conf.row(f"{kind}_stf_freqfactor", 2, ntw), np.float32)
        angles = kind == "moment" and conf.get("moment_mech_input") != "moment"
        width = 5 + (3 if angles else ncomp)
        r = _rows(_after(lines, f"<anchor_{kind}>", path), n * ntw, width, path).reshape(n, ntw, width)
        axis = r[:, -1, 0:3].copy()  # the last window's line sets the point
        axis[:, :2] *= np.pi / 180.0
        s.t0 = r[:, -1, 3].astype(np.float32)
        amp = r[:, :, 4]
        if angles:
            comps = np.stack(angle2moment(r[:, :, 5], r[:, :, 6], r[:, :, 7]))  # (6, n, ntw)
        else:
            comps = r[:, :, 5:5 + ncomp].transpose(2, 0, 1)
        s.comps = (comps * amp[None]).reshape(ncomp, n * ntw).astype(np.float32)
        out.append(SourceSet(s, axis))
    return out[0], out[1], hyper


def _nearest(c, v):
    return int(np.argmin(np.abs(np.asarray(c, np.float64) - v)))


def locate_sources(ss, x, y, z, hyper):
    """Fill indx (1-based ghosted) and shift of a SourceSet on the grid, and
    the located radius in its axis; a point outside the interior raises."""
    x, y, z = (np.asarray(a, np.float32).astype(np.float64) for a in (x, y, z))
    nk2 = len(z) - LENFD - 1
    n = ss.src.n
    indx = np.zeros((n, 3), np.int32)
    shift = np.zeros((n, 3), np.float32)
    for q in range(n):
        x0, y0, z0 = (float(np.float32(v)) for v in ss.axis[q])
        if not (x.min() <= x0 <= x.max() and y.min() <= y0 <= y.max()):
            raise ValueError(f"source {q + 1} at {ss.axis[q]} is out of the computational domain")
        i, j = _nearest(x, x0), _nearest(y, y0)
        if z.min() <= z0 <= z.max():
            k = _nearest(z, z0)
        elif z0 > hyper:
            k = nk2
            z0 = z[k]
        else:
            raise ValueError(f"source {q + 1} at {ss.axis[q]} is out of the computational domain")
        if not (LENFD <= i < len(x) - LENFD and LENFD <= j < len(y) - LENFD and LENFD <= k <= nk2):
            raise ValueError(f"source {q + 1} at {ss.axis[q]} is out of the computational domain")
        for d, (c, v, p) in enumerate(((x, x0, i), (y, y0, j), (z, z0, k))):
            shift[q, d] = (v - c[p]) / (c[p + 1] - c[p]) if v > c[p] else (v - c[p]) / (c[p] - c[p - 1])
        indx[q] = (i + 1, j + 1, k + 1)
        ss.axis[q, 2] = z0
    ss.src.indx = indx
    ss.src.shift = shift
    return ss


# ---------------------------------------------------------------------------
# stations
# ---------------------------------------------------------------------------
@dataclass
class Stations:
    xyz: np.ndarray  # (n, 3): colatitude, longitude (rad), radius (m)
    ids: np.ndarray  # (n, 2): (receiver number, 0) or (point on line, line number)
    tinv: int
    hyper: float


def read_stations(path):
    """io_pt_read (mod_io.F90:112-152): recv_NNN = x y z (degrees, degrees,
    m); line_NNN = x0 y0 z0 dx dy dz n."""
    conf = fio.read_conf(path)
    nrec = conf.get("number_of_recv", 2, int)
    nline = conf.get("number_of_inline", 2, int)
    tinv = conf.get("tinv_of_seismo", 2, int)
    hyper = conf.get("topo_hyper_height", 2, float)
    pts, ids = [], []
    for n in range(1, nrec + 1):
        pts.append(conf.row(f"recv_{n:03d}", 2, 3))
        ids.append((n, 0))
    for n in range(1, nline + 1):
        v = conf.row(f"line_{n:03d}", 2, 6)
        m = conf.get(f"line_{n:03d}", 8, int)
        for q in range(m):
            pts.append([v[d] + v[3 + d] * q for d in range(3)])
            ids.append((q + 1, n))
    xyz = np.array(pts, np.float64).reshape(-1, 3)
    xyz[:, :2] *= np.pi / 180.0
    return Stations(xyz.astype(np.float32), np.array(ids, np.int32).reshape(-1, 2), tinv, hyper)


def locate_stations(st, x, y, z):
    """receiver_locate: nearest node (k = nk2 above topo_hyper_height);
    points outside the interior are left out.  Returns (indx 1-based
    ghosted, gindx 1-based interior, coord, grid, ids)."""
    x, y, z = (np.asarray(a, np.float32) for a in (x, y, z))
    nx, ny, nz = len(x), len(y), len(z)
    nk2 = nz - LENFD - 1
    keep = []
    for q, (x0, y0, z0) in enumerate(st.xyz):
        i, j = int(np.argmin(np.abs(x - x0))), int(np.argmin(np.abs(y - y0)))
        k = int(np.argmin(np.abs(z - z0))) if z0 <= st.hyper else nk2
        if LENFD <= i < nx - LENFD and LENFD <= j < ny - LENFD and LENFD <= k <= nk2:
            keep.append((q, i, j, k))
    indx = np.array([(i + 1, j + 1, k + 1) for _, i, j, k in keep], np.int32).reshape(-1, 3)
    gindx = indx - LENFD
    coord = st.xyz[[q for q, *_ in keep]].reshape(-1, 3)
    grid = np.array([(x[i], y[j], z[k]) for _, i, j, k in keep], np.float32).reshape(-1, 3)
    ids = st.ids[[q for q, *_ in keep]].reshape(-1, 2)
    return indx, gindx, coord, grid, ids


# ---------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------
def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        raise SystemExit("usage: python -m fwant_amd.prep SeisFD3D.conf [grid] [metric] [media] [source] [station]")
    path, steps = argv[0], argv[1:] or ["grid", "metric", "media", "source", "station"]
    conf = fio.read_conf(path)
    base = os.path.dirname(os.path.abspath(path))
    rel = lambda k: os.path.join(base, conf.get(k))  # noqa: E731
    ni, nj, nk = (conf.get(k, 2, int) for k in ("ni", "nj", "nk"))
    gdir = rel("GRID_ROOT")
    if "grid" in steps:
        os.makedirs(gdir, exist_ok=True)
        x, y, z = grid_coords(rel("GRID_CONF"), ni, nj, nk)
        fio.write_coord(fio._fnm(gdir, "coord"), x, y, z)
        print(f"grid: {len(x)} x {len(y)} x {len(z)} points")
    c = fio._read(fio._fnm(gdir, "coord"), ("x", "y", "z"))
    x, y, z = (c[k].astype(np.float64) for k in "xyz")
    if "metric" in steps:
        from .solver import build_metric  # the metric is built on the GPU in FP64
        m = models.Model("grid", ni, nj, nk, conf.get("stept", 2, float), 1, True, x, y, z,
                         *(np.ones((ni + 6, nj + 6, nk + 6), np.float32, order="F") for _ in range(3)))
        fio.write_metric(fio._fnm(gdir, "metric"), build_metric(m), ni + 6, nj + 6, nk + 6)
        print("metric written")
    if "media" in steps:
        mc = read_media_conf(rel("MEDIA_CONF"))
        lam, mu, rho = effective_media(mc, x, y, z)
        os.makedirs(rel("MEDIA_ROOT"), exist_ok=True)
        fio.write_media(fio._fnm(rel("MEDIA_ROOT"), "media"), lam, mu, rho)
        cf = cfl_check(x, y, z, lam, mu, rho)
        print(f"Maximum allowed time step is {cf.dtmax:.6g} located on {cf.indx}; Vp and dL are {cf.vp:.6g} "
              f"{cf.length:.6g}")
        stept = conf.get("stept", 2, float)
        if cf.dtmax < stept:
            print(f"Serious Error: stept>dtmax {stept} {cf.dtmax}")
    if "source" in steps:
        f, mo, hyper = read_src_conf(rel("SOURCE_CONF"))
        for ss in (f, mo):
            locate_sources(ss, x, y, z, hyper)
        os.makedirs(rel("SOURCE_ROOT"), exist_ok=True)
        fio.write_source(fio._fnm(rel("SOURCE_ROOT"), "source"), f.src, mo.src, f.axis, mo.axis)
        print(f"sources: {f.src.n} forces, {mo.src.n} moments")
    if "station" in steps:
        st = read_stations(path)
        indx, gindx, coord, grid, ids = locate_stations(st, x, y, z)
        os.makedirs(rel("STATION_ROOT"), exist_ok=True)
        fio.write_station(fio._fnm(rel("STATION_ROOT"), "station"), indx, st.tinv, gindx, coord, grid, ids)
        print(f"stations: {len(indx)} of {len(st.xyz)} located")


if __name__ == "__main__":
    main()
