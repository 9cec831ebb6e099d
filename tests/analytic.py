"""Closed-form elastodynamic solutions for the whole-solver pins of the
oracle (test infrastructure; nothing in the product imports this).

  - stokes_*: Stokes' solution for a point force and for a point moment
    tensor in a homogeneous, isotropic full space (Aki & Richards,
    Quantitative Seismology, 2nd ed., eqs. 4.23 and 4.29), with the near-,
    intermediate- and far-field terms;
  - halfspace_*: the response at the free surface of a homogeneous
    half-space to a vertical point force at depth d >= 0 (Lamb's problem
    for d = 0), by frequency-wavenumber integration of the exact P-SV
    solution: for each (omega, k) the 4x4 motion-stress system of a
    homogeneous medium is solved with the source jump and the
    traction-free surface, the wavenumber integral is a sum over a
    discrete wavenumber set (Bouchon's method: sources repeated every L
    along the surface, so the Hankel integral becomes a series), and the
    frequency integral an inverse FFT with complex frequency omega - i*w_I,
    whose damping exp(-w_I t) is undone afterwards.

Source-time functions follow the reference's Ricker (mod_src.F90:509-530):
R(t) = (u^2/4 - 1/2) exp(-u^2/4) sqrt(pi)/2, u = 2 pi f0 (t - tc).
"""
import numpy as np

SQPI2 = np.sqrt(np.pi) / 2.0


def ricker(t, f0, tc, d=0):
    """Ricker of the reference and its first (d=1) / second (d=2) time derivative."""
    w = 2.0 * np.pi * f0
    u = w * (np.asarray(t, np.float64) - tc)
    e = np.exp(-u * u / 4.0)
    if d == 0:
        return (u * u / 4.0 - 0.5) * e * SQPI2
    if d == 1:
        return w * (u / 2.0) * (1.5 - u * u / 4.0) * e * SQPI2
    if d == 2:
        return w * w * (0.75 - 0.75 * u * u + u ** 4 / 16.0) * e * SQPI2
    raise ValueError(d)


_GL_X, _GL_W = np.polynomial.legendre.leggauss(96)


def _near(g, t, ta, tb):
    """int_{ta}^{tb} tau g(t - tau) dtau for every t (Gauss-Legendre, 96 nodes)."""
    tau = 0.5 * (tb - ta) * _GL_X + 0.5 * (tb + ta)
    w = 0.5 * (tb - ta) * _GL_W
    return (g(t[:, None] - tau[None, :]) * (tau * w)[None, :]).sum(1)


def stokes_force(x, F, t, alpha, beta, rho, g):
    """Displacement u_i(x, t) (shape [nt, 3]) at offset x (3,) from a point
    force F (3,) g(t) at the origin; eq. 4.23.  For the velocity pass the
    time derivative of the force's time function as g."""
    r = float(np.linalg.norm(x))
    gam = np.asarray(x, np.float64) / r
    GG = np.outer(gam, gam)
    I = np.eye(3)
    ta, tb = r / alpha, r / beta
    near = _near(g, t, ta, tb) / (4 * np.pi * rho * r ** 3)
    fp = g(t - ta) / (4 * np.pi * rho * alpha ** 2 * r)
    fs = g(t - tb) / (4 * np.pi * rho * beta ** 2 * r)
    F = np.asarray(F, np.float64)
    return (near[:, None] * ((3 * GG - I) @ F)[None, :] + fp[:, None] * (GG @ F)[None, :]
            - fs[:, None] * ((GG - I) @ F)[None, :])


def stokes_moment(x, M, t, alpha, beta, rho, g, gd):
    """Displacement u_n(x, t) ([nt, 3]) at offset x from a point moment tensor
    M (3x3, symmetric) with time function g (gd its derivative); eq. 4.29.
    For the velocity pass (g', g'')."""
    r = float(np.linalg.norm(x))
    c = np.asarray(x, np.float64) / r
    I = np.eye(3)
    M = np.asarray(M, np.float64)
    ggg = np.einsum("n,p,q->npq", c, c, c)
    gd_pq = np.einsum("n,pq->npq", c, I)
    gp_nq = np.einsum("p,nq->npq", c, I)
    gq_np = np.einsum("q,np->npq", c, I)
    NN = 15 * ggg - 3 * gd_pq - 3 * gp_nq - 3 * gq_np
    NIP = 6 * ggg - gd_pq - gp_nq - gq_np
    NIS = -(6 * ggg - gd_pq - gp_nq - 2 * gq_np)
    NFP = ggg
    NFS = -(ggg - gq_np)  # -(g_n g_p - d_np) g_q
    cN, cIP, cIS, cFP, cFS = (np.einsum("npq,pq->n", T, M) for T in (NN, NIP, NIS, NFP, NFS))
    ta, tb = r / alpha, r / beta
    k = 1.0 / (4 * np.pi * rho)
    near = k * _near(g, t, ta, tb) / r ** 4
    ip = k * g(t - ta) / (alpha ** 2 * r ** 2)
    is_ = k * g(t - tb) / (beta ** 2 * r ** 2)
    fp = k * gd(t - ta) / (alpha ** 3 * r)
    fs = k * gd(t - tb) / (beta ** 3 * r)
    return (near[:, None] * cN + ip[:, None] * cIP + is_[:, None] * cIS + fp[:, None] * cFP
            + fs[:, None] * cFS)


# ---------------------------------------------------------------------------
# Vertical point force in a homogeneous half-space (or full space), by
# frequency-wavenumber integration.  z points down, the free surface is z = 0;
# time dependence exp(i w t), complex frequency w = w_R - i w_I.
#
# Potentials u = grad(phi) + curl curl(psi e_z) with phi = Phi(z) J0(k r),
# psi = Psi(z) J0(k r) give, per wavenumber k (Hankel transforms of order 0
# for u_z and sigma_zz, of order 1 for u_r and sigma_rz):
#   U_z = Phi' + k^2 Psi,         U_r = -k (Phi + Psi')
#   T_z = mu (2k^2 - kb^2) Phi + 2 mu k^2 Psi'
#   T_r = -mu k (2 Phi' + (2k^2 - kb^2) Psi)
# with ka = w/alpha, kb = w/beta, Phi'' = na^2 Phi, na^2 = k^2 - ka^2 (nb
# likewise), Re(na), Re(nb) > 0.  A unit point force F e_z at depth d is the
# jump [T_z] = T_z(d+) - T_z(d-) = -F / (2 pi) (the order-0 Hankel
# transform of delta(x) delta(y) is 1 / (2 pi)); U_z, U_r, T_r are continuous.
# A downward load F on the free surface (d = 0) is T_z(0) = -F / (2 pi).
# ---------------------------------------------------------------------------
def _basis(kind, s, E, k, kb2, mu):
    """(U_z, U_r, T_z, T_r) of one exponential solution with value E and
    z-derivative s E of its potential (kind 'P': Phi, 'S': Psi)."""
    q = 2 * k * k - kb2
    if kind == "P":
        return np.stack([s * E, -k * E, mu * q * E, -2 * mu * k * s * E])
    return np.stack([k * k * E, -k * s * E, 2 * mu * k * k * s * E, -mu * k * q * E])


def fk_vertical_force(k, w, d, zr, alpha, beta, rho, surface=True):
    """Per-(w, k) transforms (U_z, U_r) [nw, nk] at receiver depth zr of a
    unit downward point force at depth d; half-space when surface, else full
    space.  zr <= d (receiver above or at the source depth)."""
    mu = rho * beta ** 2
    K = np.asarray(k, np.float64)[None, :]
    W = np.asarray(w)[:, None]
    kb2 = (W / beta) ** 2
    na = np.sqrt(K * K - (W / alpha) ** 2 + 0j)
    nb = np.sqrt(K * K - kb2 + 0j)
    na = np.where(na.real < 0, -na, na)
    nb = np.where(nb.real < 0, -nb, nb)
    shape = np.broadcast(K, W).shape
    one = np.ones(shape, complex)
    f = -1.0 / (2.0 * np.pi)
    if d == 0.0:
        # region z >= 0: down-going a3 (P), b3 (S); traction at z = 0
        P = _basis("P", -na, one, K, kb2, mu)
        S = _basis("S", -nb, one, K, kb2, mu)
        A = np.stack([np.stack([P[2], S[2]], -1), np.stack([P[3], S[3]], -1)], -2)
        rhs = np.stack([f * one, 0 * one], -1)[..., None]
        a3, b3 = np.moveaxis(np.linalg.solve(A, rhs)[..., 0], -1, 0)
        ea, eb = np.exp(-na * zr), np.exp(-nb * zr)
        v = a3 * _basis("P", -na, ea, K, kb2, mu) + b3 * _basis("S", -nb, eb, K, kb2, mu)
        return v[0], v[1]
    ead, ebd = np.exp(-na * d), np.exp(-nb * d)
    # unknowns x = (a1, a2, b1, b2, a3, b3):
    #   region A (z < d): a1 e^{-na z} + a2 e^{na (z-d)}, b1 e^{-nb z} + b2 e^{nb (z-d)}
    #   region B (z > d): a3 e^{-na (z-d)}, b3 e^{-nb (z-d)}
    # without the surface, region A holds only the up-going a2, b2 (a1 = b1 = 0)
    zero = np.zeros(shape, complex)
    at_d_A = [_basis("P", -na, ead, K, kb2, mu), _basis("P", na, one, K, kb2, mu),
              _basis("S", -nb, ebd, K, kb2, mu), _basis("S", nb, one, K, kb2, mu)]
    at_d_B = [_basis("P", -na, one, K, kb2, mu), _basis("S", -nb, one, K, kb2, mu)]
    rows = []
    rhs = []
    for r in range(4):  # B - A at z = d
        rows.append([-at_d_A[0][r], -at_d_A[1][r], -at_d_A[2][r], -at_d_A[3][r], at_d_B[0][r], at_d_B[1][r]])
        rhs.append(f * one if r == 2 else zero)
    if surface:
        at_0 = [_basis("P", -na, one, K, kb2, mu), _basis("P", na, ead, K, kb2, mu),
                _basis("S", -nb, one, K, kb2, mu), _basis("S", nb, ebd, K, kb2, mu)]
        for r in (2, 3):
            rows.append([at_0[0][r], at_0[1][r], at_0[2][r], at_0[3][r], zero, zero])
            rhs.append(zero)
    else:
        rows.append([one, zero, zero, zero, zero, zero])
        rows.append([zero, zero, one, zero, zero, zero])
        rhs += [zero, zero]
    A = np.stack([np.stack(rr, -1) for rr in rows], -2)
    x = np.linalg.solve(A, np.stack(rhs, -1)[..., None])[..., 0]
    a1, a2, b1, b2 = (x[..., q] for q in range(4))
    v = (a1 * _basis("P", -na, np.exp(-na * zr), K, kb2, mu) + a2 * _basis("P", na, np.exp(na * (zr - d)), K, kb2, mu)
         + b1 * _basis("S", -nb, np.exp(-nb * zr), K, kb2, mu) + b2 * _basis("S", nb, np.exp(nb * (zr - d)), K, kb2, mu))
    return v[0], v[1]


def surface_static(alpha, beta, rho):
    """Large-k limits k U_z, k U_r of the surface response to a unit downward
    surface load (every frequency): Boussinesq's u_z = (1 - nu) / (2 pi mu r),
    u_r = -(1 - 2 nu) / (4 pi mu r)."""
    mu = rho * beta ** 2
    lam = rho * alpha ** 2 - 2 * mu
    nu = lam / (2 * (lam + mu))
    return (1 - nu) / (2 * np.pi * mu), -(1 - 2 * nu) / (4 * np.pi * mu)


def fk_time_series(kern, rs, t, g, T, dw_frac=0.25, fmax=None, kmax=None, nk=None, static=None):
    """Time series of u_z(r, t) and u_r(r, t) for the radial distances rs,
    for a force time function g(t) (vectorised), from kern(k, w) -> (Uz, Ur)
    per unit force.  T: length of the time window (sets the damping w_I =
    pi / T and the frequency step); fmax: highest frequency (Hz); kmax, nk:
    wavenumber range and count (midpoint rule in k: the discrete-wavenumber
    sum).  static = (Sz, Sr): large-k limits k U_z, k U_r (a surface source
    seen at the surface), subtracted from the kernel under the sum and added
    back exactly (Hankel transforms of 1/k: 1/r for both orders)."""
    from scipy.special import j0, j1
    wI = np.pi / T
    dw = dw_frac * 2 * np.pi / T
    wR = np.arange(0.0, 2 * np.pi * fmax + dw, dw)
    w = wR - 1j * wI
    # spectrum of g at complex frequency (trapezoid over a fine time grid)
    ts = np.linspace(min(0.0, t[0]), T, 8 * int(T * fmax * 8) + 1)
    dts = ts[1] - ts[0]
    G = (g(ts)[None, :] * np.exp(-1j * w[:, None] * ts[None, :])).sum(1) * dts
    dk = kmax / nk
    k = (np.arange(nk) + 0.5) * dk  # midpoint rule (the k = 0 term carries weight k = 0)
    wk = np.full(nk, dk)
    uz = np.zeros((len(rs), len(t)))
    ur = np.zeros((len(rs), len(t)))
    for q0 in range(0, len(w), 16):
        ws = w[q0:q0 + 16]
        Uz, Ur = kern(k, ws)
        if static is not None:
            Uz = Uz - static[0] / k[None, :]
            Ur = Ur - static[1] / k[None, :]
        for n, r in enumerate(rs):
            Z = (Uz * (j0(k * r) * k * wk)[None, :]).sum(1) * G[q0:q0 + 16]
            R = (Ur * (j1(k * r) * k * wk)[None, :]).sum(1) * G[q0:q0 + 16]
            ph = np.exp(1j * wR[q0:q0 + 16, None] * t[None, :])
            fac = np.where(wR[q0:q0 + 16] == 0, 0.5, 1.0)[:, None] * dw / np.pi
            uz[n] += (Z[:, None] * ph * fac).real.sum(0)
            ur[n] += (R[:, None] * ph * fac).real.sum(0)
    damp = np.exp(wI * t)[None, :]
    uz, ur = uz * damp, ur * damp
    if static is not None:
        gt = g(t)
        for n, r in enumerate(rs):
            uz[n] += static[0] / r * gt
            ur[n] += static[1] / r * gt
    return uz, ur
