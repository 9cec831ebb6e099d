"""Sensitivity-kernel time integral on the GPU (host mirror of
code/srcF.kernel/SI_ker_sta.F90's numeric subroutines).

Functions take device arrays (fwant_amd._lib.DeviceArray, allocated through
the library's own HIP runtime) with the reference's series semantics; the
layout is [component][sample][point] (point fastest), i.e. the
transpose-free form of the reference's (i, j, k, n) arrays.  Work runs on
the default stream; the synchronous copies of DeviceArray order it.

  stress2strain(E, lam, mu)                 SI_ker_sta.F90:1185-1211
  sgt2scatter(R, S, lam, mu, dt, param)     :1257-1337 (one receiver component;
                                            param "lamu": SI_ker_sta_lamu.F90)
  vel2scatter(R, S, rho, dt)                SI_ker_sta_rho.F90:1478-1552
  eoshift(E, shift)                         SI_ker_EQ.F90:574-587
  reparam(mode, lam, mu, Kap, Kaq, Kbp, Kbq)  SI_ker_{lamu2kamu,lamu2vpvs,vpvs2kamu}.F90
  scatter_filter(w0, b, a) / scatter_filtfilt   :1433-1475
  scatter2kernel(V, U, Sa, Sb, nt1, nt2, dt, M0)  :1375-1431
  scatter_rotate(X, Y, Z, rot)              :1354-1373
  sgt_decode(fmt, codes, scales, subc, nslab, out)   packed BF16/FP16 SGT -> FP32
  pair_kernel(...)                          the per-block pipeline of :574-798
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, kcheck


def _ptr(t):
    if not isinstance(t, DeviceArray):
        raise ValueError("DeviceArray required")
    return C.c_void_p(t.ptr)


def _series(t):
    """(ncomp, knt, npt) series -> base pointer, component stride, time stride."""
    nc, knt, npt = t.shape
    cs, ts = t.strides if t.strides is not None else (knt * npt, npt)
    return _ptr(t), cs, ts, knt, npt


def snapshot_series(ptr, nrec, nf, npts):
    """The records of a snapshot ([record][field][point] storage) as a
    (field, sample, point) series view, without copying."""
    return DeviceArray.wrap(ptr, (nf, nrec, npts), np.float32, strides=(npts, nf * npts))


def _stream():
    return None


def stress2strain(E, lam, mu):
    """In-place stress -> strain of a (6, knt, npt) float32 series."""
    L = _lib.lib()
    p, cs, ts, knt, npt = _series(E)
    kcheck(L.fwk_stress2strain(p, cs, ts, npt, knt, _ptr(lam), _ptr(mu), _stream()), "fwk_stress2strain")
    return E


def sgt2scatter(R, S, lam, mu, dt, param="vpvs"):
    """Scattered series (knt, npt) of one receiver component: ("vpvs")
    alpha and beta (SI_ker_sta.F90:1290-1331) or ("lamu") lambda and mu
    (SI_ker_sta_lamu.F90:1290-1336)."""
    L = _lib.lib()
    rp, rcs, rts, knt, npt = _series(R)
    sp, scs, sts, knt2, npt2 = _series(S)
    if (knt, npt) != (knt2, npt2):
        raise ValueError("R and S shapes differ")
    A = DeviceArray((knt, npt))
    B = DeviceArray((knt, npt))
    kcheck(L.fwk_scatter_param({"vpvs": 0, "lamu": 1}[param], rp, rcs, rts, sp, scs, sts, npt, knt, float(dt),
                               _ptr(lam), _ptr(mu), _ptr(A), _ptr(B), _stream()), "fwk_scatter_param")
    return A, B


def vel2scatter(R, S, rho, dt):
    """Density scattered series (knt, npt) from 3-component velocity series
    (SI_ker_sta_rho.F90:1478-1552)."""
    L = _lib.lib()
    rp, rcs, rts, knt, npt = _series(R)
    sp, scs, sts, knt2, npt2 = _series(S)
    if (knt, npt) != (knt2, npt2) or R.shape[0] < 3 or S.shape[0] < 3:
        raise ValueError("R and S must be (3, knt, npt) series of the same shape")
    out = DeviceArray((knt, npt))
    kcheck(L.fwk_vscatter(rp, rcs, rts, sp, scs, sts, npt, knt, float(dt), _ptr(rho), _ptr(out), _stream()),
           "fwk_vscatter")
    return out


def eoshift(E, shift):
    """Fortran eoshift along time of a (ncomp, knt, npt) series in place (SI_ker_EQ.F90:574-587)."""
    p, cs, ts, knt, npt = _series(E)
    kcheck(_lib.lib().fwk_eoshift(p, cs, ts, npt, knt, E.shape[0], int(shift), _stream()), "fwk_eoshift")
    return E


def reparam(mode, lam, mu, Kap, Kaq, Kbp, Kbq):
    """In-place kernel re-parameterisation: "lamu2kamu", "lamu2vpvs" or "vpvs2kamu"."""
    m = {"lamu2kamu": 0, "lamu2vpvs": 1, "vpvs2kamu": 2}[mode]
    kcheck(_lib.lib().fwk_reparam(m, Kap.numel(), _ptr(lam), _ptr(mu), _ptr(Kap), _ptr(Kaq), _ptr(Kbp), _ptr(Kbq),
                                  _stream()), "fwk_reparam")


def _filter(mode, w0, b, a):
    L = _lib.lib()
    knt, npt = w0.shape
    b = np.ascontiguousarray(b, np.float64)
    a = np.ascontiguousarray(a, np.float64)
    if len(a) != len(b):
        raise ValueError("filter a and b must have the same length (NOD1)")
    w = DeviceArray(w0.shape)
    kcheck(L.fwk_filter(mode, npt, knt, len(b), _lib.dptr(b), _lib.dptr(a), _ptr(w0), _ptr(w), _stream()),
           "fwk_filter")
    return w


def scatter_filter(w0, b, a):
    """P1: causal IIR with zero front padding and odd end padding."""
    return _filter(1, w0, b, a)


def scatter_filtfilt(w0, b, a):
    """P2: zero-phase forward-backward IIR with odd padding (NFACT = 3*NOD)."""
    return _filter(2, w0, b, a)


def scatter2kernel(V, U, Sa, Sb, nt1, nt2, dt, M0):
    """Phase/amplitude kernels (Kap, Kaq, Kbp, Kbq), each (npt,); V, U host series."""
    L = _lib.lib()
    knt, npt = Sa.shape
    V = np.ascontiguousarray(V, np.float32)
    U = np.ascontiguousarray(U, np.float32)
    out = [DeviceArray((npt,)) for _ in range(4)]
    kcheck(L.fwk_scatter2kernel(npt, knt, int(nt1), int(nt2), float(dt), float(M0), _lib.fptr(V), _lib.fptr(U),
                                _ptr(Sa), _ptr(Sb), *[_ptr(o) for o in out], _stream()), "fwk_scatter2kernel")
    return out


def scatter_rotate(X, Y, Z, rot):
    """In-place rotation of three series by the 3x3 matrix rot (rot[r, c])."""
    L = _lib.lib()
    r = np.ascontiguousarray(np.asarray(rot, np.float32).ravel(order="F"))
    kcheck(L.fwk_rotate(X.numel(), _ptr(X), _ptr(Y), _ptr(Z), _lib.fptr(r), _stream()), "fwk_rotate")


def filter_series_host(x, b, a, mode):
    """syn_filter (mode 1) / syn_filtfilt (mode 2) of one host series
    (SI_ker_sta.F90:1477-1557): FP64 IIR with the same padding as the
    scattered series."""
    x = np.asarray(x, np.float64)
    b = np.asarray(b, np.float64)
    a = np.asarray(a, np.float64)
    nod1 = len(b)
    nfact = 3 * (nod1 - 1)
    knt = len(x)
    end = 2 * x[-1] - x[knt - 1 - np.arange(1, nfact + 1)]
    front = np.zeros(nfact) if mode == 1 else (2 * x[0] - x[1:nfact + 1])[::-1]
    v = np.concatenate([front, x, end])

    def iir(u):
        y = np.zeros_like(u)
        for n in range(len(u)):
            acc = b[0] * u[n]
            for m in range(1, min(nod1, n + 1)):
                acc = acc + b[m] * u[n - m] - a[m] * y[n - m]
            y[n] = acc / a[0]
        return y

    y = iir(v) if mode == 1 else iir(iir(v)[::-1])[::-1]
    return y[nfact:nfact + knt].astype(np.float32)


def sgt_decode(fmt, codes, scales, subc, nslab, out):
    """Restore packed SGT slabs (fw_add_snapshot_packed layout) to FP32:
    codes, scales and out are device pointers or DeviceArrays."""
    f = _lib.SGT_FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
    p = lambda t: C.c_void_p(t) if isinstance(t, int) else _ptr(t)  # noqa: E731
    kcheck(_lib.lib().fwk_sgt_decode(f, p(codes), p(scales), int(subc[0]), int(subc[1]), int(subc[2]), int(nslab),
                                     p(out), _stream()), "fwk_sgt_decode")


def pair_kernel(Rx, Ry, Rz, S, lam, mu, dt, V, U, nt1, nt2, M0, comp, filt=None):
    """One receiver-component kernel of a station-event pair (SI_ker_sta.F90
    :574-798): stress SGTs -> strain, scattered series, optional filter
    (("P1"|"P2", b, a)), trapezoid kernel.  Rx/Ry/Rz/S are (6, knt, npt)
    stress series (Rx..Rz may be None when unused); comp in {0, 1, 2}
    selects the receiver component (x, y, z)."""
    R = (Rx, Ry, Rz)[comp]
    Rs = stress2strain(R.copy(), lam, mu)
    Ss = stress2strain(S.copy(), lam, mu)
    A, B = sgt2scatter(Rs, Ss, lam, mu, dt)
    if filt is not None:
        kind, b, a = filt
        mode = 1 if kind == "P1" else 2
        A, B = _filter(mode, A, b, a), _filter(mode, B, b, a)
        V, U = filter_series_host(V, b, a, mode), filter_series_host(U, b, a, mode)
    return scatter2kernel(V, U, A, B, nt1, nt2, dt, M0)
