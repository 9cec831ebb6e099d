"""ctypes front-end of the CPU oracle (oracle/fw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


# ORACLE_SANITIZE=1: load the AddressSanitizer / UndefinedBehaviorSanitizer
# build (make san); the process must have libasan preloaded.
SANITIZE = os.environ.get("ORACLE_SANITIZE") == "1"


def build():
    subprocess.run(["make", "-s", "-C", HERE] + (["san"] if SANITIZE else []), check=True)


def _load(prec):
    path = os.path.join(BUILD, f"liborc{prec}{'_san' if SANITIZE else ''}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    R = C.c_float if prec == 32 else C.c_double
    rp = C.POINTER(R)
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    p = f"orc{prec}_"
    sig = {
        "create": (vp, [C.c_int, C.c_int, C.c_int, R, C.c_int, C.c_int]),
        "destroy": (None, [vp]),
        "set_grid": (None, [vp] + [rp] * 8),
        "build_metric": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp]),
        "set_media": (None, [vp, rp, rp, rp]),
        "set_pml": (None, [vp, ip, rp, rp, rp]),
        "get_pml_profiles": (None, [vp, rp, rp, rp]),
        "get_freenode": (C.c_int, [vp]),
        "set_force": (None, [vp, C.c_int, C.c_int, C.c_int, ip, rp, rp, rp, rp, rp, rp, rp]),
        "set_moment": (None, [vp, C.c_int, C.c_int, C.c_int, ip, rp, rp, rp, rp, rp]),
        "set_receivers": (None, [vp, C.c_int, C.c_int, ip]),
        "step": (None, [vp, C.c_int]),
        "run": (C.c_int, [vp, C.c_int, C.c_int, rp]),
        "get_field": (None, [vp, C.c_int, rp]),
        "set_cycle_origin": (None, [vp, C.c_int]),
        "set_field": (None, [vp, C.c_int, rp]),
        "ncell": (C.c_long, [vp]),
        "rhs": (None, [vp, C.c_int, C.c_int, C.c_int, rp]),
        "rk_stage": (None, [C.c_int, R, R, R, C.c_long, rp, rp, rp, rp]),
        "rk_coef": (None, [rp, rp]),
        "op_apply": (R, [C.c_int, C.c_int, rp]),
        "stf": (R, [C.c_int, C.c_int, R, R, R, C.c_int, R]),
        "norm_delt": (None, [rp, R, R, R]),
        "stage_pattern": (None, [C.c_int, C.c_int, ip]),
        "stress2strain": (None, [C.c_long, C.c_int, rp, rp, rp, rp, rp, rp, rp, rp]),
        "scatter1": (None, [C.c_long, C.c_int, R, C.POINTER(rp), C.POINTER(rp), rp, rp, rp, rp]),
        "series_filter": (None, [C.c_int, C.c_long, C.c_int, C.c_int, dp, dp, rp, rp]),
        "scatter2kernel": (None, [C.c_long, C.c_int, C.c_int, C.c_int, R, R, rp, rp, rp, rp, rp, rp, rp, rp]),
        "scatter_rotate": (None, [C.c_long, C.c_int, rp, rp, rp, rp]),
        "scatter1_lamu": (None, [C.c_long, C.c_int, R, C.POINTER(rp), C.POINTER(rp), rp, rp, rp, rp]),
        "vscatter1": (None, [C.c_long, C.c_int, R, C.POINTER(rp), C.POINTER(rp), rp, rp]),
        "eoshift": (None, [C.c_long, C.c_int, C.c_int, rp, rp]),
        "reparam": (None, [C.c_int, C.c_long, rp, rp, rp, rp]),
        "assemble": (None, [C.c_long, vp, vp, vp, vp, vp]),
        "lsqr": (C.c_long, [C.c_long, C.c_long, vp, vp, vp, vp, C.c_double, C.c_double, C.c_long, vp, vp, vp]),
        "sgt_encode": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, vp, vp, vp]),
        "sgt_decode": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, vp, vp, vp]),
        "f32_to_bf16": (C.c_uint16, [C.c_float]),
        "f32_to_fp16": (C.c_uint16, [C.c_float]),
        "fp16_to_f32": (C.c_float, [C.c_uint16]),
        "set_mutation": (None, [C.c_int]),
        "src_terms": (None, [vp, C.c_int, R, rp, rp]),
        "build_metric3d": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
        "set_metric3d": (None, [vp, rp, rp, rp, rp, rp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, p + name)
        f.restype = res
        f.argtypes = args
    L.orc_filter.restype = None
    L.orc_filter.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
    return L, R, p


_LIBS = {}


class Lib:
    def __init__(self, prec=32):
        if prec not in _LIBS:
            _LIBS[prec] = _load(prec)
        self.L, self.R, self.p = _LIBS[prec]
        self.prec = prec
        self.dtype = np.float32 if prec == 32 else np.float64

    def __getattr__(self, name):
        return getattr(self.L, self.p + name)

    def ptr(self, a):
        assert a.dtype == self.dtype and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
        return a.ctypes.data_as(C.POINTER(self.R))

    def arr(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=self.dtype).ravel(order="K"))


def sgt_encode(fmt, x, subc):
    """Oracle packing of FP32 SGT slabs x[..., npts] (fmt "bf16" | "fp16"):
    returns codes (uint16, same shape) and scales [..., nblocks]."""
    f = {"bf16": 1, "fp16": 2}[fmt]
    x = np.ascontiguousarray(x, np.float32)
    npts = int(np.prod(subc))
    assert x.shape[-1] == npts
    nslab = x.size // npts
    nblk = int(np.prod([(c + 15) // 16 for c in subc]))
    codes = np.zeros(x.shape, np.uint16)
    scales = np.zeros(x.shape[:-1] + (nblk,), np.float32)
    Lib(32).sgt_encode(f, int(subc[0]), int(subc[1]), int(subc[2]), nslab, x.ctypes.data, codes.ctypes.data,
                       scales.ctypes.data)
    return codes, scales


def sgt_decode(fmt, codes, scales, subc):
    f = {"bf16": 1, "fp16": 2}[fmt]
    codes = np.ascontiguousarray(codes, np.uint16)
    scales = np.ascontiguousarray(scales, np.float32)
    npts = int(np.prod(subc))
    out = np.zeros(codes.shape, np.float32)
    Lib(32).sgt_decode(f, int(subc[0]), int(subc[1]), int(subc[2]), codes.size // npts, codes.ctypes.data,
                       scales.ctypes.data, out.ctypes.data)
    return out


class mutation:
    """Context manager planting mistake `m` in the oracle (both precisions)
    for a control test; see fw_oracle.c for the list."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        for prec in (32, 64):
            Lib(prec).set_mutation(self.m)
        return self

    def __exit__(self, *a):
        for prec in (32, 64):
            Lib(prec).set_mutation(0)


def build_metric(model):
    """Oracle FP64 metric of seis3d_metric.F90:82-112 (returned as float64)."""
    L = Lib(64).L
    nx, ny, nz = model.nx, model.ny, model.nz
    out = [np.zeros(n) for n in (nx, nx, nx, ny, nz)]
    d = C.POINTER(C.c_double)
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (model.x, model.y, model.z))
    L.orc64_build_metric(nx, ny, nz, x.ctypes.data_as(d), y.ctypes.data_as(d), z.ctypes.data_as(d),
                         *[o.ctypes.data_as(d) for o in out])
    return dict(zip(("xsin", "xcot", "xi_x", "eta_y", "zeta_z"), out))


def build_metric3d(model):
    """Oracle FP64 mode C metric of a terrain-following model (model.r3):
    slopes p, q, zeta_r and 1/r, each float64 (nx, ny, nz)."""
    L = Lib(64).L
    nx, ny, nz = model.nx, model.ny, model.nz
    d = C.POINTER(C.c_double)
    x, y = (np.ascontiguousarray(a, np.float64) for a in (model.x, model.y))
    r3 = np.asfortranarray(model.r3, np.float64)
    out = [np.zeros((nx, ny, nz), np.float64, order="F") for _ in range(4)]
    L.orc64_build_metric3d(nx, ny, nz, x.ctypes.data_as(d), y.ctypes.data_as(d), r3.ctypes.data_as(d),
                           *[o.ctypes.data_as(d) for o in out])
    return dict(zip(("p", "q", "zz", "rr"), out))


class OracleSolver:
    """The oracle solver on a Model (same inputs as fwant_amd.solver.Solver)."""

    def __init__(self, model, prec=32, metric=None, nthreads=None, metric3d=None):
        self.lib = lib = Lib(prec)
        dt = lib.dtype
        self.model = m = model
        nthreads = nthreads or int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))
        self.h = lib.create(m.ni, m.nj, m.nk, m.stept, 1 if m.freenode else 0, nthreads)
        if metric is None:
            metric = build_metric(m)
        self.metric = {k: np.asarray(v) for k, v in metric.items()}
        g = [np.ascontiguousarray(a, dt) for a in (m.x, m.y, m.z, metric["xsin"], metric["xcot"],
                                                   metric["xi_x"], metric["eta_y"], metric["zeta_z"])]
        self._keep = g
        lib.set_grid(self.h, *[lib.ptr(a) for a in g])
        media = [np.asfortranarray(a, dt).ravel(order="F") for a in (m.lam, m.mu, m.rho)]
        lib.set_media(self.h, *[lib.ptr(a) for a in media])
        if getattr(m, "r3", None) is not None:
            # mode C: the oracle's own FP64 metric, or a given one (p, q, zz, rr)
            met3 = build_metric3d(m) if metric3d is None else metric3d
            arrs = [np.asfortranarray(a, dt).ravel(order="F") for a in
                    (m.r3, met3["p"], met3["q"], met3["zz"], met3["rr"])]
            self._keep += arrs
            lib.set_metric3d(self.h, *[lib.ptr(a) for a in arrs])
        an = np.array(m.abs_number, np.int32)
        av, bm, am = (np.array(v, dt) for v in (m.abs_velocity, m.cfs_bmax, m.cfs_amax))
        lib.set_pml(self.h, an.ctypes.data_as(C.POINTER(C.c_int)), lib.ptr(av), lib.ptr(bm), lib.ptr(am))
        ip = C.POINTER(C.c_int)
        f = m.force
        if f.n > 0:
            idx = np.ascontiguousarray(f.indx, np.int32).ravel() - 1
            arrs = [np.ascontiguousarray(a, dt).ravel() for a in (f.shift, f.t0, f.stf_time, f.stf_freq,
                                                                   f.comps[0], f.comps[1], f.comps[2])]
            self._keep += [idx] + arrs
            lib.set_force(self.h, f.n, f.ntw, f.stf_id, idx.ctypes.data_as(ip), *[lib.ptr(a) for a in arrs])
        mo = m.moment
        if mo.n > 0:
            idx = np.ascontiguousarray(mo.indx, np.int32).ravel() - 1
            arrs = [np.ascontiguousarray(a, dt).ravel() for a in (mo.shift, mo.t0, mo.stf_time, mo.stf_freq,
                                                                   mo.comps)]
            self._keep += [idx] + arrs
            lib.set_moment(self.h, mo.n, mo.ntw, mo.stf_id, idx.ctypes.data_as(ip), *[lib.ptr(a) for a in arrs])
        self.npt = len(m.rec_indx)
        if self.npt:
            ri = np.ascontiguousarray(m.rec_indx, np.int32).ravel() - 1
            self._keep.append(ri)
            lib.set_receivers(self.h, self.npt, m.rec_tinv, ri.ctypes.data_as(ip))
        self.ntime = 0
        self.records = []

    def run(self, nsteps):
        lib = self.lib
        out = np.zeros((nsteps + 1, max(self.npt, 1), 9), lib.dtype)
        n = lib.run(self.h, self.ntime, nsteps, lib.ptr(out))
        self.ntime += nsteps
        if self.npt:
            self.records.append(out[:n, : self.npt].copy())
        return out[:n, : self.npt]

    def receivers(self):
        if not self.records:
            return np.zeros((0, self.npt, 9), self.lib.dtype)
        return np.concatenate(self.records, 0)

    def set_cycle_origin(self, origin):
        self.lib.set_cycle_origin(self.h, int(origin))

    def wavefield(self, f):
        from fwant_amd._lib import FIELDS
        fi = FIELDS.index(f) if isinstance(f, str) else int(f)
        m = self.model
        out = np.zeros(m.nx * m.ny * m.nz, self.lib.dtype)
        self.lib.get_field(self.h, fi, self.lib.ptr(out))
        return out.reshape((m.nx, m.ny, m.nz), order="F")

    def set_wavefield(self, f, arr):
        from fwant_amd._lib import FIELDS
        fi = FIELDS.index(f) if isinstance(f, str) else int(f)
        a = np.asfortranarray(arr, self.lib.dtype).ravel(order="F")
        self.lib.set_field(self.h, fi, self.lib.ptr(a))

    def rhs(self, xf, yf, zf):
        m = self.model
        n = m.nx * m.ny * m.nz
        out = np.zeros(9 * n, self.lib.dtype)
        self.lib.rhs(self.h, xf, yf, zf, self.lib.ptr(out))
        return out.reshape((9, m.nz, m.ny, m.nx)).transpose(0, 3, 2, 1)

    def src_terms(self, ntime, tinc):
        """Source terms of one stage at time (ntime + tinc) dt: h [9][nx][ny][nz]
        and the surface planes TxSrc TySrc TzSrc VxSrc VySrc VzSrc [6][nx][ny]."""
        m = self.model
        n = m.nx * m.ny * m.nz
        h = np.zeros(9 * n, self.lib.dtype)
        surf = np.zeros(6 * m.nx * m.ny, self.lib.dtype)
        self.lib.src_terms(self.h, int(ntime), float(tinc), self.lib.ptr(h), self.lib.ptr(surf))
        return (h.reshape((9, m.nz, m.ny, m.nx)).transpose(0, 3, 2, 1),
                surf.reshape((6, m.ny, m.nx)).transpose(0, 2, 1))

    def pml_profiles(self):
        m = self.model
        n = m.nx + m.ny + m.nz
        A, B, D = (np.zeros(n, self.lib.dtype) for _ in range(3))
        self.lib.get_pml_profiles(self.h, self.lib.ptr(A), self.lib.ptr(B), self.lib.ptr(D))
        return A, B, D

    def close(self):
        if self.h:
            self.lib.destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
