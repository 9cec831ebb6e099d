"""After the kernels: assembly onto inversion blocks and the damped,
smoothed least-squares inversion (code/srcF.kernel/SI_ker2assm.F90,
inv.LSQR/solver_damp_smooth.f90, inv.LSQR/mod_LSQR.f90), on the GPU.

  assemble(K, blocks)              B(n) = sum_pixels K * volume   (:300-320)
  compact(B, scale, ker0)          the (index, value) pairs export_G_compact
                                   keeps: |scale*B(n)| > ker0      (:545-572)
  write_G_compact / read_G_compact the Fortran unformatted record file
  System                           the rows lldrow adds: data rows (kernel
                                   coefficients above ker_thres, station and
                                   event terms, data weight), damping rows
                                   lamb * I, smoothing rows eta * S
                                   (solver_damp_smooth.f90:135-262)
  lsqr(system, ...)                slsqr + lgtsol on the GPU (fwl_*)
"""
import ctypes as C
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DeviceArray, kcheck


@dataclass
class Blocks:
    """Inversion blocks in CSR form: block n owns pixels ptr[n]..ptr[n+1]-1
    (0-based point indices into the kernel volume) with their volumes
    (indx_of_pix / num_of_pix / volume of the block file)."""
    ptr: np.ndarray
    pix: np.ndarray
    vol: np.ndarray

    @property
    def n(self):
        return len(self.ptr) - 1


def assemble(K, blocks):
    """Block sums of a kernel (DeviceArray or host float32 array of the
    snapshot points) -> host float64 array [nblk]."""
    Kd = K if isinstance(K, DeviceArray) else DeviceArray.from_host(np.ascontiguousarray(K, np.float32))
    ptr = DeviceArray.from_host(np.ascontiguousarray(blocks.ptr, np.int64))
    pix = DeviceArray.from_host(np.ascontiguousarray(blocks.pix, np.int64))
    vol = DeviceArray.from_host(np.ascontiguousarray(blocks.vol, np.float32))
    B = DeviceArray((blocks.n,), np.float64)
    kcheck(_lib.lib().fwk_assemble(Kd.numel(), C.c_void_p(Kd.ptr), blocks.n, C.c_void_p(ptr.ptr), C.c_void_p(pix.ptr),
                                   C.c_void_p(vol.ptr), C.c_void_p(B.ptr), None), "fwk_assemble")
    return B.to_host()


def compact(B, scale, ker0):
    """(1-based block indices, float32 values) with |scale*B| > ker0, as export_G_compact keeps them."""
    v = (np.asarray(B, np.float32) * np.float32(scale)).astype(np.float32)
    keep = np.nonzero(np.abs(v) > np.float32(ker0))[0]
    return keep.astype(np.int32) + 1, v[keep]


def _record(f, payload):
    f.write(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))


def write_G_compact(path, num_blk, ker0, idx, val):
    """Fortran unformatted sequential file: (num_blk, ker0), (ncoef),
    (idx(m), val(m), m = 1..ncoef)."""
    idx = np.asarray(idx, np.int32)
    val = np.asarray(val, np.float32)
    pairs = np.empty(len(idx), dtype=[("i", "<i4"), ("v", "<f4")])
    pairs["i"], pairs["v"] = idx, val
    with open(path, "wb") as f:
        _record(f, struct.pack("<if", num_blk, ker0))
        _record(f, struct.pack("<i", len(idx)))
        _record(f, pairs.tobytes())


def read_G_compact(path):
    with open(path, "rb") as f:
        data = f.read()
    recs, o = [], 0
    while o < len(data):
        (n,) = struct.unpack_from("<i", data, o)
        recs.append(data[o + 4:o + 4 + n])
        o += n + 8
    num_blk, ker0 = struct.unpack("<if", recs[0])
    (ncoef,) = struct.unpack("<i", recs[1])
    pairs = np.frombuffer(recs[2], dtype=[("i", "<i4"), ("v", "<f4")], count=ncoef)
    return num_blk, ker0, pairs["i"].copy(), pairs["v"].copy()


@dataclass
class System:
    """Rows of the damped, smoothed least-squares system in CSR (0-based columns)."""
    ncol: int
    ptr: list = field(default_factory=lambda: [0])
    col: list = field(default_factory=list)
    val: list = field(default_factory=list)
    rhs: list = field(default_factory=list)

    def add_row(self, cols, vals, rhs):
        """lldrow: a row with at least one coefficient (empty rows are skipped)."""
        if len(cols) == 0:
            return
        self.col.extend(int(c) for c in cols)
        self.val.extend(float(v) for v in vals)
        self.ptr.append(len(self.col))
        self.rhs.append(float(rhs))

    def add_data(self, kernels, rhs, weight=1.0, ker_thres=0.0, stations=(), events=(), num_mval=None,
                 num_sta=0, weig_sta=1.0, weig_evt=1.0):
        """One measurement row: kernels is a list of (1-based idx, value) per
        component, offset by num_blk per component; coefficients with
        |value| <= ker_thres dropped; station / event terms (signed 1-based
        ids) after the model values; row and rhs scaled by the data weight."""
        cols, vals = [], []
        off = 0
        for idx, v, nblk in kernels:
            for i, x in zip(idx, v):
                if abs(float(x)) > ker_thres:
                    cols.append(off + int(i) - 1)
                    vals.append(float(x))
            off += nblk
        nm = off if num_mval is None else num_mval
        for s in stations:
            cols.append(nm + abs(int(s)) - 1)
            vals.append(float(np.copysign(weig_sta, s)))
        for e in events:
            cols.append(nm + num_sta + abs(int(e)) - 1)
            vals.append(float(np.copysign(weig_evt, e)))
        self.add_row(cols, [x * weight for x in vals], rhs * weight)

    def add_damping(self, lamb):
        for i in range(self.ncol):
            self.add_row([i], [lamb], 0.0)

    def add_smoothing(self, rows, eta, num_cmp=1, num_blk=None):
        """rows: (1-based block indices, coefficients) of the smoothing
        operator, applied to every component."""
        for idx, coef in rows:
            for k in range(num_cmp):
                self.add_row([(k * (num_blk or 0)) + int(i) - 1 for i in idx], [float(c) * eta for c in coef], 0.0)

    def arrays(self):
        return (np.asarray(self.ptr, np.int64), np.asarray(self.col, np.int64), np.asarray(self.val, np.float64),
                np.asarray(self.rhs, np.float64))


def lsqr(system, atol=1.0e-7, btol=1.0e-7, maxit=30000):
    """slsqr + lgtsol on the GPU: (x, standard errors, iterations, residual norm)."""
    L = _lib.lib()
    ptr, col, val, rhs = system.arrays()
    h = C.c_void_p()
    rc = L.fwl_create(len(rhs), system.ncol, len(val), ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                      rhs.ctypes.data, C.byref(h))
    if rc:
        raise _lib.FwError(f"fwl_create failed ({rc}): {L.fwl_last_error(None).decode()}")
    try:
        x = np.zeros(system.ncol)
        err = np.zeros(system.ncol)
        it, res = C.c_long(), C.c_double()
        rc = L.fwl_solve(h, atol, btol, maxit, x.ctypes.data, err.ctypes.data, C.byref(it), C.byref(res))
        if rc:
            raise _lib.FwError(f"fwl_solve failed ({rc}): {L.fwl_last_error(None).decode()}")
        return x, err, it.value, res.value
    finally:
        L.fwl_destroy(h)
