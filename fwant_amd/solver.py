"""Host-side driver over libfwant, mirroring the reference's solver driver.

``Solver`` follows the control flow of seis3d_wave (code/srcF/
seis3d_wave_8op_p3.F90:34-594): the set-up calls of :40-76 map to the
constructor (grid -> metric, media, sources, receivers, abs_init), the time
loop of :92-594 to ``run`` (one ``fw_step`` per time step; the step index
selects the F/B block of the 8-step cycle), and io_seismo_put /
io_wave_export to the receiver and snapshot buffers read back with
``receivers()`` / ``snapshot()``.

Errors are raised as ``FwError`` (the reference stops or aborts).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FwConfig, check, fptr, iptr, dptr


class Solver:
    def __init__(self, model, metric=None, max_records=None, snapshots=(), batch=None, metric3d_out=False):
        """model: fwant_amd.models.Model.  metric: None -> build on the GPU in
        FP64 (fw_build_metric); or a dict with xsin, xcot, xi_x, eta_y, zeta_z
        (float32) as read from metric_mpi*.nc.  snapshots: iterable of dicts
        with subs, subc, subt (1-based), tinv, fields (names), max_records and
        optionally format ("fp32", "bf16" or "fp16": 16-bit codes with a
        per-16^3-block FP32 scale).  batch: None, or a Batch (member of each
        force / moment source of the model and the steps each member runs):
        the members advance together in the same launches.  A model with r3
        (a terrain-following grid) runs mode C (fw_build_metric3d);
        metric3d_out keeps its FP32 metric (p, q, zz, rr) in self.metric3d."""
        L = _lib.lib()
        self.L = L
        self.model = model
        cfg = FwConfig()
        cfg.ni, cfg.nj, cfg.nk = model.ni, model.nj, model.nk
        cfg.stept = model.stept
        cfg.freenode = 1 if model.freenode else 0
        for n in range(6):
            cfg.abs_number[n] = int(model.abs_number[n])
            cfg.abs_velocity[n] = float(model.abs_velocity[n])
            cfg.cfs_bmax[n] = float(model.cfs_bmax[n])
            cfg.cfs_amax[n] = float(model.cfs_amax[n])
        h = C.c_void_p()
        check(L.fw_create(C.byref(cfg), C.byref(h)), None, "fw_create")
        self.h = h
        nx, ny, nz, fr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.fw_dims(h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(fr))
        self.nx, self.ny, self.nz, self.freenode = nx.value, ny.value, nz.value, bool(fr.value)
        x = np.ascontiguousarray(model.x, np.float64)
        y = np.ascontiguousarray(model.y, np.float64)
        z = np.ascontiguousarray(model.z, np.float64)
        if metric is None:
            m = {k: np.zeros(n, np.float32) for k, n in
                 (("xsin", self.nx), ("xcot", self.nx), ("xi_x", self.nx), ("eta_y", self.ny), ("zeta_z", self.nz))}
            check(L.fw_build_metric(h, dptr(x), dptr(y), dptr(z), fptr(m["xsin"]), fptr(m["xcot"]),
                                    fptr(m["xi_x"]), fptr(m["eta_y"]), fptr(m["zeta_z"])), h, "fw_build_metric")
            self.metric = m
        else:
            m = {k: np.ascontiguousarray(metric[k], np.float32) for k in ("xsin", "xcot", "xi_x", "eta_y", "zeta_z")}
            xf, yf, zf = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
            check(L.fw_set_grid(h, fptr(xf), fptr(yf), fptr(zf), fptr(m["xsin"]), fptr(m["xcot"]),
                                fptr(m["xi_x"]), fptr(m["eta_y"]), fptr(m["zeta_z"])), h, "fw_set_grid")
            self.metric = m
        self.metric3d = None
        if getattr(model, "r3", None) is not None:
            # mode C: the terrain-following grid's 3-D metric, built on the GPU in FP64
            r3 = np.asfortranarray(model.r3, np.float64).ravel(order="F")
            shp = (self.nx, self.ny, self.nz)
            if metric3d_out:
                m3 = {k: np.zeros(self.nx * self.ny * self.nz, np.float32) for k in ("p", "q", "zz", "rr")}
                check(L.fw_build_metric3d(h, dptr(x), dptr(y), dptr(r3), fptr(m3["p"]), fptr(m3["q"]),
                                          fptr(m3["zz"]), fptr(m3["rr"])), h, "fw_build_metric3d")
                self.metric3d = {k: v.reshape(shp, order="F") for k, v in m3.items()}
            else:
                nul = C.POINTER(C.c_float)()
                check(L.fw_build_metric3d(h, dptr(x), dptr(y), dptr(r3), nul, nul, nul, nul), h, "fw_build_metric3d")
        lam = np.asfortranarray(model.lam, np.float32).ravel(order="F")
        mu = np.asfortranarray(model.mu, np.float32).ravel(order="F")
        rho = np.asfortranarray(model.rho, np.float32).ravel(order="F")
        check(L.fw_set_media(h, fptr(lam), fptr(mu), fptr(rho)), h, "fw_set_media")
        f = model.force
        if f.n > 0:
            idx = np.ascontiguousarray(f.indx, np.int32).ravel()
            sh = np.ascontiguousarray(f.shift, np.float32).ravel()
            cs = [np.ascontiguousarray(f.comps[q], np.float32) for q in range(3)]
            check(L.fw_set_force(h, f.n, f.ntw, f.stf_id, iptr(idx), fptr(sh), fptr(np.ascontiguousarray(f.t0, np.float32)),
                                 fptr(np.ascontiguousarray(f.stf_time, np.float32)),
                                 fptr(np.ascontiguousarray(f.stf_freq, np.float32)),
                                 fptr(cs[0]), fptr(cs[1]), fptr(cs[2])), h, "fw_set_force")
        mo = model.moment
        if mo.n > 0:
            idx = np.ascontiguousarray(mo.indx, np.int32).ravel()
            sh = np.ascontiguousarray(mo.shift, np.float32).ravel()
            M6 = np.ascontiguousarray(mo.comps, np.float32).ravel()
            check(L.fw_set_moment(h, mo.n, mo.ntw, mo.stf_id, iptr(idx), fptr(sh),
                                  fptr(np.ascontiguousarray(mo.t0, np.float32)),
                                  fptr(np.ascontiguousarray(mo.stf_time, np.float32)),
                                  fptr(np.ascontiguousarray(mo.stf_freq, np.float32)), fptr(M6)), h, "fw_set_moment")
        npt = len(model.rec_indx)
        self.npt = npt
        if npt > 0:
            mr = max_records if max_records is not None else model.ref_steps() // model.rec_tinv + 8
            ri = np.ascontiguousarray(model.rec_indx, np.int32).ravel()
            check(L.fw_set_receivers(h, npt, model.rec_tinv, iptr(ri), int(mr)), h, "fw_set_receivers")
        self.snaps = []
        for s in snapshots:
            mask = 0
            for name in s["fields"]:
                mask |= 1 << _lib.FIELDS.index(name)
            sid = C.c_int()
            subs = np.array(s["subs"], np.int32)
            subc = np.array(s["subc"], np.int32)
            subt = np.array(s["subt"], np.int32)
            fmt = s.get("format", "fp32")
            if fmt == "fp32":
                check(L.fw_add_snapshot(h, iptr(subs), iptr(subc), iptr(subt), int(s["tinv"]), mask,
                                        int(s["max_records"]), C.byref(sid)), h, "fw_add_snapshot")
            else:
                check(L.fw_add_snapshot_packed(h, iptr(subs), iptr(subc), iptr(subt), int(s["tinv"]), mask,
                                               int(s["max_records"]), _lib.SGT_FORMATS[fmt], C.byref(sid)), h,
                      "fw_add_snapshot_packed")
            nblk = int(np.prod([(c + 15) // 16 for c in s["subc"]]))
            self.snaps.append(dict(s, id=sid.value, nf=len(s["fields"]), npts=int(np.prod(subc)), nblk=nblk,
                                   format=fmt))
        self.nbatch = 1
        if batch is not None:
            fm = np.ascontiguousarray(batch.force_member, np.int32)
            mm = np.ascontiguousarray(batch.moment_member, np.int32)
            ns = np.ascontiguousarray(batch.nsteps, np.int32)
            check(L.fw_set_batch(h, len(ns), iptr(fm), iptr(mm), iptr(ns)), h, "fw_set_batch")
            self.nbatch = len(ns)
        check(L.fw_init(h), h, "fw_init")
        self.ntime = 0

    # -- time stepping ------------------------------------------------------
    def run(self, nsteps, ntime0=None):
        t0 = self.ntime if ntime0 is None else ntime0
        check(self.L.fw_run(self.h, int(t0), int(nsteps)), self.h, "fw_run")
        self.ntime = t0 + nsteps

    def select(self, member):
        """Address member `member` of a batch in the output / wavefield calls."""
        check(self.L.fw_select_member(self.h, int(member)), self.h, "fw_select_member")

    def set_cycle_origin(self, origin):
        """Step ntime uses F/B block (ntime - origin) mod 8 (a reference-style
        restart at ntime0 passes origin = ntime0)."""
        check(self.L.fw_set_cycle_origin(self.h, int(origin)), self.h, "fw_set_cycle_origin")

    def checkpoint(self):
        """Complete state between steps (wavefield + PML auxiliaries) as a float32 array."""
        n = C.c_size_t()
        check(self.L.fw_checkpoint_size(self.h, C.byref(n)), self.h, "fw_checkpoint_size")
        buf = np.zeros(n.value // 4, np.float32)
        check(self.L.fw_checkpoint_save(self.h, buf.ctypes.data_as(C.c_void_p)), self.h, "fw_checkpoint_save")
        return buf

    def restore(self, buf):
        buf = np.ascontiguousarray(buf, np.float32)
        n = C.c_size_t()
        check(self.L.fw_checkpoint_size(self.h, C.byref(n)), self.h, "fw_checkpoint_size")
        if buf.nbytes != n.value:
            raise ValueError(f"checkpoint holds {buf.nbytes} bytes, this solver needs {n.value}")
        check(self.L.fw_checkpoint_load(self.h, buf.ctypes.data_as(C.c_void_p)), self.h, "fw_checkpoint_load")

    def use_graphs(self, on=True):
        check(self.L.fw_use_graphs(self.h, 1 if on else 0), self.h, "fw_use_graphs")

    def set_width(self, width):
        check(self.L.fw_set_batch_width(self.h, int(width)), self.h, "fw_set_batch_width")

    def autotune(self, trial_steps=4):
        """Pick the fastest tile shape and batch width; returns (tile, width, ms per step).
        fw_autotune restores the zero initial state and empty records, so the
        step counter starts again at 0: autotune before run(), or expect the
        next run() to start from rest at ntime 0."""
        t, w, ms = C.c_int(), C.c_int(), C.c_float()
        check(self.L.fw_autotune(self.h, int(trial_steps), C.byref(t), C.byref(w), C.byref(ms)), self.h,
              "fw_autotune")
        self.ntime = 0
        return t.value, w.value, ms.value

    def sync(self):
        check(self.L.fw_sync(self.h), self.h, "fw_sync")

    # -- outputs --------------------------------------------------------------
    def receivers(self):
        n = C.c_int()
        check(self.L.fw_get_receivers(self.h, None, C.byref(n)), self.h, "fw_get_receivers")
        out = np.zeros((n.value, self.npt, 9), np.float32)
        if n.value:
            check(self.L.fw_get_receivers(self.h, fptr(out), C.byref(n)), self.h, "fw_get_receivers")
        return out

    def snapshot(self, k=0):
        s = self.snaps[k]
        n = C.c_int()
        check(self.L.fw_get_snapshot(self.h, s["id"], None, C.byref(n)), self.h, "fw_get_snapshot")
        out = np.zeros((n.value, s["nf"], s["npts"]), np.float32)
        if n.value:
            check(self.L.fw_get_snapshot(self.h, s["id"], fptr(out), C.byref(n)), self.h, "fw_get_snapshot")
        return out

    def snapshot_device(self, k=0):
        s = self.snaps[k]
        p = C.c_void_p()
        n = C.c_int()
        npts = C.c_long()
        check(self.L.fw_snapshot_device(self.h, s["id"], C.byref(p), C.byref(n), C.byref(npts)), self.h,
              "fw_snapshot_device")
        return p.value, n.value, npts.value

    def snapshot_packed(self, k=0):
        """codes [nrec][nf][npts] (uint16) and scales [nrec][nf][nblocks] of a packed snapshot."""
        s = self.snaps[k]
        n = C.c_int()
        nb = C.c_long()
        check(self.L.fw_get_snapshot_packed(self.h, s["id"], None, None, C.byref(n), C.byref(nb)), self.h,
              "fw_get_snapshot_packed")
        codes = np.zeros((n.value, s["nf"], s["npts"]), np.uint16)
        scales = np.zeros((n.value, s["nf"], nb.value), np.float32)
        if n.value:
            check(self.L.fw_get_snapshot_packed(self.h, s["id"], codes.ctypes.data_as(C.c_void_p), fptr(scales),
                                                C.byref(n), C.byref(nb)), self.h, "fw_get_snapshot_packed")
        return codes, scales

    def snapshot_packed_device(self, k=0):
        """(codes pointer, scales pointer, records, points, blocks) of a packed snapshot in HBM."""
        s = self.snaps[k]
        pc, ps = C.c_void_p(), C.c_void_p()
        n, npts, nb = C.c_int(), C.c_long(), C.c_long()
        check(self.L.fw_snapshot_packed_device(self.h, s["id"], C.byref(pc), C.byref(ps), C.byref(n), C.byref(npts),
                                               C.byref(nb)), self.h, "fw_snapshot_packed_device")
        return pc.value, ps.value, n.value, npts.value, nb.value

    def wavefield(self, f):
        fi = _lib.FIELDS.index(f) if isinstance(f, str) else int(f)
        out = np.zeros(self.nx * self.ny * self.nz, np.float32)
        check(self.L.fw_get_wavefield(self.h, fi, fptr(out)), self.h, "fw_get_wavefield")
        return out.reshape((self.nx, self.ny, self.nz), order="F")

    def set_wavefield(self, f, arr):
        fi = _lib.FIELDS.index(f) if isinstance(f, str) else int(f)
        a = np.asfortranarray(arr, np.float32).ravel(order="F")
        check(self.L.fw_set_wavefield(self.h, fi, fptr(a)), self.h, "fw_set_wavefield")

    def pml_profiles(self):
        n = self.nx + self.ny + self.nz
        A, B, D = (np.zeros(n, np.float32) for _ in range(3))
        check(self.L.fw_get_pml_profiles(self.h, fptr(A), fptr(B), fptr(D)), self.h, "fw_get_pml_profiles")
        return A, B, D

    def timing(self, on=True):
        self.L.fw_enable_timing(self.h, 1 if on else 0)

    def get_timing(self):
        s, t, n = C.c_double(), C.c_double(), C.c_int()
        self.L.fw_get_timing(self.h, C.byref(s), C.byref(t), C.byref(n))
        return s.value, t.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.L.fw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def build_metric(model):
    """The FP32 metric of `model`'s grid built on the GPU in FP64 (as
    seis3d_metric writes it to metric_mpi*.nc)."""
    m = models_lite_copy(model)
    s = Solver(m)
    met = s.metric
    s.close()
    return met


def models_lite_copy(model):
    """A copy of `model` without sources and receivers (for set-up only work)."""
    import copy
    from . import models
    m = copy.copy(model)
    m.force = models.Sources()
    m.moment = models.Sources(comps=np.zeros((6, 0), np.float32))
    m.rec_indx = np.zeros((0, 3), np.int32)
    return m

