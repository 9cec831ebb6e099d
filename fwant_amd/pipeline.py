"""Station-pair sensitivity kernels end to end: virtual-source solves that
store the 6-component stress SGT, then the kernel time integral per pair
(the work of code/srcF/seis3d_wave_8op_p3.F90 per station followed by
code/srcF.kernel/SI_ker_sta.F90 per pair).

Two ways to produce the same kernels:
  plain    one solve per station, launch by launch, FP32 SGT; with
           host_sgt the SGT of each station leaves the GPU after its solve
           and each pair's two SGTs come back for the integral (the
           reference's SGT-file round trip, through host memory: at the
           configured record lengths the FP32 SGT of all stations does not
           fit in HBM beside the solver)
  batched  all stations in one batched solve (members share the model;
           step graphs; tile shape and batch width from the autotuner),
           SGT stored as BF16 or FP16 codes with per-block scales, kept in
           HBM and restored on the device before the integral

Both return the kernels (Kap, Kaq, Kbp, Kbq per pair, on the host) and a
dict of wall times: setup (solver construction: metric, media, sources,
PML; for the batched path also the autotuner), steps (the time stepping
with SGT capture) and kernel (SGT restore and the pair integrals).
"""
import time

import numpy as np

from . import kernel, models
from ._lib import DeviceArray, device_sync
from .solver import Solver

FIELDS6 = ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")


def sgt_snapshot(model, tinv, step, fmt="fp32", max_records=None):
    """Whole interior on a grid decimated by `step`, every `tinv` steps."""
    subc = tuple(max(1, n // step) for n in (model.ni, model.nj, model.nk))
    return dict(subs=(4, 4, 4), subc=subc, subt=(step, step, step), tinv=tinv, fields=FIELDS6,
                max_records=max_records or (model.nt // tinv + 2), format=fmt)


def _media_on(model, snap):
    sl = tuple(slice(s - 1, s - 1 + c * t, t) for s, c, t in zip(snap["subs"], snap["subc"], snap["subt"]))
    lam = np.asarray(model.lam[sl], np.float32).ravel(order="F")
    mu = np.asarray(model.mu[sl], np.float32).ravel(order="F")
    return DeviceArray.from_host(lam), DeviceArray.from_host(mu)


def _pair(RA, SB, dl, dm, dt, V, U, win, M0, filt):
    out = kernel.pair_kernel(None, None, RA, SB, dl, dm, dt, V, U, win[0], win[1], M0, 2, filt=filt)
    return [o.to_host() for o in out]


def plain(model, stations, nsteps, pairs, tinv=10, step=2, V=None, U=None, win=None, M0=1e16, filt=None,
          host_sgt=False):
    """One solve per station (FP32 SGT), then the pair kernels.  host_sgt:
    each station's SGT goes to host memory after its solve (timed as
    sgt_out) and each pair's two SGTs back to the GPU (part of kernel)."""
    sgts, solvers = {}, []
    t_setup = t_steps = t_out = 0.0
    used = {x for p in pairs for x in p}
    for q, st in enumerate(stations):
        t0 = time.perf_counter()
        m, _ = models.virtual_sources(model, [st], [nsteps[q]])
        m.rec_indx = np.zeros((0, 3), np.int32)
        snap = sgt_snapshot(m, tinv, step, "fp32", max_records=max(nsteps) // tinv + 2)
        s = Solver(m, snapshots=[snap], max_records=1)
        s.sync()
        t1 = time.perf_counter()
        s.run(nsteps[q])
        s.sync()
        t_setup += t1 - t0
        t2 = time.perf_counter()
        t_steps += t2 - t1
        if host_sgt:
            if q in used:
                sgts[q] = s.snapshot(0)  # (nrec, 6, npts) on the host
            s.close()
            t_out += time.perf_counter() - t2
        else:
            ptr, nrec, npts = s.snapshot_device(0)
            sgts[q] = (ptr, nrec, npts)
            solvers.append(s)
    t0 = time.perf_counter()
    dl, dm = _media_on(model, snap)
    ks = []
    for a, b in pairs:
        if host_sgt:
            knt = min(sgts[a].shape[0], sgts[b].shape[0])
            da, db = DeviceArray.from_host(sgts[a][:knt]), DeviceArray.from_host(sgts[b][:knt])
            npts = sgts[a].shape[2]
            RA = kernel.snapshot_series(da.ptr, knt, 6, npts)
            SB = kernel.snapshot_series(db.ptr, knt, 6, npts)
        else:
            knt = min(sgts[a][1], sgts[b][1])
            RA = kernel.snapshot_series(sgts[a][0], knt, 6, sgts[a][2])
            SB = kernel.snapshot_series(sgts[b][0], knt, 6, sgts[b][2])
        ks.append(_pair(RA, SB, dl, dm, model.stept * tinv, V[:knt], U[:knt], win, M0, filt))
    device_sync()
    t_kernel = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return ks, dict(setup=t_setup, steps=t_steps, sgt_out=t_out, kernel=t_kernel)


def batched(model, stations, nsteps, pairs, fmt="bf16", tinv=10, step=2, V=None, U=None, win=None, M0=1e16,
            filt=None, tune_steps=4):
    """All stations in one batched solve (graphs, autotuned), packed SGT."""
    t0 = time.perf_counter()
    m, b = models.virtual_sources(model, stations, nsteps)
    m.rec_indx = np.zeros((0, 3), np.int32)
    snap = sgt_snapshot(m, tinv, step, fmt, max_records=max(nsteps) // tinv + 2)
    s = Solver(m, snapshots=[snap], max_records=1, batch=b)
    tune = s.autotune(tune_steps) if tune_steps else (0, 0, 0.0)
    s.use_graphs(True)
    s.sync()
    t1 = time.perf_counter()
    s.run(max(nsteps))
    s.sync()
    t_setup, t_steps = t1 - t0, time.perf_counter() - t1
    t0 = time.perf_counter()
    dl, dm = _media_on(model, snap)
    npts = int(np.prod(snap["subc"]))

    def decoded(q):
        # a station's FP32 SGT, decoded for one pair and freed after it: at
        # the configured record lengths the FP32 SGT of every station at once
        # would not fit in HBM beside the 16-bit store
        s.select(q)
        pc, ps, nrec, _, _ = s.snapshot_packed_device(0)
        y = DeviceArray((nrec, 6, npts), np.float32)
        kernel.sgt_decode(fmt, pc, ps, snap["subc"], nrec * 6, y)
        return y, nrec

    ks = []
    for a, bb in pairs:
        (ya, na), (yb, nb) = decoded(a), decoded(bb)
        knt = min(na, nb)
        RA = kernel.snapshot_series(ya.ptr, knt, 6, npts)
        SB = kernel.snapshot_series(yb.ptr, knt, 6, npts)
        ks.append(_pair(RA, SB, dl, dm, model.stept * tinv, V[:knt], U[:knt], win, M0, filt))
        device_sync()
        ya.free()
        yb.free()
    device_sync()
    t_kernel = time.perf_counter() - t0
    s.close()
    return ks, dict(setup=t_setup, steps=t_steps, kernel=t_kernel, tile=tune[0], width=tune[1])
