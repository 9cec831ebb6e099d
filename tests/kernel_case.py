"""A TomoKernel.conf kernel-stage case and its oracle pipeline, shared by
the runner's GPU test (SGTs from the GPU solver) and its CPU test (random
SGT records, oracle kernels in place of the device layer).

The case: a 12 x 10 x 6 stress snapshot with stride 2, written as the
reference's time-chunked sgt_001_mpi000000_nNNNNN.nc files (10 records per
file, 36 of the 40 records used) for one station (x, y, z) and two events;
a strided kernel sub-grid of 12 x 8 x 3 points over 6 blocks; picks on all
six components (P1, P2 and N; rotated components; a pick outside the
record; two picks sharing a band and filter; comment lines); an event with
no picks.  The oracle recomputes every pick on the whole sub-grid at once:
stress2strain, scatter1, scatter_rotate for the rotated components (with
the synthetics rotated in FP32), series_filter for P1/P2 (series and
synthetics) and scatter2kernel, with the pick windows worked out by hand.
"""
import ctypes as C
import os

import numpy as np
from scipy import signal
from scipy.io import netcdf_file

STRESS = ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")
SNAP = dict(subs=(8, 8, 10), subc=(12, 10, 6), subt=(2, 2, 2), tinv=2, fields=STRESS, max_records=40)
KNT, TCNT, M0 = 36, 10, 1.0e16
ROT = np.array([[0.8, -0.6, 0.0], [0.6, 0.8, 0.0], [0.0, 0.0, 1.0]], np.float32)  # rot[r, c]
BANDS = {1: signal.butter(2, 0.3), 2: signal.butter(3, 0.15)}
# event, picknum(6), picks per component: (band, filter, t1, t2, outdir, expected window or None)
EVENTS = [
    ("EVT1", (1, 0, 3, 1, 1, 0), {
        0: [(1, "P2", 0.85, 2.95, "k/e1x", (8, 30))],
        2: [(2, "P1", 0.55, 3.05, "k/e1z1", (5, 31)), (2, "P1", 1.25, 2.15, "k/e1z2", (12, 22)),
            (2, "P1", 3.75, 3.9, "k/e1zskip", None)],
        3: [(0, "N", 0.35, 2.45, "k/e1r", (3, 25))],
        4: [(1, "P2", -1.0, 1.0e9, "k/e1t", (1, 36))],
    }),
    ("EVT2", (0, 1, 0, 0, 0, 0), {1: [(1, "N", 1.05, 3.35, "k/e2y", (10, 34))]}),
    ("EVT3", (0, 0, 0, 0, 0, 0), {}),
]


def model():
    from fwant_amd import models
    return models.small_test(nt=80, seed=31, moment=False)


def solver_records(m):
    """Station SGTs (forces x, y, z at a surface point) and event SGTs
    (moment sources) from the GPU solver: {name: (40, 6, npts)}."""
    from fwant_amd import models
    from fwant_amd.solver import Solver

    def solve(force=None, moment=None):
        mm = models.small_test(nt=80, seed=31, moment=False)
        mm.force = force if force is not None else models.Sources(n=0, ntw=1)
        mm.moment = moment if moment is not None else models.Sources(n=0, ntw=1)
        s = Solver(mm, snapshots=[SNAP])
        s.run(80)
        r = s.snapshot(0)
        s.close()
        return r

    def point(indx, comps, f0):
        s = models.Sources(n=1, ntw=1, stf_id=models.STF_RICKER)
        s.indx = np.array([indx], np.int32)
        s.shift = np.zeros((1, 3), np.float32)
        s.t0 = np.zeros(1, np.float32)
        s.stf_time = np.array([1.2], np.float32)
        s.stf_freq = np.array([f0], np.float32)
        s.comps = np.asarray(comps, np.float32).reshape(-1, 1)
        return s

    rec = {}
    sta = (3 + m.ni // 2 + 1, 3 + m.nj // 2 + 1, m.nk2 + 1)
    for c, d in enumerate("xyz"):
        print(f"station SGT {d}", flush=True)
        f = np.zeros(3)
        f[c] = 1.0e15
        rec[d] = solve(force=point(sta, f, 0.8))
    for e, (name, mom) in enumerate((("EVT1", (1, -2, 0.5, 0.7, -0.3, 1.1)), ("EVT2", (-0.4, 0.9, 1.3, 0.2, 0.8, -1.0)))):
        print(f"event SGT {name}", flush=True)
        loc = (3 + m.ni // 3 + 1 + 6 * e, 3 + 2 * m.nj // 3 + 1 - 4 * e, 3 + m.nk // 2 + 1)
        rec[name] = solve(moment=point(loc, np.array(mom) * 1e15, 0.6))
    return rec


def random_records(seed=0):
    """Random stress records of the same shapes (for the host flow)."""
    rng = np.random.default_rng(seed)
    npts = int(np.prod(SNAP["subc"]))
    return {n: (rng.standard_normal((40, 6, npts)) * 1e5).astype(np.float32) for n in ("x", "y", "z", "EVT1", "EVT2")}


def write_case(root, m, rec, metric=None):
    """The run directory: SeisFD3D.conf and inputs, the SGT files, the
    synthetics, the filter files, TomoKernel.conf and the lists."""
    from fwant_amd import io as fio
    from fwant_amd import kernel_run as kr
    fio.write_inputs(m, root, metric=metric, snapshots=[SNAP], snap_tcnt=[TCNT])
    for d in "xyz":
        sd = os.path.join(root, "sgt", "STA1", f"sta_{d}")
        os.makedirs(sd, exist_ok=True)
        fio.write_snapshots(sd, 1, SNAP, TCNT, rec[d], m.stept)
    for name in ("EVT1", "EVT2"):
        d = os.path.join(root, "event", name, "out")
        os.makedirs(d, exist_ok=True)
        fio.write_snapshots(d, 1, SNAP, TCNT, rec[name], m.stept)
    kdt = np.float32(SNAP["tinv"]) * np.float32(m.stept)
    T = np.arange(1, KNT + 1, dtype=np.float32) * kdt
    syn = {}
    for e, (name, *_rest) in enumerate(EVENTS):
        for c, d in enumerate("xyz"):
            V = np.sin(2 * np.pi * (0.3 + 0.05 * c + 0.02 * e) * T).astype(np.float32)
            U = (np.cos(2 * np.pi * (0.25 + 0.04 * c) * T) * np.exp(-T / 3)).astype(np.float32)
            syn[name, c] = (V, U)
            np.savetxt(os.path.join(root, f"syn_{name}_{d}.txt"), np.stack([T, V, U], 1), fmt="%.9e")
    for q, (b, a) in BANDS.items():
        kr.write_filter(os.path.join(root, f"band{q}.txt"), b, a)
    with open(os.path.join(root, "TomoKernel.conf"), "w") as f:
        f.write("MAIN_CONF = SeisFD3D.conf\nsnap_id = 1\nfirst_run = T\n"
                "sub_start = 1 2 1\nsub_count = -1 8 -1\nsub_stride = 1 1 2\n"
                "block_from = 1 1 1\nblock_size = 5 4 -1\n"
                f"SGT_TIMEDIM_SIZE = {KNT}\nSGT_M0 = {M0!r}\n"
                "SGT_ROOT = sgt\nSGT_STAX = sta_x\nSGT_STAY = sta_y\nSGT_STAZ = sta_z\n"
                "EVENT_ROOT = event\nEVENT_OUTD = out\n"
                "maximum_of_pick_per_component = 2\nnumber_of_freqband = 3\n"
                "<filter_file>\nband1.txt\nband2.txt\nnone\n")
    with open(os.path.join(root, "evt.list"), "w") as f:
        for name, pk, picks in EVENTS:
            f.write(f"# event {name}\n{name} " + " ".join(map(str, pk)) + " "
                    + " ".join(repr(float(v)) for v in ROT.ravel(order="F")) + "\n")
            for d in "xyz":
                f.write(f"syn_{name}_{d}.txt\n")
            for c in range(6):
                for q, (band, filt, t1, t2, out, _w) in enumerate(picks.get(c, [])):
                    if q == 1:
                        f.write("# a comment between picks\n")
                    f.write(f"{band} {filt} {t1} {t2} {out}\n")
    with open(os.path.join(root, "sta.list"), "w") as f:
        f.write("# station list\nSTA1 evt.list\n")
    return syn, T


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def P(lib, a):
    return a.ctypes.data_as(C.POINTER(lib.R))


def oracle_kernels(m, rec, syn, T):
    """Every pick's kernel on the whole sub-grid through the FP32 oracle."""
    from oracle.pyoracle import Lib
    lib = Lib(32)
    D = C.POINTER(C.c_double)
    sc = SNAP["subc"]
    ii, jj, kk = np.arange(0, 12), np.arange(1, 9), np.arange(0, 6, 2)  # kernel points, 0-based snapshot indices

    def at(r):  # (nrec, 6, npts) -> (6, KNT, npt) at the kernel points, i fastest
        x = r[:KNT].reshape(KNT, 6, sc[2], sc[1], sc[0])[:, :, kk][:, :, :, jj][:, :, :, :, ii]
        return np.ascontiguousarray(x.reshape(KNT, 6, -1).transpose(1, 0, 2))

    sel = np.ix_(*[np.array([SNAP["subs"][d] - 1 + SNAP["subt"][d] * v for v in (ii, jj, kk)[d]]) for d in range(3)])
    lam = np.ascontiguousarray(m.lam[sel].ravel(order="F"), np.float32)
    mu = np.ascontiguousarray(m.mu[sel].ravel(order="F"), np.float32)
    npt = lam.size
    kdt = float(T[0])

    def strain(r):
        E = [np.ascontiguousarray(x) for x in at(r)]
        lib.stress2strain(npt, KNT, *[P(lib, e) for e in E], P(lib, lam), P(lib, mu))
        return E

    sta = [strain(rec[d]) for d in "xyz"]
    out = {}
    for name, pk, picks in EVENTS:
        if not picks:
            continue
        S = strain(rec[name])
        A, B = [], []
        for c in range(3):
            a, b = np.zeros((KNT, npt), np.float32), np.zeros((KNT, npt), np.float32)
            lib.scatter1(npt, KNT, kdt, (C.POINTER(lib.R) * 6)(*[P(lib, e) for e in sta[c]]),
                         (C.POINTER(lib.R) * 6)(*[P(lib, e) for e in S]), P(lib, lam), P(lib, mu), P(lib, a), P(lib, b))
            A.append(a)
            B.append(b)
        V = [syn[name, c][0].copy() for c in range(3)]
        U = [syn[name, c][1].copy() for c in range(3)]
        rcm = np.ascontiguousarray(ROT.ravel(order="F"))
        for c in range(6):
            if c == 3 and any(p > 0 for p in pk[3:6]):
                V = [V[0] * ROT[0, q] + V[1] * ROT[1, q] + V[2] * ROT[2, q] for q in range(3)]
                U = [U[0] * ROT[0, q] + U[1] * ROT[1, q] + U[2] * ROT[2, q] for q in range(3)]
                lib.scatter_rotate(npt, KNT, P(lib, A[0]), P(lib, A[1]), P(lib, A[2]), P(lib, rcm))
                lib.scatter_rotate(npt, KNT, P(lib, B[0]), P(lib, B[1]), P(lib, B[2]), P(lib, rcm))
            for band, filt, t1, t2, outdir, w in picks.get(c, []):
                if w is None:
                    continue
                q = c % 3
                Vv, Uu, Sa, Sb = V[q], U[q], A[q], B[q]
                if filt != "N":
                    bf, af = (np.ascontiguousarray(x, np.float64) for x in BANDS[band])
                    mode = 1 if filt == "P1" else 2

                    def flt(x, n):
                        y = np.zeros_like(x)
                        lib.series_filter(mode, n, KNT, len(bf), bf.ctypes.data_as(D), af.ctypes.data_as(D),
                                          P(lib, np.ascontiguousarray(x)), P(lib, y))
                        return y
                    Vv, Uu, Sa, Sb = flt(Vv, 1), flt(Uu, 1), flt(Sa, npt), flt(Sb, npt)
                ref = [np.zeros(npt, np.float32) for _ in range(4)]
                lib.scatter2kernel(npt, KNT, w[0], w[1], kdt, M0, P(lib, np.ascontiguousarray(Vv)),
                                   P(lib, np.ascontiguousarray(Uu)), P(lib, np.ascontiguousarray(Sa)),
                                   P(lib, np.ascontiguousarray(Sb)), *[P(lib, r) for r in ref])
                out[outdir] = ref
    return out


def read_kernel(path):
    from fwant_amd.kernel_run import KERNEL_VARS
    with netcdf_file(path, "r", mmap=False) as nc:
        return [np.array(nc.variables[v][:]) for v in KERNEL_VARS]




def n_picks():
    return sum(1 for _n, _p, picks in EVENTS for c in picks for p in picks[c] if p[-1] is not None)


def check_run(root, st, want):
    """The kernel files of a first run against the oracle; returns them."""
    from fwant_amd import kernel_run as kr
    assert st.skipped == 6 and st.kernels == 6 * n_picks() and len(want) == n_picks()  # per block
    assert not os.path.exists(os.path.join(root, "k", "e1zskip"))
    mask = np.zeros((6, 10, 12), bool)
    mask[0:6:2, 1:9, 0:12] = True  # (K, J, I) of the kernel points
    saved = {}
    for outdir, ref in want.items():
        path = os.path.join(root, outdir, "kernel_001_mpi000000_n00000.nc")
        got = read_kernel(path)
        saved[path] = got
        for name, g, r in zip(kr.KERNEL_VARS, got, ref):
            e = rel(g[mask], r)  # K, J, I order with I fastest
            print(f"{outdir} {name}: rel-L2 {e:.2e}", flush=True)
            assert np.abs(r).max() > 0 and e < 1e-5
            assert not g[~mask].any()
    return saved


def check_append_run(root, saved):
    """A later run with first_run off from block 2 1 1 writes the same values
    into the existing files."""
    from fwant_amd import kernel_run as kr
    conf = os.path.join(root, "TomoKernel.conf")
    with open(conf) as f:
        text = f.read().replace("first_run = T", "first_run = F").replace("block_from = 1 1 1", "block_from = 2 1 1")
    with open(conf, "w") as f:
        f.write(text)
    st2 = kr.run(conf, "sta.list")
    assert st2.kernels == 4 * n_picks()
    for path, got in saved.items():
        for g0, g1 in zip(got, read_kernel(path)):
            np.testing.assert_allclose(g1, g0, rtol=1e-6, atol=0)
