"""The kernel stage driven by TomoKernel.conf: host mirror of the main
program of code/srcF.kernel/SI_ker_sta.F90 (:135-821) over this library's
device kernels (fwant_amd.kernel).

    python -m fwant_amd.kernel_run [TomoKernel.conf] [--sgt-list LIST]

Inputs, as the reference program reads them:
  TomoKernel.conf (init_kernel :1047-1084, init_filter :1086-1112): the
    main SeisFD3D.conf (MAIN_CONF), the snapshot id (snap_id), first_run,
    the kernel sub-grid of the snapshot (sub_start/sub_count/sub_stride),
    its blocking (block_from/block_size), the number of SGT records
    (SGT_TIMEDIM_SIZE), SGT_M0, the station and event SGT directories
    (SGT_ROOT, SGT_STAX/Y/Z, EVENT_ROOT, EVENT_OUTD), number_of_freqband and,
    after the line <filter_file>, one filter file (or 'none') per band; a
    filter file holds NOD1, then NOD1 lines "a b" in the format (2(f20.16));
  SeisFD3D.conf: the snapshot's sub-grid, interval and records per file,
    stept and MEDIA_ROOT (lambda and mu at the kernel points);
  the station list (first line of stdin, as :150-151 reads it, or
    --sgt-list): lines "KSTANM evt_list";
  each event list: "KEVNM picknum(6) rot_mat(9, column-major)", three
    synthetic files (knt lines "T V U"), then, for each component with
    picks, picknum lines "band P1|P2|N t1 t2 outdir";
  the station SGTs SGT_ROOT/KSTANM/SGT_STA{X,Y,Z}/sgt_NNN_mpi000000_nMMMMM.nc
    and the event SGT EVENT_ROOT/KEVNM/EVENT_OUTD/sgt_..., records split
    over files of the snapshot's records-per-file.
Output: outdir/kernel_NNN_mpi000000_n00000.nc per pick (phase_Vp,
amplitude_Vp, phase_Vs, amplitude_Vs on the snapshot's grid; the file is
created at block (1, 1, 1) when first_run is set, otherwise it must exist),
written block by block at the kernel points.

The flow follows the reference: stations, then blocks (k, j, i; i fastest;
from block_from), then events.  Per block the station SGTs of the
components an event needs are read once and turned into strain
(stress2strain :1185-1211); per event the event SGT is read, turned into
strain and scattered against each needed station component
(sgt2scatter3 :1257-1336); components 4-6 use the synthetics and the
scattered series rotated by rot_mat (syn_rotate/scatter_rotate
:1339-1373); each pick filters (P1 causal / P2 zero-phase / N none;
recomputed when the band or the filter changes, :736-757) and integrates
the kernel over its window (scatter2kernel :1375-1431).  The pick window:
t1 > Tmax or t2 < Tmin skips the pick; t1 <= Tmin starts at sample 1,
else at the last sample with T <= t1; t2 >= Tmax ends at knt, else at the
first sample with T >= t2 (:686-710), with T the time column of the last
synthetic file read (init_synthetic overwrites T, :1114-1125) and
Tmin/Tmax = kdt, knt*kdt, kdt = snapshot interval * stept.

Where the reference's behaviour is undefined this runner stops or skips
instead: a skipped pick writes nothing (the reference's kernel loop reuses
the previous pick's window and file handle for it); a filter on a band
with no filter file, a band id above number_of_freqband, an unknown filter
name or a missing kernel file raise an error.  Paths are relative to the
directory of TomoKernel.conf (the reference runs in that directory).  One
rank: the reference's non-MPI build with mpi id 0 0 0.
"""
import argparse
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np
from scipy.io import netcdf_file

from . import io as fio
from . import kernel
from ._lib import DeviceArray

STRESS = ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")
KERNEL_VARS = ("phase_Vp", "amplitude_Vp", "phase_Vs", "amplitude_Vs")
FILTERS = {"P1": 1, "P2": 2, "N": 0}


def _fortran_f(field_text, d=16):
    """One Fortran Fw.d input field: blank is 0, no decimal point means d
    implied decimals."""
    s = field_text.strip()
    if not s:
        return 0.0
    if any(c in s for c in ".eEdD"):
        return float(s.replace("d", "e").replace("D", "e"))
    return int(s) * 10.0 ** (-d)


def read_filter(path):
    """(b, a) of a filter file: NOD1, then NOD1 lines of a and b in the
    format (2(f20.16)) (init_filter, SI_ker_sta.F90:1100-1108)."""
    with open(path) as f:
        lines = f.read().splitlines()
    nod1 = int(lines[0].split()[0])
    if nod1 < 1 or len(lines) < 1 + nod1:
        raise ValueError(f"{path}: NOD1 = {nod1} but {len(lines) - 1} coefficient lines")
    a = np.array([_fortran_f(ln[0:20]) for ln in lines[1:1 + nod1]])
    b = np.array([_fortran_f(ln[20:40]) for ln in lines[1:1 + nod1]])
    return b, a


def write_filter(path, b, a):
    """The file read_filter reads."""
    with open(path, "w") as f:
        f.write(f"{len(b)}\n")
        for av, bv in zip(a, b):
            f.write(f"{av:20.16f}{bv:20.16f}\n")


@dataclass
class KernelConf:
    root: str
    main_conf: str
    kid: int
    first_run: bool
    bstart: tuple
    bcount: tuple
    bstride: tuple
    blkfst: tuple
    blksiz: tuple
    knt: int
    M0: float
    sgt_root: str
    sgt_sta: tuple
    event_root: str
    event_outd: str
    num_pick: int
    bands: list = field(default_factory=list)  # (b, a) or None per band

    def path(self, p):
        return p if os.path.isabs(p) else os.path.join(self.root, p)


def read_kernel_conf(path):
    """TomoKernel.conf (init_kernel and init_filter, SI_ker_sta.F90:1047-1112)."""
    with open(path) as f:
        text = f.read()
    conf = fio.Conf(text)
    root = os.path.dirname(os.path.abspath(path))
    row = lambda k: tuple(conf.row(k, 2, 3, int))  # noqa: E731
    num_freq = conf.get("number_of_freqband", 2, int)
    kc = KernelConf(root=root, main_conf=conf.get("MAIN_CONF"), kid=conf.get("snap_id", 2, int),
                    first_run=conf.get("first_run", 2, bool), bstart=row("sub_start"), bcount=row("sub_count"),
                    bstride=row("sub_stride"), blkfst=row("block_from"), blksiz=row("block_size"),
                    knt=conf.get("SGT_TIMEDIM_SIZE", 2, int), M0=conf.get("SGT_M0", 2, float),
                    sgt_root=conf.get("SGT_ROOT"),
                    sgt_sta=(conf.get("SGT_STAX"), conf.get("SGT_STAY"), conf.get("SGT_STAZ")),
                    event_root=conf.get("EVENT_ROOT"), event_outd=conf.get("EVENT_OUTD"),
                    num_pick=conf.get("maximum_of_pick_per_component", 2, int))
    if kc.knt < 2:
        raise ValueError("SGT_TIMEDIM_SIZE must be >= 2")
    # the file names follow the line whose first word is <filter_file>
    lines = text.splitlines()
    at = next((q for q, ln in enumerate(lines) if fio._words(ln)[:1] == ["<filter_file>"]), None)
    if at is None and num_freq > 0:
        raise KeyError("TomoKernel.conf: there is no <filter_file> on column 1")
    names = [ln.split()[0].strip("'\"") for ln in lines[at + 1:] if ln.strip()][:num_freq] if num_freq else []
    if len(names) < num_freq:
        raise ValueError(f"TomoKernel.conf: {num_freq} bands but {len(names)} filter files")
    kc.bands = [None if n == "none" else read_filter(kc.path(n)) for n in names]
    return kc


@dataclass
class Grid:
    """The kernel sub-grid in the snapshot's index space and its blocks."""
    snap: dict
    tcnt: int
    stept: float
    media: str
    subnc_s: tuple
    subnc_c: tuple
    subnc_t: tuple
    blksiz: tuple
    blknum: tuple

    @property
    def kdt(self):
        return np.float32(self.snap["tinv"]) * np.float32(self.stept)


def plan_grid(kc):
    """The snapshot of snap_id in the main conf and the kernel sub-grid of
    one rank holding all of it (SI_ker_sta.F90:188-247): start bstart, count
    (snap_subc - bstart)/bstride + 1 capped by sub_count (-1: all), stride
    bstride; block_size -1 or above the count is the count."""
    conf = fio.read_conf(kc.path(kc.main_conf))
    ni, nj, nk = (conf.get(k, 2, int) for k in ("ni", "nj", "nk"))
    snaps, tcnts = fio._snapshot_defs(conf, ni, nj, nk, 3 + nk)
    if not 1 <= kc.kid <= len(snaps):
        raise ValueError(f"snap_id {kc.kid} not in 1..{len(snaps)}")
    snap, tcnt = snaps[kc.kid - 1], tcnts[kc.kid - 1]
    if not all(f in snap["fields"] for f in STRESS):
        raise ValueError(f"snapshot {kc.kid} holds no stress fields")
    s, c, t = [], [], []
    for d in range(3):
        if kc.bstart[d] < 1 or kc.bstride[d] < 1 or kc.bstart[d] > snap["subc"][d]:
            raise ValueError(f"sub_start/sub_stride {kc.bstart} {kc.bstride} outside snapshot {snap['subc']}")
        n = (snap["subc"][d] - kc.bstart[d]) // kc.bstride[d] + 1
        if kc.bcount[d] != -1 and kc.bcount[d] < n:
            n = kc.bcount[d]
        s.append(kc.bstart[d])
        c.append(n)
        t.append(kc.bstride[d])
    bs = tuple(cc if (b == -1 or b > cc) else b for b, cc in zip(kc.blksiz, c))
    if min(bs) < 1 or min(c) < 1:
        raise ValueError(f"empty kernel grid: count {c}, block size {bs}")
    nb = tuple((cc + b - 1) // b for cc, b in zip(c, bs))
    return Grid(snap, tcnt, conf.get("stept", 2, float), kc.path(conf.get("MEDIA_ROOT")),
                tuple(s), tuple(c), tuple(t), bs, nb)


def blocks(kc, g):
    """(index, bsubs, bsubc, bsubt) of each block in the reference's order
    (k, j, i from block_from, i fastest; :277-328): bsubs in the snapshot's
    1-based index space."""
    for bk in range(kc.blkfst[2], g.blknum[2] + 1):
        for bj in range(kc.blkfst[1], g.blknum[1] + 1):
            for bi in range(kc.blkfst[0], g.blknum[0] + 1):
                b = (bi, bj, bk)
                bsubs = tuple((b[d] - 1) * g.blksiz[d] * g.subnc_t[d] + g.subnc_s[d] for d in range(3))
                bsubc = tuple(min(g.blksiz[d], g.subnc_c[d] - (bsubs[d] - g.subnc_s[d]) // g.subnc_t[d])
                              for d in range(3))
                yield b, bsubs, bsubc, tuple(g.subnc_t)


def pick_window(T, t1, t2, kdt, knt):
    """1-based (nt1, nt2) of a pick, or None when the pick is skipped
    (SI_ker_sta.F90:686-710)."""
    T = np.asarray(T, np.float32)
    t1, t2 = np.float32(t1), np.float32(t2)
    tmin, tmax = np.float32(kdt), np.float32(knt) * np.float32(kdt)
    if t1 > tmax:
        return None
    if t1 <= tmin:
        n1 = 1
    else:
        idx = np.nonzero(T <= t1)[0]
        if not len(idx):
            raise ValueError(f"no synthetic time <= t1 = {t1}")
        n1 = int(idx[np.argmax(T[idx])]) + 1
    if t2 < tmin:
        return None
    if t2 >= tmax:
        n2 = knt
    else:
        idx = np.nonzero(T >= t2)[0]
        if not len(idx):
            raise ValueError(f"no synthetic time >= t2 = {t2}")
        n2 = int(idx[np.argmin(T[idx])]) + 1
    return n1, n2


@dataclass
class Pick:
    band: int
    filt: str
    t1: float
    t2: float
    outdir: str


@dataclass
class Event:
    name: str
    picknum: tuple
    rot: np.ndarray  # rot[r, c] = rot_mat(r+1, c+1)
    syn: tuple


def _records(f):
    """Lines of an open file with '#' lines skipped (blank lines too)."""
    for line in f:
        s = line.strip()
        if s and not s.startswith("#"):
            yield s


def parse_event_header(line):
    """KEVNM picknum(6) rot_mat(9) (:364-384)."""
    w = line.split()
    if len(w) < 16:
        raise ValueError(f"event line needs a name, 6 pick counts and 9 rotation values: {line!r}")
    return w[0], tuple(int(v) for v in w[1:7]), np.array([float(v) for v in w[7:16]],
                                                         np.float32).reshape(3, 3, order="F")


def parse_pick(line, num_freq):
    """band filter t1 t2 outdir (:655-713)."""
    w = line.split(None, 4)
    if len(w) < 5:
        raise ValueError(f"pick line needs band, filter, t1, t2 and an output directory: {line!r}")
    band = int(w[0])
    if band > num_freq:
        raise ValueError("frequency band id overflow")
    if w[1] not in FILTERS:
        raise ValueError(f"pickfilt error: {w[1]!r} is not P1, P2 or N")
    return Pick(band, w[1], float(w[2]), float(w[3]), w[4].strip())


def read_synthetic(path, knt):
    """T, V, U of a synthetic file of knt lines (init_synthetic :1114-1125)."""
    a = np.loadtxt(path, dtype=np.float64, max_rows=knt, ndmin=2)
    if a.shape[0] < knt or a.shape[1] < 3:
        raise ValueError(f"{path}: {knt} lines of T V U wanted, {a.shape} read")
    a = a.astype(np.float32)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


def sgt_file(root, kid, n):
    return os.path.join(root, f"sgt_{kid:03d}_{fio.SUFFIX}_n{n:05d}.nc")


def read_sgt(root, kid, tcnt, knt, bsubs, bsubc, bsubt):
    """(6, knt, npt) stress records of the block, i fastest, from the
    files sgt_<kid>_..._n00001.. of tcnt records each (:405-463): file n
    gives its first min(tcnt, knt - read) records."""
    npt = int(np.prod(bsubc))
    out = np.empty((6, knt, npt), np.float32)
    sl = tuple(slice(s - 1, s + (c - 1) * t, t) for s, c, t in zip(bsubs[::-1], bsubc[::-1], bsubt[::-1]))
    m, n = 0, 0
    per = tcnt if tcnt < knt else knt
    while m < knt:
        n += 1
        mt = min(per, knt - m)
        path = sgt_file(root, kid, n)
        with netcdf_file(path, "r", mmap=False) as nc:
            for q, v in enumerate(STRESS):
                var = nc.variables[v]
                if var.shape[0] < mt:
                    raise ValueError(f"{path}: {v} has {var.shape[0]} records, {mt} wanted")
                out[q, m:m + mt] = np.asarray(var[(slice(0, mt),) + sl], np.float32).reshape(mt, npt)
                del var
        m += mt
    return out


def read_media(path, subs, subc, subt):
    """lambda, mu at the block's points (media_mpi000000.nc, ghosted
    1-based indices), i fastest."""
    sl = tuple(slice(s - 1, s + (c - 1) * t, t) for s, c, t in zip(subs[::-1], subc[::-1], subt[::-1]))
    with netcdf_file(fio._fnm(path, "media"), "r", mmap=False) as nc:
        lam = np.asarray(nc.variables["lambda"][sl], np.float32).ravel()
        mu = np.asarray(nc.variables["mu"][sl], np.float32).ravel()
    return lam, mu


def _rotate_host(X, Y, Z, R):
    """syn_rotate (:1339-1352): v' = R^T v in FP32, left to right."""
    return (X * R[0, 0] + Y * R[1, 0] + Z * R[2, 0], X * R[0, 1] + Y * R[1, 1] + Z * R[2, 1],
            X * R[0, 2] + Y * R[1, 2] + Z * R[2, 2])


@dataclass
class Stats:
    kernels: int = 0
    points: int = 0
    t_read: float = 0.0
    t_gpu: float = 0.0
    t_write: float = 0.0
    files: list = field(default_factory=list)
    skipped: int = 0


def run(kconf_path, sgt_list=None, log=print):
    """Run the kernel stage; returns Stats (kernels written, timings)."""
    kc = read_kernel_conf(kconf_path)
    g = plan_grid(kc)
    knt, kdt, M0 = kc.knt, g.kdt, kc.M0
    T0 = (np.arange(1, knt + 1, dtype=np.float32) * kdt).astype(np.float32)
    num_freq = len(kc.bands)
    st = Stats()
    kname = f"kernel_{kc.kid:03d}_{fio.SUFFIX}_n00000.nc"
    with open(kc.path(sgt_list)) as sf:
        stations = [ln.split()[:2] for ln in _records(sf)]
    for ksta, evt_list in stations:
        log(f"station {ksta}: {evt_list}")
        for bidx, bsubs, bsubc, bsubt in blocks(kc, g):
            npt = int(np.prod(bsubc))
            subs = tuple(g.snap["subs"][d] + (bsubs[d] - 1) * g.snap["subt"][d] for d in range(3))
            subt = tuple(g.snap["subt"][d] * bsubt[d] for d in range(3))
            t = time.perf_counter()
            lam_h, mu_h = read_media(g.media, subs, bsubc, subt)
            st.t_read += time.perf_counter() - t
            lam, mu = DeviceArray.from_host(lam_h), DeviceArray.from_host(mu_h)
            sta = [None, None, None]  # station strain per receiver component, read once per block
            with open(kc.path(evt_list)) as ef:
                rec = _records(ef)
                raw = iter(ef)
                for line in rec:
                    ev = Event(*parse_event_header(line), syn=tuple(next(raw).strip() for _ in range(3)))
                    pk = ev.picknum
                    need = [pk[c] > 0 or any(p > 0 for p in pk[3:6]) for c in range(3)]
                    V, U, T = [None] * 3, [None] * 3, T0
                    for c in range(3):
                        if need[c]:
                            T, V[c], U[c] = read_synthetic(kc.path(ev.syn[c]), knt)
                    if all(p <= 0 for p in pk):
                        continue
                    t = time.perf_counter()
                    for c in range(3):
                        if need[c] and sta[c] is None:
                            h = read_sgt(kc.path(os.path.join(kc.sgt_root, ksta, kc.sgt_sta[c])), kc.kid, g.tcnt,
                                         knt, bsubs, bsubc, bsubt)
                            st.t_read += time.perf_counter() - t
                            sta[c] = kernel.stress2strain(DeviceArray.from_host(h), lam, mu)
                            t = time.perf_counter()
                    h = read_sgt(kc.path(os.path.join(kc.event_root, ev.name, kc.event_outd)), kc.kid, g.tcnt, knt,
                                 bsubs, bsubc, bsubt)
                    st.t_read += time.perf_counter() - t
                    t = time.perf_counter()
                    S = kernel.stress2strain(DeviceArray.from_host(h), lam, mu)
                    del h
                    A, B = [None] * 3, [None] * 3
                    for c in range(3):
                        if need[c]:
                            A[c], B[c] = kernel.sgt2scatter(sta[c], S, lam, mu, kdt)
                    del S
                    st.t_gpu += time.perf_counter() - t
                    for ncmp in range(6):
                        c = ncmp % 3
                        if ncmp == 3 and any(p > 0 for p in pk[3:6]):
                            V = list(_rotate_host(*V, ev.rot))
                            U = list(_rotate_host(*U, ev.rot))
                            kernel.scatter_rotate(A[0], A[1], A[2], ev.rot)
                            kernel.scatter_rotate(B[0], B[1], B[2], ev.rot)
                        if pk[ncmp] <= 0:
                            continue
                        picks = [parse_pick(next(rec), num_freq) for _ in range(pk[ncmp])]
                        plan = []
                        for p in picks:
                            w = pick_window(T, p.t1, p.t2, kdt, knt)
                            if w is None:
                                log(f"  skip pick {p} (window outside 1..{knt})")
                                st.skipped += 1
                                plan.append(None)
                                continue
                            out = os.path.join(kc.path(p.outdir), kname)
                            if kc.first_run and bidx == (1, 1, 1):
                                os.makedirs(os.path.dirname(out), exist_ok=True)
                                fio.grid3d_create(out, g.snap["subc"], KERNEL_VARS, "Finite frequency kernel")
                            elif not os.path.exists(out):
                                raise FileNotFoundError(f"{out}: kernel file missing (first_run is off or the run "
                                                        "does not start at block 1 1 1)")
                            plan.append((p, w, out))
                        cache = {}
                        for item in plan:
                            if item is None:
                                continue
                            p, (nt1, nt2), out = item
                            t = time.perf_counter()
                            key = (p.band, p.filt)
                            if key not in cache:
                                mode = FILTERS[p.filt]
                                if mode == 0:
                                    cache = {key: (V[c], U[c], A[c], B[c])}
                                else:
                                    band = kc.bands[p.band - 1] if p.band >= 1 else None
                                    if band is None:
                                        raise ValueError(f"band {p.band} has no filter file")
                                    b, a = band
                                    cache = {key: (kernel.filter_series_host(V[c], b, a, mode),
                                                   kernel.filter_series_host(U[c], b, a, mode),
                                                   kernel._filter(mode, A[c], b, a),
                                                   kernel._filter(mode, B[c], b, a))}
                            VV, UU, Sa, Sb = cache[key]
                            K = [k.to_host() for k in kernel.scatter2kernel(VV, UU, Sa, Sb, nt1, nt2, kdt, M0)]
                            st.t_gpu += time.perf_counter() - t
                            t = time.perf_counter()
                            for name, k in zip(KERNEL_VARS, K):
                                fio.grid3d_put(out, name, k, bsubs, bsubc, bsubt)
                            st.t_write += time.perf_counter() - t
                            st.kernels += 1
                            st.points += npt
                            if out not in st.files:
                                st.files.append(out)
            log(f"  block {bidx} of {g.blknum}: {npt} points")
    return st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("conf", nargs="?", default="TomoKernel.conf")
    ap.add_argument("--sgt-list", help="station list (default: the first line of stdin)")
    a = ap.parse_args(argv)
    sgt_list = a.sgt_list
    if sgt_list is None:
        sgt_list = sys.stdin.readline().strip()
        mpi = sys.stdin.readline().split()  # the non-MPI build's "input mpi id:"
        if len(mpi) >= 3 and any(int(v) != 0 for v in mpi[:3]):
            raise SystemExit("one rank: the mpi id must be 0 0 0")
    if not sgt_list:
        raise SystemExit("no station list given")
    t = time.perf_counter()
    st = run(a.conf, sgt_list)
    dt = time.perf_counter() - t
    print(f"{st.kernels} kernel blocks ({st.points} points) in {dt:.2f} s: read {st.t_read:.2f} s, "
          f"device {st.t_gpu:.2f} s, write {st.t_write:.2f} s; {len(st.files)} files")


if __name__ == "__main__":
    main()
