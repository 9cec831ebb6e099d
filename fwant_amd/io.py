"""The reference's run inputs and outputs: the SeisFD3D.conf parameter file
and the per-rank NetCDF-classic files (single rank: suffix mpi000000).

  read_conf / Conf           string_conf semantics (code/srcF/mod_string.F90:
                             38-66, 140-186): '#' starts a comment, '=', tab
                             and '|' separate words, column 1 is the key
  load_run(conf_path)        SeisFD3D.conf + coord_/metric_/media_/source_/
                             station_ files -> (Model, metric, snapshots,
                             output root), as para_init, grid_*_import,
                             media_import, src_import, io_pt_import,
                             abs_init and io_snap_read read them
  write_inputs(model, root)  the same files from a Model (the schemas the
                             preprocessors seis3d_grid/metric/source/station
                             and seis3d_media write)
  write_seismo(...)          seismo_mpi000000.nc (mod_nfseis.F90:81-126,
                             mod_io.F90:430-497)
  write_snapshots(...)       sgt_NNN_ / vel_NNN_mpi000000_nNNNNN.nc files of
                             tcnt records each (mod_nfseis.F90:131-203,
                             mod_io.F90:505-645, 825-836)
  grid3d_create / grid3d_put a 3-D grid file (kernel_NNN_ files) created
                             once and written block by block in place
                             (mod_nfseis.F90:208-255, 450-462)

NetCDF-classic reading and writing goes through scipy.io.netcdf_file,
except the in-place block writes of grid3d_put (numpy.memmap at the
variable's offset, read from the file's header).
Arrays keep the reference's Fortran order: a variable of Fortran shape
(I, J, K) is stored with the NetCDF dimensions (K, J, I).
"""
import os
import struct
from dataclasses import dataclass, field

import numpy as np
from scipy.io import netcdf_file

from . import models

SUFFIX = "mpi000000"
PUNCT = "=\t|"


# ---------------------------------------------------------------------------
# parameter file
# ---------------------------------------------------------------------------
def _words(line):
    line = line.lstrip()
    c = line.find("#")
    if c >= 0:
        line = line[:c]
    for ch in PUNCT:
        line = line.replace(ch, " ")
    return line.split()


class Conf:
    """Key -> words of the first line whose first word is the key (the
    reference's strip_conf_file scans from the top and stops there)."""

    def __init__(self, text):
        self.lines = {}
        for raw in text.splitlines():
            w = _words(raw)
            if w and w[0] not in self.lines:
                self.lines[w[0]] = w

    def has(self, key):
        return key in self.lines

    def get(self, key, col=2, typ=str, default=None):
        """Word `col` (1-based, the key is column 1) of the key's line."""
        if key not in self.lines:
            if default is not None:
                return default
            raise KeyError(f"SeisFD3D.conf: there is no {key} on column 1")
        w = self.lines[key]
        if col > len(w):
            raise KeyError(f"SeisFD3D.conf: {key} has only {len(w)} columns, column {col} wanted")
        v = w[col - 1]
        if typ is bool:
            return v.strip(".").upper() in ("T", "TRUE")
        if typ is int:
            return int(float(v)) if ("." in v or "e" in v.lower()) else int(v)
        return typ(v)

    def row(self, key, first, count, typ=float):
        return [self.get(key, first + q, typ) for q in range(count)]


def read_conf(path):
    with open(path) as f:
        return Conf(f.read())


# ---------------------------------------------------------------------------
# NetCDF helpers
# ---------------------------------------------------------------------------
def _read(path, names):
    with netcdf_file(path, "r", mmap=False) as nc:
        return {n: np.array(nc.variables[n][:]) for n in names}


def _fortran3(a):
    """NetCDF (K, J, I) variable -> Fortran-ordered (I, J, K) array."""
    return np.asfortranarray(np.asarray(a).T)


def _fnm(root, stem):
    return os.path.join(root, f"{stem}_{SUFFIX}.nc")


# ---------------------------------------------------------------------------
# inputs
# ---------------------------------------------------------------------------
@dataclass
class RunSetup:
    model: "models.Model"
    metric: dict
    snapshots: list = field(default_factory=list)
    output_root: str = "."
    snap_tcnt: list = field(default_factory=list)
    # restart keys (mod_io.F90:101-103): rest files under checkpoint_root
    # every checkpoint_tinv steps; run_from_checkpoint > 0 starts from
    # rest_t<that step>
    checkpoint_root: str = "."
    checkpoint_tinv: int = 0
    run_from_checkpoint: int = 0


def _sources(path, kind):
    """force / moment tables of source_mpi000000.nc (mod_src.F90:150-185)."""
    with netcdf_file(path, "r", mmap=False) as nc:
        n = int(getattr(nc, f"number_of_{kind}"))
        s = models.Sources()
        if n <= 0:
            if kind == "moment":
                s.comps = np.zeros((6, 0), np.float32)
            return s
        ntw = nc.dimensions[f"{kind}_time_window"]
        s.n, s.ntw = n, int(ntw)
        s.stf_id = int(getattr(nc, f"{kind}_stf_id"))
        s.indx = np.array(nc.variables[f"{kind}_indx"][:], np.int32).reshape(n, 3)
        s.shift = np.array(nc.variables[f"{kind}_shift"][:], np.float32).reshape(n, 3)
        s.t0 = np.array(nc.variables[f"{kind}_start_time"][:], np.float32).reshape(n)
        s.stf_time = np.array(nc.variables[f"{kind}_stf_time"][:], np.float32).reshape(s.ntw)
        s.stf_freq = np.array(nc.variables[f"{kind}_stf_freq"][:], np.float32).reshape(s.ntw)
        names = ("Fx", "Fy", "Fz") if kind == "force" else ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz")
        # Fortran (ntw, n): window index fastest
        s.comps = np.stack([np.array(nc.variables[v][:], np.float32).reshape(-1) for v in names])
        return s


def _snapshot_defs(conf, ni, nj, nk, nk2_1based):
    """snap_NNN lines (mod_io.F90:180-199) with corr_indx (:300-340): -1 =
    first interior index / to the end, -2 = the free surface; indices in
    the line are 1-based interior indices."""
    out, tcnt = [], []
    n = conf.get("number_of_snap", 2, int, default=0)
    first = (4, 4, 4)
    last = (3 + ni, 3 + nj, 3 + nk)
    for q in range(1, n + 1):
        key = f"snap_{q:03d}"
        subs = conf.row(key, 2, 3, int)
        subc = conf.row(key, 5, 3, int)
        subt = conf.row(key, 8, 3, int)
        tinv = conf.get(key, 11, int)
        cnt = conf.get(key, 12, int)
        var = conf.get(key, 13, str)
        s = []
        for d in range(3):
            if d == 2 and subs[d] == -2:
                s.append(nk2_1based)
            elif subs[d] == -1:
                s.append(first[d])
            else:
                s.append(subs[d] + 3)
        c = list(subc)
        for d in range(3):
            if c[d] == -1:
                c[d] = (last[d] - s[d]) // subt[d] + 1
        fields = []
        if "V" in var:
            fields += ["Vx", "Vy", "Vz"]
        if "T" in var:
            fields += ["Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"]
        out.append(dict(subs=tuple(s), subc=tuple(c), subt=tuple(subt), tinv=tinv, fields=tuple(fields)))
        tcnt.append(cnt)
    return out, tcnt


def load_run(conf_path):
    """Everything the reference driver reads before its time loop."""
    conf = read_conf(conf_path)
    base = os.path.dirname(os.path.abspath(conf_path))
    root = lambda k: os.path.join(base, conf.get(k, 2))  # noqa: E731
    dims = conf.row("dims", 2, 3, int) if conf.has("dims") else [1, 1, 1]
    if dims != [1, 1, 1]:
        raise ValueError("this build runs one rank per solve: dims must be 1 1 1")
    ni, nj, nk, nt = (conf.get(k, 2, int) for k in ("ni", "nj", "nk", "nt"))
    stept = conf.get("stept", 2, float)
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    c = _read(_fnm(root("GRID_ROOT"), "coord"), ("x", "y", "z"))
    met = _read(_fnm(root("GRID_ROOT"), "metric"), ("xsin", "xcot", "xi_x", "eta_y", "zeta_z"))
    med = _read(_fnm(root("MEDIA_ROOT"), "media"), ("lambda", "mu", "rho"))
    for k, n in (("x", nx), ("y", ny), ("z", nz)):
        if c[k].shape != (n,):
            raise ValueError(f"coord {k} has {c[k].shape}, expected ({n},)")
    abs_number = tuple(conf.row("abs_number", 2, 6, int))
    m = models.Model("conf", ni, nj, nk, stept, nt, abs_number[5] <= 0,
                     c["x"].astype(np.float64), c["y"].astype(np.float64), c["z"].astype(np.float64),
                     _fortran3(med["lambda"]).astype(np.float32), _fortran3(med["mu"]).astype(np.float32),
                     _fortran3(med["rho"]).astype(np.float32),
                     abs_number=abs_number,
                     abs_velocity=tuple(conf.row("abs_velocity", 2, 6, float)),
                     cfs_bmax=tuple(conf.row("CFS_bmax", 2, 6, float)),
                     cfs_amax=tuple(conf.row("CFS_amax", 2, 6, float)))
    src = _fnm(root("SOURCE_ROOT"), "source")
    m.force = _sources(src, "force")
    m.moment = _sources(src, "moment")
    st = _fnm(root("STATION_ROOT"), "station")
    with netcdf_file(st, "r", mmap=False) as nc:
        npt = nc.variables["indx"].shape[0]
        m.rec_indx = np.array(nc.variables["indx"][:], np.int32).reshape(npt, 3)
        m.rec_tinv = int(nc.tinv)
    metric = {k: v.astype(np.float32) for k, v in met.items()}
    snaps, tcnt = _snapshot_defs(conf, ni, nj, nk, 3 + nk)
    ck_root = root("CHECKPOINT_ROOT") if conf.has("CHECKPOINT_ROOT") else root("OUTPUT_ROOT")
    ck_tinv = conf.get("checkpoint_tinv", 2, int) if conf.has("checkpoint_tinv") else 0
    ck_from = conf.get("run_from_checkpoint", 2, int) if conf.has("run_from_checkpoint") else 0
    if ck_tinv < 0 or ck_from < 0:
        raise ValueError("checkpoint_tinv and run_from_checkpoint must be >= 0")
    return RunSetup(m, metric, snaps, root("OUTPUT_ROOT"), tcnt, ck_root, ck_tinv, ck_from)


def rest_path(root, ntime):
    """rest_tNNNNN_mpi000000.nc of step ntime (get_fnm_rest, mod_io.F90)."""
    return os.path.join(root, f"rest_t{ntime:05d}_{SUFFIX}.nc")


def run_plan(nt, ntime0, tinv):
    """Steps of a reference-style run from ntime0 and the steps after which
    it writes rest files.  The driver tests ntime > nt only at the top of
    each 8-step block (seis3d_wave_8op_p3.F90:92-95), so it runs
    8*ceil((nt - ntime0 + 1)/8) steps; io_rest_export writes rest_t<ntime>
    after every step whose new ntime is a multiple of checkpoint_tinv
    (mod_io.F90:364-374; tinv 0: never)."""
    steps = 8 * ((nt - ntime0 + 1 + 7) // 8) if ntime0 <= nt else 0
    marks = [t for t in range(ntime0 + 1, ntime0 + steps + 1) if tinv > 0 and t % tinv == 0]
    return steps, marks


STF_NAMES = {100: "bell_int", 110: "triangle_int", 120: "gauss_int", 200: "ricker", 210: "B(shift)",
             220: "gauss", 230: "bell", 240: "step", 250: "delta"}  # stf_id2name, mod_src.F90:651-677


def write_coord(path, x, y, z):
    """coord_mpi000000.nc: x (colatitude, rad), y (longitude, rad), z
    (radius, m) on the ghosted axes (seis3d_grid.F90:135-177)."""
    with netcdf_file(path, "w") as nc:
        for n, v in zip("IJK", (len(x), len(y), len(z))):
            nc.createDimension(n, v)
        for n, d, v in (("x", "I", x), ("y", "J", y), ("z", "K", z)):
            nc.createVariable(n, "f4", (d,))[:] = np.asarray(v, np.float32)
    return path


def write_metric(path, metric, nx, ny, nz):
    with netcdf_file(path, "w") as nc:
        for n, v in zip("IJK", (nx, ny, nz)):
            nc.createDimension(n, v)
        for n, d in (("xsin", "I"), ("xcot", "I"), ("xi_x", "I"), ("eta_y", "J"), ("zeta_z", "K")):
            nc.createVariable(n, "f4", (d,))[:] = np.asarray(metric[n], np.float32)
    return path


def write_media(path, lam, mu, rho):
    """media_mpi000000.nc: lambda, mu, rho on the ghosted grid (media_create)."""
    nx, ny, nz = np.asarray(lam).shape
    with netcdf_file(path, "w", version=2) as nc:
        for n, v in zip("IJK", (nx, ny, nz)):
            nc.createDimension(n, v)
        for n, a in (("lambda", lam), ("mu", mu), ("rho", rho)):
            nc.createVariable(n, "f4", ("K", "J", "I"))[:] = np.asarray(a, np.float32).T
    return path


def write_source(path, force, moment, force_axis=None, moment_axis=None):
    """source_mpi000000.nc (srcnode_skel and the puts of seis3d_source.F90:
    211-306, 428-493): per kind the point and time-window dimensions, the
    component tables (window fastest), 1-based ghosted indices, located
    coordinates (axis), shifts, start times, STF factors, id and name."""
    with netcdf_file(path, "w") as nc:
        nc.createDimension("geo_dimension", 3)
        for kind, s, names, axis in (("force", force, ("Fx", "Fy", "Fz"), force_axis),
                                     ("moment", moment, ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"), moment_axis)):
            setattr(nc, f"number_of_{kind}", np.int32(s.n))
            if s.n <= 0:
                continue
            nc.createDimension(f"{kind}_point", s.n)
            nc.createDimension(f"{kind}_time_window", s.ntw)
            for q, v in enumerate(names):
                nc.createVariable(v, "f4", (f"{kind}_point", f"{kind}_time_window"))[:] = \
                    np.asarray(s.comps[q], np.float32).reshape(s.n, s.ntw)
            nc.createVariable(f"{kind}_indx", "i4", (f"{kind}_point", "geo_dimension"))[:] = s.indx
            if axis is not None:
                nc.createVariable(f"{kind}_axis", "f4", (f"{kind}_point", "geo_dimension"))[:] = \
                    np.asarray(axis, np.float32).reshape(s.n, 3)
            nc.createVariable(f"{kind}_shift", "f4", (f"{kind}_point", "geo_dimension"))[:] = s.shift
            nc.createVariable(f"{kind}_start_time", "f4", (f"{kind}_point",))[:] = s.t0
            nc.createVariable(f"{kind}_stf_time", "f4", (f"{kind}_time_window",))[:] = s.stf_time
            nc.createVariable(f"{kind}_stf_freq", "f4", (f"{kind}_time_window",))[:] = s.stf_freq
            setattr(nc, f"{kind}_stf_id", np.int32(s.stf_id))
            setattr(nc, f"{kind}_stf_type", STF_NAMES.get(int(s.stf_id), "unknown"))
    return path


def write_station(path, indx, tinv, gindx=None, coord=None, grid=None, ids=None):
    """station_mpi000000.nc (station_skel, seis3d_station.F90:107-150):
    1-based ghosted indices, interior indices, requested and grid
    coordinates and (number, line) ids per receiver; tinv attribute."""
    indx = np.asarray(indx, np.int32).reshape(-1, 3)
    npt = len(indx)
    with netcdf_file(path, "w") as nc:
        nc.createDimension("num_pt", None)
        nc.createDimension("geo_dim", 3)
        nc.createDimension("twoelem", 2)
        vs = {"indx": ("i4", indx), "gindx": ("i4", gindx), "coord": ("f4", coord), "grid": ("f4", grid),
              "id": ("i4", ids)}
        for name, (t, a) in vs.items():
            if name != "indx" and a is None:
                continue
            v = nc.createVariable(name, t, ("num_pt", "twoelem" if name == "id" else "geo_dim"))
            if npt:
                v[:npt] = np.asarray(a, np.float32 if t == "f4" else np.int32)
        nc.tinv = np.int32(tinv)
    return path


def write_inputs(model, root, metric=None, snapshots=(), snap_tcnt=None, nt=None, checkpoint_tinv=0,
                 run_from_checkpoint=0):
    """Write SeisFD3D.conf and the input files of `model` under `root`
    (directories input/{grid,media,source,station}, output/ and, with the
    restart keys, checkpoint/)."""
    dirs = {k: os.path.join(root, "input", k) for k in ("grid", "media", "source", "station")}
    for d in list(dirs.values()) + [os.path.join(root, "output")]:
        os.makedirs(d, exist_ok=True)
    m = model
    write_coord(_fnm(dirs["grid"], "coord"), m.x, m.y, m.z)
    if metric is None:
        from .solver import build_metric
        metric = build_metric(m)
    write_metric(_fnm(dirs["grid"], "metric"), metric, m.nx, m.ny, m.nz)
    write_media(_fnm(dirs["media"], "media"), m.lam, m.mu, m.rho)
    write_source(_fnm(dirs["source"], "source"), m.force, m.moment)
    write_station(_fnm(dirs["station"], "station"), m.rec_indx, m.rec_tinv)
    nt = m.nt if nt is None else nt
    lines = [
        "# SeisFD3D.conf written by fwant_amd.io.write_inputs",
        "dims = 1 1 1",
        f"ni = {m.ni}", f"nj = {m.nj}", f"nk = {m.nk}", f"nt = {nt}",
        f"stept = {m.stept!r}",
        "abs_number = " + " ".join(str(int(v)) for v in m.abs_number),
        "abs_velocity = " + " ".join(repr(float(v)) for v in m.abs_velocity),
        "CFS_bmax = " + " ".join(repr(float(v)) for v in m.cfs_bmax),
        "CFS_amax = " + " ".join(repr(float(v)) for v in m.cfs_amax),
        "GRID_ROOT = input/grid", "MEDIA_ROOT = input/media", "SOURCE_ROOT = input/source",
        "STATION_ROOT = input/station", "OUTPUT_ROOT = output",
        f"number_of_snap = {len(snapshots)}",
    ]
    if checkpoint_tinv or run_from_checkpoint:
        os.makedirs(os.path.join(root, "checkpoint"), exist_ok=True)
        lines += ["CHECKPOINT_ROOT = checkpoint", f"checkpoint_tinv = {int(checkpoint_tinv)}",
                  f"run_from_checkpoint = {int(run_from_checkpoint)}"]
    for q, s in enumerate(snapshots, 1):
        var = ("V" if any(f.startswith("V") for f in s["fields"]) else "") + \
              ("T" if any(f.startswith("T") for f in s["fields"]) else "")
        subs = [v - 3 for v in s["subs"]]
        cnt = (snap_tcnt or [100] * len(snapshots))[q - 1]
        lines.append(f"snap_{q:03d} = " + " ".join(str(v) for v in (*subs, *s["subc"], *s["subt"], s["tinv"], cnt))
                     + f" {var}")
    path = os.path.join(root, "SeisFD3D.conf")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


# ---------------------------------------------------------------------------
# outputs
# ---------------------------------------------------------------------------
SEISMO_VARS = ("Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")


def write_seismo(root, records, stept, tinv, ntime0=0):
    """records [nrecord][num_pt][9]: record m (1-based) is step ntime0 +
    m*tinv at time (ntime0 + m*tinv)*stept (io_seismo_put; ntime0 > 0 for a
    run restarted from a rest file)."""
    records = np.asarray(records, np.float32)
    nrec, npt = records.shape[:2]
    path = _fnm(root, "seismo")
    with netcdf_file(path, "w") as nc:
        nc.title = "seismograms"
        nc.createDimension("time", None)  # the record dimension is defined first here
        nc.createDimension("num_pt", npt)
        t = nc.createVariable("time", "f4", ("time",))
        vs = [nc.createVariable(v, "f4", ("time", "num_pt")) for v in SEISMO_VARS]
        if nrec:
            t[:nrec] = ((ntime0 + np.arange(1, nrec + 1) * tinv) * np.float32(stept)).astype(np.float32)
            for f, v in enumerate(vs):
                v[:nrec] = records[:, :, f]
    return path


def write_snapshots(root, snap_no, snap, tcnt, records, stept):
    """Snapshot `snap_no` (1-based) of definition `snap` (fields in
    increasing field order, records [nrec][nfield][npts]) as files of tcnt
    records: sgt_NNN_* for the stress fields, vel_NNN_* for the velocities,
    file n0 holding records (n0-1)*tcnt+1 .. n0*tcnt."""
    from ._lib import FIELDS
    records = np.asarray(records, np.float32)
    nrec = records.shape[0]
    subc = tuple(int(v) for v in snap["subc"])
    order = [f for f in FIELDS if f in snap["fields"]]
    groups = (("vel_", ("Vx", "Vy", "Vz")), ("sgt_", ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")))
    paths = []
    subs = np.array(snap["subs"], np.int32) - 3  # interior (out_*) indices
    subt = np.array(snap["subt"], np.int32)
    sube = subs + subt * (np.array(subc, np.int32) - 1)
    for prefix, names in groups:
        if not any(n in order for n in names):
            continue
        for first in range(0, nrec, tcnt):
            n0 = first // tcnt + 1
            path = os.path.join(root, f"{prefix}{snap_no:03d}_{SUFFIX}_n{n0:05d}.nc")
            cnt = min(tcnt, nrec - first)
            with netcdf_file(path, "w", version=2) as nc:
                nc.title = "snap of velocity feilds" if prefix == "vel_" else "snap of stress feilds"
                nc.createDimension("time", None)
                for d, v in zip("IJK", subc):
                    nc.createDimension(d, v)
                nc.stept = np.float32(stept * snap["tinv"])
                for a, v in (("gsubs", subs), ("gsubc", subc), ("gsubt", subt), ("gsube", sube),
                             ("subs", subs), ("subc", subc), ("subt", subt), ("sube", sube)):
                    setattr(nc, a, np.asarray(v, np.int32))
                t = nc.createVariable("time", "f4", ("time",))
                t[:cnt] = (np.arange(first + 1, first + cnt + 1) * snap["tinv"] * np.float32(stept)).astype(np.float32)
                for n in names:
                    if n not in order:
                        continue
                    q = order.index(n)
                    v = nc.createVariable(n, "f4", ("time", "K", "J", "I"))
                    v[:cnt] = records[first:first + cnt, q].reshape(cnt, subc[2], subc[1], subc[0])
            paths.append(path)
    return paths


def write_rest(root, ntime, fields, state=None, cycle_origin=None):
    """rest_tNNNNN_mpi000000.nc (io_rest_export, mod_io.F90:376-399): the 9
    ghosted wavefields (dims I, J, K).  With `state` (Solver.checkpoint())
    the file also carries the complete state for an exact restart, which
    the reference's reader ignores."""
    path = rest_path(root, ntime)
    nx, ny, nz = np.asarray(fields["Vx"]).shape
    with netcdf_file(path, "w", version=2) as nc:
        nc.title = "Restart file generated by seis3d_wave"
        for d, v in zip("IJK", (nx, ny, nz)):
            nc.createDimension(d, v)
        for name in SEISMO_VARS:
            nc.createVariable(name, "f4", ("K", "J", "I"))[:] = np.asarray(fields[name], np.float32).T
        if state is not None:
            nc.createDimension("state", len(state))
            nc.createVariable("fwant_state", "f4", ("state",))[:] = np.asarray(state, np.float32)
            nc.cycle_origin = np.int32(0 if cycle_origin is None else cycle_origin)
        nc.ntime = np.int32(ntime)
    return path


# ---------------------------------------------------------------------------
# 3-D grid files written block by block (kernel_NNN_ files)
# ---------------------------------------------------------------------------
_NC_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8"}


def _pad4(b):
    return b + b"\0" * (-len(b) % 4)


def _nc_name(s):
    e = s.encode()
    return struct.pack(">I", len(e)) + _pad4(e)


def classic_layout(path):
    """{name: (begin, shape, big-endian dtype)} of the fixed-size variables
    of a classic (CDF-1) or 64-bit-offset (CDF-2) NetCDF file, parsed from
    its header (the header grammar of the NetCDF classic format
    specification); record variables are left out."""
    with open(path, "rb") as f:
        buf = f.read(1 << 20)
    if buf[:3] != b"CDF" or buf[3] not in (1, 2):
        raise ValueError(f"{path}: not a classic or 64-bit-offset NetCDF file")
    off64 = buf[3] == 2
    p = 8

    def u32():
        nonlocal p
        v = struct.unpack_from(">I", buf, p)[0]
        p += 4
        return v

    def name():
        nonlocal p
        n = u32()
        s = buf[p:p + n].decode()
        p += (n + 3) // 4 * 4
        return s

    def atts():
        nonlocal p
        u32()
        for _ in range(u32()):
            name()
            t = u32()
            ne = u32()
            p += (ne * np.dtype(_NC_TYPES[t]).itemsize + 3) // 4 * 4

    u32()
    dims = []
    for _ in range(u32()):
        dims.append((name(), u32()))
    atts()
    out = {}
    u32()
    for _ in range(u32()):
        nm = name()
        ids = [u32() for _ in range(u32())]
        atts()
        t = u32()
        u32()  # vsize
        if off64:
            begin = struct.unpack_from(">Q", buf, p)[0]
            p += 8
        else:
            begin = u32()
        shape = tuple(dims[i][1] for i in ids)
        if ids and dims[ids[0]][1] == 0:
            continue  # record variable
        out[nm] = (begin, shape, np.dtype(_NC_TYPES[t]))
    return out


def grid3d_create(path, subc, names, title):
    """A 64-bit-offset NetCDF file with dimensions I, J, K = subc and float
    variables `names` (K, J, I), all zero (nfseis_grid3d_def/defvar,
    code/srcF/mod_nfseis.F90:208-255; the reference creates its files with
    no fill, so points it never writes are undefined there).  The header is
    written here and the data region is extended with zeros, so a large
    file costs no memory."""
    ni, nj, nk = (int(v) for v in subc)
    vsize = 4 * ni * nj * nk
    hdr = b"CDF\x02" + struct.pack(">I", 0)
    hdr += struct.pack(">II", 0x0A, 3) + b"".join(_nc_name(d) + struct.pack(">I", n)
                                                  for d, n in zip("IJK", (ni, nj, nk)))
    t = title.encode()
    hdr += struct.pack(">II", 0x0C, 1) + _nc_name("title") + struct.pack(">II", 2, len(t)) + _pad4(t)
    var_len = sum(len(_nc_name(v)) + 4 + 12 + 8 + 8 + 8 for v in names)
    begin0 = len(hdr) + 8 + var_len
    hdr += struct.pack(">II", 0x0B, len(names))
    for q, v in enumerate(names):
        hdr += _nc_name(v) + struct.pack(">IIII", 3, 2, 1, 0) + struct.pack(">II", 0, 0)
        hdr += struct.pack(">II", 5, vsize) + struct.pack(">Q", begin0 + q * vsize)
    assert len(hdr) == begin0
    with open(path, "wb") as f:
        f.write(hdr)
        f.truncate(begin0 + len(names) * vsize)
    return path


def grid3d_put(path, name, data, subs, subc, subt):
    """Write `data` (the points of the sub-grid subs/subc/subt, 1-based,
    i fastest) into variable `name` of a grid file in place (nfseis_put,
    code/srcF/mod_nfseis.F90:450-462)."""
    begin, shape, dt = classic_layout(path)[name]
    sl = []
    for s, c, t, n in zip(subs[::-1], subc[::-1], subt[::-1], shape):
        if s < 1 or c < 1 or t < 1 or s + (c - 1) * t > n:
            raise ValueError(f"{path}: {name} sub-grid {subs} {subc} {subt} outside {shape[::-1]}")
        sl.append(slice(s - 1, s + (c - 1) * t, t))
    mm = np.memmap(path, dtype=dt, mode="r+", offset=begin, shape=shape)
    mm[tuple(sl)] = np.asarray(data, np.float32).reshape(subc[2], subc[1], subc[0])
    mm.flush()
    del mm


def read_rest(path):
    """(fields, state or None, cycle origin or None) of a rest_ file."""
    with netcdf_file(path, "r", mmap=False) as nc:
        fields = {n: _fortran3(np.array(nc.variables[n][:])) for n in SEISMO_VARS}
        state = np.array(nc.variables["fwant_state"][:]) if "fwant_state" in nc.variables else None
        origin = int(nc.cycle_origin) if hasattr(nc, "cycle_origin") else None
    return fields, state, origin
