"""The reference's run files: SeisFD3D.conf parsing with string_conf's rules,
the input NetCDF schemas (round trip Model -> files -> Model), and the
output schemas and file naming of seismo_ and sgt_/vel_ files."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from fwant_amd import io, models

CONF = """# a comment line
ni = 40   # trailing comment
nj	36
nk|32
stept = 0.05
abs_number = 8 8 8 8 8 0
ni = 99
flag = .true.
"""


def test_conf_parsing_rules():
    c = io.Conf(CONF)
    assert c.get("ni", 2, int) == 40          # first matching line wins
    assert c.get("nj", 2, int) == 36          # tab separates
    assert c.get("nk", 2, int) == 32          # '|' separates
    assert c.get("stept", 2, float) == pytest.approx(0.05)
    assert c.row("abs_number", 2, 6, int) == [8, 8, 8, 8, 8, 0]
    assert c.get("flag", 2, bool) is True
    with pytest.raises(KeyError):
        c.get("nt")
    with pytest.raises(KeyError):
        c.get("ni", 3)


def small_case(tmp_path):
    from oracle.pyoracle import build_metric
    m = models.small_test(ni=24, nj=22, nk=20, nt=30, pml=5, seed=3)
    met = {k: v.astype(np.float32) for k, v in build_metric(m).items()}
    snaps = [dict(subs=(6, 5, 7), subc=(8, 7, 6), subt=(2, 2, 2), tinv=3,
                  fields=("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")),
             dict(subs=(4, 4, 3 + 20), subc=(24, 22, 1), subt=(1, 1, 1), tinv=5, fields=("Vx", "Vy", "Vz"))]
    path = io.write_inputs(m, str(tmp_path), metric=met, snapshots=snaps, snap_tcnt=[4, 100])
    return m, met, snaps, path


def test_inputs_round_trip(tmp_path):
    m, met, snaps, path = small_case(tmp_path)
    r = io.load_run(path)
    g = r.model
    assert (g.ni, g.nj, g.nk, g.nt) == (m.ni, m.nj, m.nk, m.nt)
    assert g.stept == pytest.approx(m.stept)
    assert g.freenode == m.freenode
    np.testing.assert_array_equal(g.x, m.x.astype(np.float32))
    np.testing.assert_array_equal(g.z, m.z.astype(np.float32))
    for a, b in ((g.lam, m.lam), (g.mu, m.mu), (g.rho, m.rho)):
        assert a.flags["F_CONTIGUOUS"] and a.shape == (m.nx, m.ny, m.nz)
        np.testing.assert_array_equal(a, b)
    for k in met:
        np.testing.assert_array_equal(r.metric[k], met[k])
    assert g.abs_number == m.abs_number
    np.testing.assert_allclose(g.abs_velocity, m.abs_velocity)
    for s, t in ((g.force, m.force), (g.moment, m.moment)):
        assert (s.n, s.ntw, s.stf_id) == (t.n, t.ntw, t.stf_id)
        np.testing.assert_array_equal(s.indx, t.indx)
        np.testing.assert_array_equal(s.shift, t.shift)
        np.testing.assert_array_equal(s.comps, t.comps)
        np.testing.assert_array_equal(s.stf_time, t.stf_time)
    np.testing.assert_array_equal(g.rec_indx, m.rec_indx)
    assert g.rec_tinv == m.rec_tinv
    assert [dict(s, fields=tuple(s["fields"])) for s in r.snapshots] == snaps
    assert r.snap_tcnt == [4, 100]
    # the files carry the schemas the reference's readers expect
    with netcdf_file(os.path.join(tmp_path, "input", "media", "media_mpi000000.nc"), "r", mmap=False) as nc:
        assert nc.variables["lambda"].dimensions == ("K", "J", "I")
    with netcdf_file(os.path.join(tmp_path, "input", "source", "source_mpi000000.nc"), "r", mmap=False) as nc:
        assert nc.variables["Fx"].dimensions == ("force_point", "force_time_window")
        assert int(nc.number_of_force) == m.force.n and int(nc.force_stf_id) == m.force.stf_id


def test_snapshot_index_conventions():
    c = io.Conf("number_of_snap = 2\nsnap_001 = -1 -1 -2 -1 -1 1 2 2 1 10 50 T\n"
                "snap_002 = 3 4 5 5 5 5 1 1 1 2 10 VT\n")
    defs, tcnt = io._snapshot_defs(c, 40, 36, 30, 3 + 30)
    a, b = defs
    assert a["subs"] == (4, 4, 33) and a["subc"] == (20, 18, 1) and a["subt"] == (2, 2, 1)
    assert a["fields"] == ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz") and tcnt == [50, 10]
    assert b["subs"] == (6, 7, 8) and len(b["fields"]) == 9


def test_output_schemas(tmp_path):
    rng = np.random.default_rng(0)
    rec = rng.standard_normal((7, 3, 9)).astype(np.float32)
    p = io.write_seismo(str(tmp_path), rec, 0.05, 2)
    assert os.path.basename(p) == "seismo_mpi000000.nc"
    with netcdf_file(p, "r", mmap=False) as nc:
        assert nc.dimensions["num_pt"] == 3 and nc.dimensions["time"] is None
        np.testing.assert_allclose(nc.variables["time"][:], np.arange(1, 8) * 2 * np.float32(0.05))
        assert nc.variables["Tzz"].dimensions == ("time", "num_pt")
        np.testing.assert_array_equal(nc.variables["Tzz"][:], rec[:, :, 5])
    snap = dict(subs=(6, 5, 7), subc=(4, 3, 2), subt=(2, 2, 2), tinv=3, fields=("Vx", "Vy", "Vz", "Txx", "Tyy",
                                                                             "Tzz", "Txy", "Txz", "Tyz"))
    recs = rng.standard_normal((5, 9, 24)).astype(np.float32)
    paths = io.write_snapshots(str(tmp_path), 1, snap, 2, recs, 0.05)
    names = sorted(os.path.basename(q) for q in paths)
    assert names == ["sgt_001_mpi000000_n00001.nc", "sgt_001_mpi000000_n00002.nc", "sgt_001_mpi000000_n00003.nc",
                     "vel_001_mpi000000_n00001.nc", "vel_001_mpi000000_n00002.nc", "vel_001_mpi000000_n00003.nc"]
    with netcdf_file(os.path.join(tmp_path, "sgt_001_mpi000000_n00002.nc"), "r", mmap=False) as nc:
        assert nc.variables["Txy"].dimensions == ("time", "K", "J", "I")
        assert nc.variables["Txy"].shape == (2, 2, 3, 4)
        np.testing.assert_array_equal(nc.variables["Txy"][1].ravel(), recs[3, 6])
        np.testing.assert_allclose(nc.variables["time"][:], [3 * 3 * np.float32(0.05), 4 * 3 * np.float32(0.05)])
        assert list(nc.subs) == [3, 2, 4] and list(nc.subc) == [4, 3, 2]
        assert float(nc.stept) == pytest.approx(0.15)


def test_restart_keys_and_step_plan(tmp_path):
    """run_from_checkpoint / checkpoint_tinv / CHECKPOINT_ROOT (mod_io.F90:101-103)
    and the steps of a reference-style run: blocks of 8 until ntime > nt,
    rest files after every step whose ntime is a multiple of checkpoint_tinv."""
    from oracle.pyoracle import build_metric
    m = models.small_test(ni=20, nj=18, nk=16, nt=31, pml=4)
    met = {k: v.astype(np.float32) for k, v in build_metric(m).items()}
    conf = io.write_inputs(m, str(tmp_path), metric=met, checkpoint_tinv=8, run_from_checkpoint=16)
    s = io.load_run(conf)
    assert (s.checkpoint_tinv, s.run_from_checkpoint) == (8, 16)
    assert s.checkpoint_root == os.path.join(str(tmp_path), "checkpoint")
    assert io.run_plan(31, 16, 8) == (16, [24, 32])
    assert io.run_plan(15, 0, 8) == (16, [8, 16])
    assert io.run_plan(47, 0, 0) == (48, [])
    assert io.run_plan(1000, 0, 0)[0] == 1008
    assert io.run_plan(20, 17, 5) == (8, [20, 25])
    assert io.run_plan(10, 11, 5) == (0, [])
    plain = io.load_run(io.write_inputs(m, str(tmp_path / "p"), metric=met))
    assert (plain.checkpoint_tinv, plain.run_from_checkpoint) == (0, 0)
