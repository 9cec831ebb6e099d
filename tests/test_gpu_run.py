"""A reference-style run from SeisFD3D.conf and NetCDF inputs
(python -m fwant_amd.run): the seismo_ and sgt_ files it writes hold the
records of the same solve driven through the Python binding, and the
receivers match the FP32 oracle to 1e-5."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from fwant_amd import io, models

pytestmark = pytest.mark.gpu


def test_run_from_conf(gpu_lib, tmp_path):
    from fwant_amd import run
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver, build_metric
    m = models.small_test(ni=30, nj=28, nk=24, nt=47, pml=6, seed=71)
    met = {k: v.astype(np.float32) for k, v in build_metric(m).items()}
    snaps = [dict(subs=(6, 6, 8), subc=(10, 9, 7), subt=(2, 2, 2), tinv=4,
                  fields=("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"))]
    conf = io.write_inputs(m, str(tmp_path), metric=met, snapshots=snaps, snap_tcnt=[5])
    print("fwant_amd.run on small_test 30x28x24, nt = 47 (48 steps)", flush=True)
    files = run.run(conf)
    out = os.path.join(tmp_path, "output")
    with netcdf_file(os.path.join(out, "seismo_mpi000000.nc"), "r", mmap=False) as nc:
        rec = np.stack([np.array(nc.variables[v][:]) for v in io.SEISMO_VARS], axis=-1)
    setup = io.load_run(conf)
    steps = setup.model.ref_steps()
    assert steps == 48 and rec.shape == (48, len(m.rec_indx), 9)
    s = Solver(setup.model, metric=setup.metric, snapshots=[dict(snaps[0], max_records=20)])
    s.run(steps)
    np.testing.assert_array_equal(rec, s.receivers())
    sg = s.snapshot(0)
    assert sg.shape[0] == 12 and len([f for f in files if "sgt_001" in f]) == 3
    with netcdf_file(os.path.join(out, "sgt_001_mpi000000_n00002.nc"), "r", mmap=False) as nc:
        np.testing.assert_array_equal(np.array(nc.variables["Tzz"][:]).reshape(5, -1), sg[5:10, 2])
    s.close()
    o = OracleSolver(setup.model, 32, metric=setup.metric)
    ref = o.run(steps).astype(np.float64)
    o.close()
    e = np.linalg.norm(rec - ref) / np.linalg.norm(ref)
    print(f"run from conf vs FP32 oracle: rel-L2 {e:.2e}")
    assert e <= 1e-5


def test_run_with_checkpoints_and_restart(gpu_lib, tmp_path):
    """checkpoint_tinv writes rest_t<ntime> every tinv steps and deletes the
    previous one; run_from_checkpoint continues from it with the reference's
    restart semantics (9 wavefields, CFS-PML memory at zero, F/B cycle
    restarted), which the FP32 oracle doing the same matches to 1e-5."""
    from fwant_amd import run
    from fwant_amd._lib import FIELDS
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver, build_metric
    m = models.small_test(ni=30, nj=28, nk=24, nt=15, pml=6, seed=72)
    met = {k: v.astype(np.float32) for k, v in build_metric(m).items()}
    conf = io.write_inputs(m, str(tmp_path), metric=met, checkpoint_tinv=8)
    print("fwant_amd.run: 16 steps with checkpoint_tinv = 8", flush=True)
    run.run(conf)
    ck = os.path.join(tmp_path, "checkpoint")
    assert sorted(os.listdir(ck)) == ["rest_t00016_mpi000000.nc"]
    first = Solver(io.load_run(conf).model, metric=met)
    first.run(16)
    fields, _, _ = io.read_rest(os.path.join(ck, "rest_t00016_mpi000000.nc"))
    for f in FIELDS:
        np.testing.assert_array_equal(fields[f], first.wavefield(f))
    first.close()
    m.nt = 31
    conf = io.write_inputs(m, str(tmp_path), metric=met, checkpoint_tinv=8, run_from_checkpoint=16)
    print("fwant_amd.run: restart from step 16, nt = 31", flush=True)
    run.run(conf)
    assert sorted(os.listdir(ck)) == ["rest_t00032_mpi000000.nc"]
    with netcdf_file(os.path.join(tmp_path, "output", "seismo_mpi000000.nc"), "r", mmap=False) as nc:
        rec = np.stack([np.array(nc.variables[v][:]) for v in io.SEISMO_VARS], axis=-1)
        t = np.array(nc.variables["time"][:])
    assert rec.shape[0] == 16
    np.testing.assert_allclose(t, (16 + np.arange(1, 17) * m.rec_tinv) * m.stept, rtol=1e-6)
    o = OracleSolver(m, 32, metric=met)
    for f in FIELDS:
        o.set_wavefield(f, fields[f])
    o.set_cycle_origin(16)
    o.ntime = 16
    ref = o.run(16).astype(np.float64)
    o.close()
    e = np.linalg.norm(rec - ref) / np.linalg.norm(ref)
    print(f"restarted run vs FP32 oracle: rel-L2 {e:.2e}")
    assert e <= 1e-5
