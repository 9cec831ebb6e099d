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
