"""Run a reference-style case: python -m fwant_amd.run SeisFD3D.conf

The flow of code/srcF/seis3d_wave_8op_p3.F90 with the solve on the GPU:
read the parameter file and the per-rank NetCDF inputs (fwant_amd.io),
optionally restart from CHECKPOINT_ROOT/rest_t<run_from_checkpoint>
(io_rest_import, mod_io.F90:405-427: the 9 wavefields, CFS-PML memory at
zero, the F/B cycle restarted at the loop top), run the steps the
reference driver runs for nt (8*ceil((nt - ntime0 + 1)/8): the exit test
sits at the top of each 8-step block), write rest_t<ntime> every
checkpoint_tinv steps and delete the previous one (io_rest_export,
mod_io.F90:342-403), then write seismo_mpi000000.nc and the sgt_/vel_
snapshot files under OUTPUT_ROOT.  After a restart the output files hold
the records of the restarted run, timed from ntime0.
"""
import argparse
import os
import sys
import time

from . import io
from ._lib import FIELDS
from .solver import Solver


def run(conf_path, graphs=True, log=print):
    setup = io.load_run(conf_path)
    m = setup.model
    ntime0 = setup.run_from_checkpoint
    steps, marks = io.run_plan(m.nt, ntime0, setup.checkpoint_tinv)
    snaps = [dict(s, max_records=steps // s["tinv"] + 1) for s in setup.snapshots]
    t0 = time.perf_counter()
    s = Solver(m, metric=setup.metric, snapshots=snaps, max_records=steps // max(m.rec_tinv, 1) + 1)
    s.use_graphs(graphs)
    if ntime0 > 0:
        fields, _, _ = io.read_rest(io.rest_path(setup.checkpoint_root, ntime0))
        for f in FIELDS:
            s.set_wavefield(f, fields[f])
        s.set_cycle_origin(ntime0)
    s.ntime = ntime0
    done = ntime0
    if marks:
        os.makedirs(setup.checkpoint_root, exist_ok=True)
    for t in marks:  # run to each rest step, export, delete the previous one
        s.run(t - done)
        done = t
        s.sync()
        io.write_rest(setup.checkpoint_root, t, {f: s.wavefield(f) for f in FIELDS})
        prev = io.rest_path(setup.checkpoint_root, t - setup.checkpoint_tinv)
        if t - setup.checkpoint_tinv > 0 and os.path.exists(prev):
            os.remove(prev)
    if ntime0 + steps > done:
        s.run(ntime0 + steps - done)
    s.sync()
    el = time.perf_counter() - t0
    os.makedirs(setup.output_root, exist_ok=True)
    out = []
    if len(m.rec_indx):
        out.append(io.write_seismo(setup.output_root, s.receivers(), m.stept, m.rec_tinv, ntime0))
    for q, sn in enumerate(snaps):
        out += io.write_snapshots(setup.output_root, q + 1, sn, setup.snap_tcnt[q], s.snapshot(q), m.stept)
    s.close()
    log(f"{m.ni}x{m.nj}x{m.nk}, {steps} steps from {ntime0} in {el:.2f}s; wrote {len(out)} files to "
        f"{setup.output_root}" + (f", {len(marks)} rest files" if marks else ""))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("conf", nargs="?", default="SeisFD3D.conf")
    ap.add_argument("--no-graphs", action="store_true")
    a = ap.parse_args(argv)
    run(a.conf, graphs=not a.no_graphs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
