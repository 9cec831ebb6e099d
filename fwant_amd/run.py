"""Run a reference-style case: python -m fwant_amd.run SeisFD3D.conf

The flow of code/srcF/seis3d_wave_8op_p3.F90 with the solve on the GPU:
read the parameter file and the per-rank NetCDF inputs (fwant_amd.io),
run the steps the reference driver runs for nt (8*ceil((nt+1)/8), the exit
test sits at the top of each 8-step block), then write seismo_mpi000000.nc
and the sgt_/vel_ snapshot files under OUTPUT_ROOT.
"""
import argparse
import os
import sys
import time

from . import io
from .solver import Solver


def run(conf_path, graphs=True, log=print):
    setup = io.load_run(conf_path)
    m = setup.model
    steps = m.ref_steps()
    snaps = [dict(s, max_records=steps // s["tinv"] + 1) for s in setup.snapshots]
    t0 = time.perf_counter()
    s = Solver(m, metric=setup.metric, snapshots=snaps, max_records=steps // max(m.rec_tinv, 1) + 1)
    s.use_graphs(graphs)
    s.run(steps)
    s.sync()
    el = time.perf_counter() - t0
    os.makedirs(setup.output_root, exist_ok=True)
    out = []
    if len(m.rec_indx):
        out.append(io.write_seismo(setup.output_root, s.receivers(), m.stept, m.rec_tinv))
    for q, sn in enumerate(snaps):
        out += io.write_snapshots(setup.output_root, q + 1, sn, setup.snap_tcnt[q], s.snapshot(q), m.stept)
    s.close()
    log(f"{m.ni}x{m.nj}x{m.nk}, {steps} steps in {el:.2f}s; wrote {len(out)} files to {setup.output_root}")
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
