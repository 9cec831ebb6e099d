"""Write the full-size golden fixture of the two-layer 100x100x50 workload:
receiver records of the FP32 and FP64 oracle over the 1008 steps the
reference driver runs for nt = 1000.  Run: python tests/golden/make_two_layer.py
(about 10 minutes on 8 cores).  Output: tests/golden/two_layer_100x100x50.npz"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fwant_amd import models  # noqa: E402
from oracle.pyoracle import OracleSolver, build_metric  # noqa: E402


def main():
    m = models.two_layer_100x100x50()
    steps = m.ref_steps()
    metric = build_metric(m)
    met32 = {k: v.astype(np.float32) for k, v in metric.items()}
    out = {}
    for prec in (32, 64):
        t = time.time()
        o = OracleSolver(m, prec, metric=met32 if prec == 32 else metric)
        rec = o.run(steps)
        out[f"rec{prec}"] = rec.astype(np.float32)
        print(f"FP{prec}: {steps} steps in {time.time() - t:.0f}s", flush=True)
        o.close()
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "two_layer_100x100x50.npz"),
                        steps=steps, rec_indx=m.rec_indx, **out)


if __name__ == "__main__":
    main()
