"""Parity at the benchmark's full size: 400x400x200 cells (gradient model,
12-point CFS-PML on 5 sides, free surface, Ricker surface force), on the
spherical grid (mode S) and on the terrain-following grid under the 20
seeded hills of the bench workload (mode C).  The nine fields start from the
same random state (ghost layers included) on the GPU and in the FP32 oracle,
so every cell, CFS-PML slab, edge tile and work item is active from the
first step; after one full 8-step F/B cycle every field of the interior is
within 1e-5 relative L2 of the oracle's.
"""
import numpy as np
import pytest

from fwant_amd import models
from fwant_amd._lib import FIELDS

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("workload", ["gradient", "hills"])
def test_full_size_random_state(gpu_lib, workload):
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver
    m = models.gradient_400x400x200() if workload == "gradient" else models.hills_400x400x200()
    print(f"{workload}: grid {m.ni}x{m.nj}x{m.nk}, PML {m.abs_number}, 8 steps from a random state", flush=True)
    if workload == "hills":
        # the GPU's FP32 metric for both (built in FP64 on each side)
        s = Solver(m, max_records=16, metric3d_out=True)
        o = OracleSolver(m, 32, metric=s.metric, nthreads=16, metric3d=s.metric3d)
    else:
        s = Solver(m, max_records=16)
        o = OracleSolver(m, 32, metric=s.metric, nthreads=16)
    rng = np.random.default_rng(2024)
    sl = (slice(3, 3 + m.ni), slice(3, 3 + m.nj), slice(3, 3 + m.nk))
    for f in range(9):
        # random values everywhere, ghost layers included: no stage writes
        # the x / y ghosts, so every stage must see the same ghost values
        a = rng.standard_normal((m.nx, m.ny, m.nz)).astype(np.float32)
        s.set_wavefield(f, a)
        o.set_wavefield(f, a)
    del a
    s.run(8)
    o.run(8)
    worst = 0.0
    for f in FIELDS:
        e = rel(s.wavefield(f)[sl], o.wavefield(f)[sl])
        print(f"{f}: {e:.2e}", flush=True)
        worst = max(worst, e)
    s.close()
    o.close()
    assert worst <= TOL
