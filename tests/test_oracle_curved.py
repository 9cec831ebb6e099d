"""Controls for the curved-grid parity cases (CPU only).

The GPU parity cases of tests/test_gpu_curved.py run on a strongly curved
grid away from the equator (colatitude 1 rad, top radius 20 km, h = 200 m).
These controls plant one mistake at a time in the oracle (a flipped cot
term of the sweep, 1/sin replaced by sin, a flipped cot term of the y-face
CFS-PML grouping, mod_macdrp.F90:928-941 and mod_abs_npml.F90:1183-1200)
and show that on that grid each one moves the state by far more than the
1e-5 parity gate, while on the equatorial 1 km grid of the older parity
cases it stays near or below the gate.
"""
import numpy as np
import pytest

from fwant_amd import models
from fwant_amd._lib import FIELDS
from oracle.pyoracle import OracleSolver, mutation

TOL = 1e-5


def interior(a, m):
    return a[3:3 + m.ni, 3:3 + m.nj, 3:3 + m.nk]


def state_after(m, steps, mut=0, seed=7):
    o = OracleSolver(m, 32, nthreads=4)
    rng = np.random.default_rng(seed)
    for f in range(9):
        o.set_wavefield(f, rng.standard_normal((m.nx, m.ny, m.nz)).astype(np.float32))
    with mutation(mut):
        o.run(steps)
    out = {f: interior(o.wavefield(f), m).astype(np.float64) for f in FIELDS}
    o.close()
    return out


def max_rel(a, b):
    return max(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f]) for f in FIELDS)


@pytest.fixture(scope="module")
def grids():
    curved = models.curved_test(ni=40, nj=36, nk=24, nt=8, seed=41)
    flat = models.small_test(ni=40, nj=36, nk=24, nt=8, seed=41)
    return {"curved": (curved, state_after(curved, 8)), "flat": (flat, state_after(flat, 8))}


@pytest.mark.parametrize("mut,name", [(1, "cot terms of hVx flipped"), (2, "1/sin -> sin in hVx"),
                                      (3, "cot term of the y-face PML grouping flipped"),
                                      (4, "cot term of E22 flipped")])
def test_curved_grid_resolves_geometric_terms(grids, mut, name):
    m, ref = grids["curved"]
    e = max_rel(state_after(m, 8, mut), ref)
    mf, reff = grids["flat"]
    ef = max_rel(state_after(mf, 8, mut), reff)
    print(f"{name}: curved grid {e:.2e}, equatorial 1 km grid {ef:.2e}")
    assert e > 100 * TOL, name
    assert ef < e / 100


def test_curved_grid_geometry():
    m = models.curved_test()
    xc = 1.0 / np.tan(m.x[3:3 + m.ni])
    assert 0.35 < xc.min() and xc.max() < 1.0
    assert m.z[3] > 0 and (m.z[-4] - m.z[3]) / m.z[-4] > 0.2
    assert m.abs_number[:5] == (8,) * 5
