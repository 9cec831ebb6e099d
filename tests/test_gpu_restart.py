"""Restart on the GPU.

  - exact (fw_checkpoint_*): a run stopped after 17 steps, saved to a
    rest_ file with its complete state and continued by a new solver for 14
    steps gives exactly the records and wavefield of an uninterrupted
    31-step run;
  - reference semantics: continuing from the 9 wavefields alone, with the
    PML memory at zero and the F/B cycle restarted at step 17, matches the
    FP32 oracle doing the same to 1e-5, and differs from the exact restart.
"""
import numpy as np
import pytest

from fwant_amd import io, models
from fwant_amd._lib import FIELDS

pytestmark = pytest.mark.gpu
N1, N2 = 17, 14


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def test_exact_and_reference_restart(gpu_lib, tmp_path):
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver
    m = models.small_test(nt=N1 + N2, seed=81, pml=6)
    print(f"restart at step {N1} of {N1 + N2}", flush=True)
    full = Solver(m)
    full.run(N1 + N2)
    want = full.receivers()
    first = Solver(m)
    first.run(N1)
    path = io.write_rest(str(tmp_path), N1, {f: first.wavefield(f) for f in FIELDS}, state=first.checkpoint(),
                         cycle_origin=0)
    fields, state, origin = io.read_rest(path)
    # exact restart
    cont = Solver(m, metric=full.metric)
    cont.restore(state)
    cont.set_cycle_origin(origin)
    cont.run(N2, ntime0=N1)
    got = cont.receivers()
    np.testing.assert_array_equal(got, want[N1:])
    for f in FIELDS:
        np.testing.assert_array_equal(cont.wavefield(f), full.wavefield(f))
    # the reference's restart: 9 wavefields, PML memory at zero, cycle restarted
    ref = Solver(m, metric=full.metric)
    for f in FIELDS:
        ref.set_wavefield(f, fields[f])
    ref.set_cycle_origin(N1)
    ref.run(N2, ntime0=N1)
    g = ref.receivers()
    o = OracleSolver(m, 32, metric=full.metric)
    for f in FIELDS:
        o.set_wavefield(f, fields[f])
    o.set_cycle_origin(N1)
    o.ntime = N1
    r = o.run(N2)
    e = rel(g, r)
    d = rel(g, want[N1:])
    print(f"reference-style restart vs FP32 oracle {e:.2e}; differs from the exact restart by {d:.2e}")
    assert e <= 1e-5
    assert d > 1e-6
    for s in (full, first, cont, ref):
        s.close()
    o.close()


def test_exact_restart_under_topography(gpu_lib):
    """The exact restart on the terrain-following grid (mode C, 3-D metric,
    hills reaching the free surface): stopped after 17 steps and continued
    by a new solver from its checkpoint, records and wavefield equal an
    uninterrupted 31-step run bit for bit."""
    from fwant_amd.solver import Solver
    m = models.hills_test(nt=N1 + N2)
    full = Solver(m)
    full.run(N1 + N2)
    want = full.receivers()
    first = Solver(m)
    first.run(N1)
    state = first.checkpoint()
    cont = Solver(m, metric=full.metric)
    cont.restore(state)
    cont.run(N2, ntime0=N1)
    np.testing.assert_array_equal(cont.receivers(), want[N1:])
    for f in FIELDS:
        np.testing.assert_array_equal(cont.wavefield(f), full.wavefield(f))
    assert np.abs(want).max() > 0
    for s in (full, first, cont):
        s.close()
