"""Batched virtual sources, step graphs and the autotuner on the GPU.

  - every member of a batch (uneven record lengths) gives exactly the
    receiver records, snapshots and final wavefield of a solve of its own
    source alone (the same arithmetic, bit for bit);
  - runs replayed from captured step graphs equal the launch-by-launch runs
    bit for bit, replay after replay, with host work and synchronisation
    between the replays;
  - after the autotuner picks a tile shape and width, a run equals a fresh
    solver's set to that configuration, bit for bit.
"""
import numpy as np
import pytest

from fwant_amd import models
from fwant_amd._lib import FIELDS

pytestmark = pytest.mark.gpu

STATIONS = [(3 + 12, 3 + 10), (3 + 25, 3 + 20), (3 + 18, 3 + 27)]
NSTEPS = [40, 64, 24]
SNAP = dict(subs=(10, 10, 20), subc=(20, 18, 10), subt=(1, 1, 1), tinv=4, fields=("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"),
            max_records=20)


def base(ni=40):
    return models.small_test(ni=ni, nj=36, nk=30, nt=64, seed=51, moment=False)


def single(member, ni=40):
    m, b = models.virtual_sources(base(ni), [STATIONS[member]], [NSTEPS[member]], f0=0.6)
    return m


# ni = 76: one wide tile and a 12-column narrow edge tile holding the x2 slab
@pytest.mark.parametrize("ni", [40, 76])
def test_members_equal_single_solves(gpu_lib, ni):
    from fwant_amd.solver import Solver
    m, b = models.virtual_sources(base(ni), STATIONS, NSTEPS, f0=0.6)
    print(f"ni={ni}: batch of {len(STATIONS)} members, record lengths {NSTEPS}", flush=True)
    snaps = [dict(SNAP), dict(SNAP, format="bf16")]
    s = Solver(m, batch=b, snapshots=snaps, max_records=80)
    s.run(max(NSTEPS))
    for q in range(len(STATIONS)):
        s.select(q)
        rb = s.receivers()
        assert rb.shape[0] == NSTEPS[q]
        one = Solver(single(q, ni), snapshots=snaps, max_records=80)
        one.run(NSTEPS[q])
        np.testing.assert_array_equal(rb, one.receivers())
        np.testing.assert_array_equal(s.snapshot(0), one.snapshot(0))
        cb, sb = s.snapshot_packed(1)
        c1, s1 = one.snapshot_packed(1)
        np.testing.assert_array_equal(cb, c1)
        np.testing.assert_array_equal(sb, s1)
        for f in FIELDS:
            np.testing.assert_array_equal(s.wavefield(f), one.wavefield(f))
        one.close()
        print(f"member {q}: {rb.shape[0]} records identical to its own solve")
    s.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replays_equal_eager(gpu_lib, batched):
    from fwant_amd.solver import Solver

    def make():
        if batched:
            m, b = models.virtual_sources(base(), STATIONS, NSTEPS, f0=0.6)
            return Solver(m, batch=b, snapshots=[dict(SNAP, format="fp16")], max_records=80)
        return Solver(models.small_test(nt=64, seed=52), snapshots=[dict(SNAP)], max_records=80)

    print(f"graphs vs eager, batched={batched}", flush=True)
    ref = make()
    ref.run(64)
    g = make()
    g.use_graphs(True)
    members = range(len(STATIONS)) if batched else [0]
    for chunk in range(4):  # 16 steps per chunk: each step graph replays twice per chunk
        g.run(16)
        g.sync()
        for q in members:  # host work between replays
            g.select(q)
            ref.select(q)
            got, want = g.receivers(), ref.receivers()
            np.testing.assert_array_equal(got, want[: got.shape[0]])
    for q in members:
        g.select(q)
        ref.select(q)
        np.testing.assert_array_equal(g.receivers(), ref.receivers())
        if batched:
            np.testing.assert_array_equal(g.snapshot_packed(0)[0], ref.snapshot_packed(0)[0])
        else:
            np.testing.assert_array_equal(g.snapshot(0), ref.snapshot(0))
        for f in FIELDS:
            np.testing.assert_array_equal(g.wavefield(f), ref.wavefield(f))
    g.close()
    ref.close()


def test_autotune_keeps_results(gpu_lib, monkeypatch):
    # the autotuned solver equals a fresh solver set to the configuration it
    # picked, bit for bit: its trial steps leave no trace (tile shapes with
    # 1 and 2 rows per thread round differently, so the reference takes the
    # same shape)
    from fwant_amd.solver import Solver
    m, b = models.virtual_sources(base(), STATIONS, NSTEPS, f0=0.6)
    t = Solver(m, batch=b, max_records=80)
    tile, width, ms = t.autotune(4)
    print(f"autotune: tile {tile}, width {width}, {ms:.3f} ms/step", flush=True)
    monkeypatch.setenv("FWANT_TILE", str(tile))
    ref = Solver(m, batch=b, max_records=80)
    ref.set_width(width)
    ref.run(64)
    t.use_graphs(True)
    t.run(64)
    for q in range(len(STATIONS)):
        t.select(q)
        ref.select(q)
        np.testing.assert_array_equal(t.receivers(), ref.receivers())
    t.close()
    ref.close()
