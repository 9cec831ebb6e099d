"""GPU parity of mode C: the spherical velocity-stress scheme on a
terrain-following grid under topography (fw_build_metric3d), against the
oracle's mode C (pinned by tests/test_oracle_topo.py).

  - the 3-D metric built on the GPU in FP64 equals the oracle's (FP32 values);
  - a flat terrain-following grid (r = z(k)) runs the spherical scheme: the
    mode C kernels agree with the spherical kernels to FP32 rounding;
  - steep hills (slopes up to ~0.3, heights 1.5-3 km on a 1 km grid) with
    CFS-PML, the free surface, a surface force on a hill, a buried force and
    a moment source: receivers and every field within 1e-5 relative L2 of the
    FP32 oracle after 96 steps, for each stage-kernel tile shape;
  - two F/B cycles from a random state (ghost layers included) under hills;
  - fw_init rejects topography inside the CFS-PML shell and fluid cells.
"""
import numpy as np
import pytest

from fwant_amd import models
from fwant_amd._lib import FIELDS

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a, np.float64) - b) / (nb if nb > 0 else 1.0)


def interior(a, m):
    return a[3:3 + m.ni, 3:3 + m.nj, 3:3 + m.nk]


def check_fields(s, o, m, tol=TOL):
    for f in FIELDS:
        ef = rel(interior(s.wavefield(f), m), interior(o.wavefield(f), m))
        print(f"{f}: {ef:.2e}", flush=True)
        assert ef <= tol, f


def test_metric3d_matches_oracle(gpu_lib):
    from fwant_amd.solver import Solver
    from oracle.pyoracle import build_metric3d
    m = models.hills_test(nt=1)
    s = Solver(m, metric3d_out=True)
    o = build_metric3d(m)
    for k in ("p", "q", "zz", "rr"):
        g = s.metric3d[k]
        want = o[k].astype(np.float32)
        e = np.abs(g - want).max() / np.abs(want).max()
        print(f"{k}: max rel diff {e:.2e}", flush=True)
        assert e <= 2e-7, k
    assert np.abs(s.metric3d["p"]).max() > 0.1
    s.close()


def test_flat_terrain_matches_the_spherical_scheme(gpu_lib):
    """r = z(k): slopes 0 and the 1-D zeta_r and 1/z bit for bit, so the mode
    C kernels compute the spherical scheme plus exact zeros; the two
    instantiations differ only in where the compiler fuses multiplies and
    adds (one rounding in a few cells near the surface force), so they agree
    to FP32 rounding."""
    from fwant_amd.solver import Solver
    m = models.small_test(ni=40, nj=36, nk=28, nt=48)
    mf = models.with_topography(m, np.zeros((m.nx, m.ny)))
    a, b = Solver(m), Solver(mf, metric3d_out=True)
    assert np.all(b.metric3d["p"] == 0.0) and np.all(b.metric3d["q"] == 0.0)
    assert np.array_equal(b.metric3d["zz"], np.broadcast_to(a.metric["zeta_z"], b.metric3d["zz"].shape))
    a.run(48)
    b.run(48)
    e = rel(b.receivers(), a.receivers())
    print(f"receivers rel-L2 {e:.2e}", flush=True)
    assert e <= 1e-6
    for f in FIELDS:
        ef = rel(interior(b.wavefield(f), m), interior(a.wavefield(f), m))
        print(f"{f}: {ef:.2e}", flush=True)
        assert ef <= 1e-6, f
    a.close()
    b.close()


@pytest.mark.parametrize("tile", ["0", "1", "2", "3"])
def test_hills_parity(gpu_lib, monkeypatch, tile):
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver
    monkeypatch.setenv("FWANT_TILE", tile)
    m = models.hills_test(ni=84, nj=40, nk=28, nt=96) if tile == "0" else models.hills_test(nt=96)
    n = m.nt
    print(f"tile={tile} grid={m.ni}x{m.nj}x{m.nk} steps={n} PML {m.abs_number} "
          f"hills max {np.max(m.r3[:, :, m.nk2] - m.z[m.nk2]):.0f} m", flush=True)
    s = Solver(m, metric3d_out=True)
    s.run(n)
    g = s.receivers()
    o = OracleSolver(m, 32, metric=s.metric, metric3d=s.metric3d)
    r = o.run(n)
    e = rel(g, r)
    print(f"receivers rel-L2 {e:.2e}", flush=True)
    assert np.linalg.norm(r) > 0 and e <= TOL
    check_fields(s, o, m)
    s.close()
    o.close()


def test_hills_random_state_parity(gpu_lib):
    from fwant_amd.solver import Solver
    from oracle.pyoracle import OracleSolver
    m = models.hills_test(ni=76, nj=40, nk=28, nt=16, seed=9)
    print(f"grid={m.ni}x{m.nj}x{m.nk} PML {m.abs_number}, 16 steps from a random state", flush=True)
    s = Solver(m, max_records=32, metric3d_out=True)
    o = OracleSolver(m, 32, metric=s.metric, metric3d=s.metric3d)
    rng = np.random.default_rng(12)
    for f in range(9):
        a = rng.standard_normal((m.nx, m.ny, m.nz)).astype(np.float32)
        s.set_wavefield(f, a)
        o.set_wavefield(f, a)
    s.run(16)
    o.run(16)
    check_fields(s, o, m)
    s.close()
    o.close()


def test_init_rejects_hills_in_the_pml_and_fluids(gpu_lib):
    from fwant_amd.solver import Solver
    m = models.small_test(ni=40, nj=36, nk=28, nt=8)
    bad = models.with_topography(m, np.zeros((m.nx, m.ny)))
    r3 = np.array(bad.r3)
    r3[4, 10, :] += 50.0  # a bump inside the x1 slab
    bad.r3 = np.asfortranarray(r3)
    with pytest.raises(RuntimeError, match="flat grid"):
        Solver(bad)
    w = models.with_topography(models.small_test(ni=40, nj=36, nk=28, nt=8, water=True), np.zeros((46, 42)))
    with pytest.raises(RuntimeError, match="fluid"):
        Solver(w)
