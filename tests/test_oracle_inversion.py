"""Pins of the oracle's inversion steps and the host-side system builder:
  - LSQR (inv.LSQR/mod_LSQR.f90:142-292) converges to the least-squares
    solution numpy.linalg.lstsq gives, for a damped and smoothed system;
  - block assembly (SI_ker2assm.F90:300-320) equals a dense sum;
  - the compact G file is the Fortran unformatted record layout and keeps
    |scale*B| > ker0 (export_G_compact, :545-572).
"""
import numpy as np

from fwant_amd import inversion
from oracle.pyoracle import Lib


def random_system(seed, nd=40, nb=25, lamb=0.3, eta=0.5):
    rng = np.random.default_rng(seed)
    s = inversion.System(ncol=nb)
    dense = []
    for r in range(nd):
        idx = np.sort(rng.choice(np.arange(1, nb + 1), size=8, replace=False))
        v = rng.standard_normal(8)
        s.add_data([(idx, v, nb)], rhs=rng.standard_normal(), weight=1.5)
    s.add_damping(lamb)
    s.add_smoothing([(np.array([i, i + 1]), np.array([1.0, -1.0])) for i in range(1, nb)], eta, num_blk=nb)
    ptr, col, val, rhs = s.arrays()
    A = np.zeros((len(rhs), nb))
    for r in range(len(rhs)):
        for q in range(ptr[r], ptr[r + 1]):
            A[r, col[q]] += val[q]
    return s, A, rhs


def test_lsqr_reaches_least_squares_solution():
    lib = Lib(64)
    s, A, b = random_system(1)
    ptr, col, val, rhs = s.arrays()
    x = np.zeros(s.ncol)
    err = np.zeros(s.ncol)
    res = np.zeros(1)
    it = lib.lsqr(len(rhs), s.ncol, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, rhs.ctypes.data, 1e-12, 1e-12,
                  30000, x.ctypes.data, err.ctypes.data, res.ctypes.data)
    want = np.linalg.lstsq(A, b, rcond=None)[0]
    assert 0 < it < 30000
    np.testing.assert_allclose(x, want, rtol=1e-8, atol=1e-10)
    assert abs(res[0] - np.linalg.norm(A @ want - b)) < 1e-8 * np.linalg.norm(b)
    assert np.all(err >= 0)


def test_system_rows():
    s = inversion.System(ncol=6)
    s.add_data([(np.array([1, 3]), np.array([0.5, 1e-9]), 2), (np.array([2]), np.array([2.0]), 2)], rhs=3.0,
               weight=2.0, ker_thres=1e-6, stations=[-1], events=[2], num_mval=4, num_sta=1, weig_sta=0.1,
               weig_evt=0.2)
    ptr, col, val, rhs = s.arrays()
    assert list(col) == [0, 3, 4, 6] and rhs[0] == 6.0
    np.testing.assert_allclose(val, [1.0, 4.0, -0.2, 0.4])


def test_assembly_and_compact_file(tmp_path):
    rng = np.random.default_rng(3)
    npt, nblk = 500, 30
    K = rng.standard_normal(npt).astype(np.float32)
    owner = rng.integers(0, nblk, npt)
    order = np.argsort(owner, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(np.bincount(owner, minlength=nblk))]).astype(np.int64)
    vol = rng.uniform(0.5, 2, npt).astype(np.float32)
    lib = Lib(64)
    B = np.zeros(nblk)
    pix = order.astype(np.int64)
    lib.assemble(nblk, ptr.ctypes.data, pix.ctypes.data, vol.ctypes.data, K.ctypes.data, B.ctypes.data)
    dense = np.zeros(nblk)
    for q, p in enumerate(pix):
        dense[owner[p]] += float(K[p]) * float(vol[q])
    np.testing.assert_allclose(B, dense, rtol=1e-12)
    idx, v = inversion.compact(B, 2.0, 1.0)
    assert np.all(np.abs(v) > 1.0) and np.all(np.abs(np.float32(2.0) * B.astype(np.float32)[idx - 1]) == np.abs(v))
    p = tmp_path / "G.Kap.unformat"
    inversion.write_G_compact(p, nblk, 1.0, idx, v)
    raw = p.read_bytes()
    assert raw[:4] == np.int32(8).tobytes() and raw[12:16] == np.int32(8).tobytes()
    nb, k0, i2, v2 = inversion.read_G_compact(p)
    assert nb == nblk and k0 == 1.0
    np.testing.assert_array_equal(i2, idx)
    np.testing.assert_array_equal(v2, v)
