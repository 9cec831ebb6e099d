"""GPU block assembly and LSQR against the oracle: assembly to 1e-12
(FP64 sums in another order), LSQR with the reference's tolerances
(1e-7) to the same solution within 1e-6 relative (the iterations differ
only by FP64 summation order), and with tight tolerances to the
least-squares solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_assembly(gpu_lib):
    from fwant_amd import inversion
    from oracle.pyoracle import Lib
    rng = np.random.default_rng(21)
    npt, nblk = 200000, 1500
    K = rng.standard_normal(npt).astype(np.float32)
    owner = rng.integers(0, nblk, npt)
    pix = np.argsort(owner, kind="stable").astype(np.int64)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(owner, minlength=nblk))]).astype(np.int64)
    vol = rng.uniform(0.5, 2, npt).astype(np.float32)
    got = inversion.assemble(K, inversion.Blocks(ptr, pix, vol))
    want = np.zeros(nblk)
    Lib(64).assemble(nblk, ptr.ctypes.data, pix.ctypes.data, vol.ctypes.data, K.ctypes.data, want.ctypes.data)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)


def test_lsqr(gpu_lib):
    from fwant_amd import inversion
    from oracle.pyoracle import Lib
    import importlib
    t = importlib.import_module("test_oracle_inversion")
    s, A, b = t.random_system(7, nd=300, nb=120)
    ptr, col, val, rhs = s.arrays()
    lib = Lib(64)
    for tol in (1e-7, 1e-13):
        x, err, it, res = inversion.lsqr(s, atol=tol, btol=tol)
        xo, eo, ro = np.zeros(s.ncol), np.zeros(s.ncol), np.zeros(1)
        ito = lib.lsqr(len(rhs), s.ncol, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, rhs.ctypes.data, tol, tol,
                       30000, xo.ctypes.data, eo.ctypes.data, ro.ctypes.data)
        e = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        print(f"tol {tol}: GPU {it} iterations, oracle {ito}; solution rel diff {e:.2e}")
        assert abs(it - ito) <= 2 and e < 1e-6
        np.testing.assert_allclose(err, eo, rtol=1e-4)
    want = np.linalg.lstsq(A, b, rcond=None)[0]
    np.testing.assert_allclose(x, want, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("maxit", [1, 5, 16, 17, 40])
def test_lsqr_iteration_limit(gpu_lib, maxit):
    """The device-side stop at itct == maxit (the host reads the stop flag
    every 16 iterations): the GPU runs exactly as many iterations as the
    oracle, for limits below, at and across that polling period, and the
    iterates agree to FP64 summation order."""
    from fwant_amd import inversion
    from oracle.pyoracle import Lib
    import importlib
    t = importlib.import_module("test_oracle_inversion")
    s, A, b = t.random_system(7, nd=300, nb=120)
    ptr, col, val, rhs = s.arrays()
    x, err, it, res = inversion.lsqr(s, atol=1e-15, btol=1e-15, maxit=maxit)
    xo, eo, ro = np.zeros(s.ncol), np.zeros(s.ncol), np.zeros(1)
    ito = Lib(64).lsqr(len(rhs), s.ncol, ptr.ctypes.data, col.ctypes.data, val.ctypes.data, rhs.ctypes.data, 1e-15,
                       1e-15, maxit, xo.ctypes.data, eo.ctypes.data, ro.ctypes.data)
    print(f"maxit {maxit}: GPU {it}, oracle {ito} iterations; rel diff {np.linalg.norm(x - xo) / np.linalg.norm(xo):.2e}")
    assert it == ito == maxit
    # FP64 sums in another order: the iterates drift apart slowly (1.1e-12
    # after 40 iterations)
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_lsqr_zero_rhs(gpu_lib):
    """b = 0: beta = alpha = 0, so rho = 0 ends the solve before the first
    iteration with x = 0 (the oracle does the same)."""
    from fwant_amd import inversion
    import importlib
    t = importlib.import_module("test_oracle_inversion")
    s, A, b = t.random_system(3, nd=60, nb=30)
    s.rhs = [0.0] * len(s.rhs)
    x, err, it, res = inversion.lsqr(s)
    assert it == 0 and np.all(x == 0.0) and res == 0.0
