"""Reduced-precision SGT storage on the GPU (fw_add_snapshot_packed,
fwk_sgt_decode):

  - the packed codes and block scales the solver writes equal the oracle's
    packing of the FP32 snapshot of the same run, bit for bit;
  - fwk_sgt_decode equals the oracle's restore bit for bit;
  - sensitivity kernels computed from BF16 / FP16 SGT stay within 1%
    relative L2 of the FP64 oracle pipeline on the FP32 SGT (the FP32 GPU
    pipeline's own error is printed beside them).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS6 = ("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")
SNAP = dict(subs=(8, 8, 10), subc=(20, 18, 8), subt=(1, 1, 1), tinv=2, fields=FIELDS6, max_records=40)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run_three(seed, fz):
    from fwant_amd import models
    from fwant_amd.solver import Solver
    m = models.small_test(nt=80, seed=seed, moment=False)
    m.force.comps[2] = fz
    snaps = [dict(SNAP), dict(SNAP, format="bf16"), dict(SNAP, format="fp16")]
    s = Solver(m, snapshots=snaps)
    s.run(80)
    out = {"fp32": s.snapshot(0), "bf16": s.snapshot_packed(1), "fp16": s.snapshot_packed(2),
           "dev": {k: s.snapshot_packed_device(i) for i, k in ((1, "bf16"), (2, "fp16"))}}
    return m, s, out


def test_packed_snapshot_bit_exact(gpu_lib):
    from fwant_amd import kernel
    from fwant_amd._lib import DeviceArray
    from oracle.pyoracle import sgt_decode, sgt_encode
    print("packed snapshots on small_test, 80 steps", flush=True)
    m, s, out = run_three(41, 1e15)
    x = out["fp32"]
    assert x.shape == (40, 6, int(np.prod(SNAP["subc"])))
    for fmt in ("bf16", "fp16"):
        codes, scales = out[fmt]
        oc, osc = sgt_encode(fmt, x, SNAP["subc"])
        np.testing.assert_array_equal(scales, osc)
        np.testing.assert_array_equal(codes, oc)
        pc, ps, nrec, npts, nblk = out["dev"][fmt]
        y = DeviceArray((nrec, 6, npts), np.float32)
        kernel.sgt_decode(fmt, pc, ps, SNAP["subc"], nrec * 6, y)
        got = y.to_host()
        np.testing.assert_array_equal(got, sgt_decode(fmt, codes, scales, SNAP["subc"]))
        print(f"{fmt}: restore rel-L2 vs FP32 snapshot {rel(got, x):.2e}")
    s.close()


def oracle_pipeline(prec, R, S, lam, mu, dt, V, U, b, a, nt1, nt2, M0):
    from oracle.pyoracle import Lib
    lib = Lib(prec)
    T = lib.dtype
    knt, npt = R.shape[1], R.shape[2]
    P = lambda v: v.ctypes.data_as(C.POINTER(lib.R))  # noqa: E731
    Rl = [np.ascontiguousarray(R[c], T) for c in range(6)]
    Sl = [np.ascontiguousarray(S[c], T) for c in range(6)]
    la, mu_ = np.ascontiguousarray(lam, T), np.ascontiguousarray(mu, T)
    lib.stress2strain(npt, knt, *[P(r) for r in Rl], P(la), P(mu_))
    lib.stress2strain(npt, knt, *[P(r) for r in Sl], P(la), P(mu_))
    A = np.zeros((knt, npt), T)
    B = np.zeros((knt, npt), T)
    lib.scatter1(npt, knt, dt, (C.POINTER(lib.R) * 6)(*[P(r) for r in Rl]), (C.POINTER(lib.R) * 6)(*[P(q) for q in Sl]),
                 P(la), P(mu_), P(A), P(B))
    D = C.POINTER(C.c_double)
    bb, aa = np.ascontiguousarray(b, np.float64), np.ascontiguousarray(a, np.float64)
    filt = lambda w, n: (lib.series_filter(2, n, knt, len(b), bb.ctypes.data_as(D), aa.ctypes.data_as(D), P(w[0]),  # noqa: E731
                                           P(w[1])), w[1])[1]
    Af = filt((A, np.zeros_like(A)), npt)
    Bf = filt((B, np.zeros_like(B)), npt)
    Vf = filt((np.ascontiguousarray(V, T), np.zeros(knt, T)), 1)
    Uf = filt((np.ascontiguousarray(U, T), np.zeros(knt, T)), 1)
    ref = [np.zeros(npt, T) for _ in range(4)]
    lib.scatter2kernel(npt, knt, nt1, nt2, dt, M0, P(Vf), P(Uf), P(Af), P(Bf), *[P(r) for r in ref])
    return ref


def test_kernel_error_from_packed_sgt(gpu_lib):
    from scipy import signal
    from fwant_amd import kernel
    from fwant_amd._lib import DeviceArray
    print("kernels from FP32 / BF16 / FP16 SGT", flush=True)
    (m, s1, o1), (_, s2, o2) = run_three(31, 1e15), run_three(32, -2e15)
    knt, _, npt = o1["fp32"].shape
    sl = tuple(slice(st - 1, st - 1 + c * t, t) for st, c, t in zip(SNAP["subs"], SNAP["subc"], SNAP["subt"]))
    lam = m.lam[sl].ravel(order="F").astype(np.float32)
    mu = m.mu[sl].ravel(order="F").astype(np.float32)
    dt = 2 * m.stept
    t = dt * np.arange(1, knt + 1)
    V = np.sin(2 * np.pi * 0.3 * t).astype(np.float32)
    U = np.cos(2 * np.pi * 0.3 * t).astype(np.float32)
    b, a = signal.butter(2, 0.3)
    nt1, nt2, M0 = 5, 36, 1e15
    T = lambda x: np.ascontiguousarray(x.transpose(1, 0, 2))  # noqa: E731  (knt, 6, npt) -> (6, knt, npt)
    ref = oracle_pipeline(64, T(o1["fp32"]), T(o2["fp32"]), lam, mu, dt, V, U, b, a, nt1, nt2, M0)
    dl, dm = DeviceArray.from_host(lam), DeviceArray.from_host(mu)

    def gpu_kernels(R, S):
        dR, dS = DeviceArray.from_host(T(R)), DeviceArray.from_host(T(S))
        return [o.to_host() for o in kernel.pair_kernel(None, None, dR, dS, dl, dm, dt, V, U, nt1, nt2, M0, 2,
                                                        filt=("P2", b, a))]

    def decoded(o, fmt):
        pc, ps, nrec, n, nblk = o["dev"][fmt]
        y = DeviceArray((nrec, 6, n), np.float32)
        kernel.sgt_decode(fmt, pc, ps, SNAP["subc"], nrec * 6, y)
        return y.to_host()

    errs = {}
    for fmt in ("fp32", "bf16", "fp16"):
        R = o1["fp32"] if fmt == "fp32" else decoded(o1, fmt)
        S = o2["fp32"] if fmt == "fp32" else decoded(o2, fmt)
        got = gpu_kernels(R, S)
        errs[fmt] = [rel(g, r) for g, r in zip(got, ref)]
        print(f"{fmt} SGT: kernel rel-L2 vs FP64 oracle (Kap Kaq Kbp Kbq) " + " ".join(f"{e:.2e}" for e in errs[fmt]))
    s1.close()
    s2.close()
    assert max(errs["fp32"]) < 1e-5
    assert max(errs["bf16"]) < 1e-2
    assert max(errs["fp16"]) < 1e-2
