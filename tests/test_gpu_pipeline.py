"""End-to-end station-pair kernels: the batched, BF16/FP16-stored pipeline
against the plain one (one FP32 solve per station) on the same stations,
record lengths and pairs; kernels agree to 1% relative L2 (the FP32 plain
pipeline itself is pinned to the FP64 oracle by test_gpu_sgt.py)."""
import numpy as np
import pytest

from fwant_amd import models

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_batched_compressed_matches_plain(gpu_lib, fmt):
    from scipy import signal
    from fwant_amd import pipeline
    m = models.small_test(ni=40, nj=36, nk=28, nt=120, seed=61, moment=False)
    stations = [(3 + 10, 3 + 9), (3 + 28, 3 + 12), (3 + 20, 3 + 26), (3 + 14, 3 + 22)]
    nsteps = [120, 100, 80, 110]
    pairs = [(0, 1), (0, 2), (1, 3), (2, 3)]
    knt = max(nsteps) // 4
    t = np.arange(1, knt + 1) * m.stept * 4
    V = np.sin(2 * np.pi * 0.4 * t).astype(np.float32)
    U = np.cos(2 * np.pi * 0.4 * t).astype(np.float32)
    b, a = signal.butter(2, 0.4)
    kw = dict(tinv=4, step=2, V=V, U=U, win=(3, 18), M0=1e16, filt=("P2", b, a))
    print(f"pipeline plain vs batched {fmt}: 4 stations, 4 pairs", flush=True)
    kp, tp = pipeline.plain(m, stations, nsteps, pairs, **kw)
    kb, tb = pipeline.batched(m, stations, nsteps, pairs, fmt=fmt, **kw)
    worst = 0.0
    for q, (x, y) in enumerate(zip(kp, kb)):
        for name, u, v in zip(("Kap", "Kaq", "Kbp", "Kbq"), x, y):
            e = np.linalg.norm(v.astype(np.float64) - u) / np.linalg.norm(u)
            worst = max(worst, e)
    print(f"{fmt}: worst kernel rel-L2 vs plain FP32 {worst:.2e}; times plain {tp}, batched {tb}")
    assert worst < 1e-2
