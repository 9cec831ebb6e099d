"""Learned SGT compression at the configuration it is specified for: the 3-D
convolutional autoencoder (6 x 16^3 blocks -> 8 x 4^3 latent, 4 layers per
side, FP32 weights, BF16 activations) trained on the GPU on stress SGT of
the batched configuration's curvilinear 256 x 256 x 128 model (4 virtual
sources at seeded stations, 200 steps, SGT every 10 steps on the 2x
decimated 128 x 128 x 64 grid), three sources for training and the fourth
held out.  Measured and printed:

  - training loss, and the held-out source's block errors against the
    method without a network (4^3 block averages), with the fraction of
    blocks within 1e-3 .. 0.3;
  - a threshold sweep of the per-block FP32 fallback: live blocks kept as
    latents, stored bytes (exactly zero blocks cost only their scales) and
    the traveltime kernels of 3 station pairs against the FP64 oracle
    pipeline on the FP32 SGT.

Asserted: training is finite and lowers the loss; the accepted fraction
grows with the threshold; at the specified 1e-3 threshold every record
stays within 1e-3 of the FP32 SGT and the kernels within 1% of the FP64
oracle.  On this SGT (1.3-3 samples per wavelength at 2x decimation:
Vs 3.2 km/s, Ricker 1 Hz) the network's held-out median error is 1.0
(4^3 averages 0.92) and no live block is accepted below 0.3: the store's
only saving is the zero blocks ahead of the waves (0.81 of FP32).
"""
import numpy as np
import pytest

from fwant_amd import models

pytestmark = pytest.mark.gpu


def test_config5_learned_compression(gpu_lib):
    from scipy import signal
    from fwant_amd import ae, kernel, pipeline
    from fwant_amd._lib import DeviceArray
    from fwant_amd.solver import Solver
    from test_gpu_configs import N3, model3
    from test_gpu_sgt import oracle_pipeline, rel
    rng = np.random.default_rng(15)
    c = 3 + N3[0] // 2
    st = [(int(c + u), int(c + v)) for u, v in rng.choice(np.arange(-12, 13), size=(4, 2))]
    assert len(set(st)) == 4
    nsteps = [200] * 4
    base = model3(200)
    m, b = models.virtual_sources(base, st, nsteps, f0=1.0)
    m.rec_indx = np.zeros((0, 3), np.int32)
    snap = pipeline.sgt_snapshot(m, 10, 2, "fp32", max_records=20)
    subc = snap["subc"]
    print(f"4 sources on {N3}, 200 steps, SGT every 10 steps on {subc}", flush=True)
    s = Solver(m, batch=b, snapshots=[snap], max_records=1)
    s.run(200)
    s.sync()
    rec = {}
    for q in range(4):
        s.select(q)
        rec[q] = s.snapshot(0)[2:]  # (nrec, 6, npts); the first records are nearly empty
        assert np.isfinite(rec[q]).all()
    s.close()
    # training set: the live blocks of sources 0-2
    blk, sc = {}, {}
    for q in range(4):
        bq, sq = ae.blocks_of(rec[q], subc)
        blk[q], sc[q] = bq.to_host().reshape(bq.shape[0], -1), sq
        bq.free()
    live = {q: np.nonzero(sc[q].max(axis=1) > 0)[0] for q in range(4)}
    train = np.concatenate([blk[q][live[q]] for q in range(3)])
    print(f"{train.shape[0]} training blocks (sources 0-2), {len(live[3])} live held-out blocks", flush=True)
    net = ae.AE(seed=5)
    tb = DeviceArray.from_host(train.reshape(-1, 6, 16, 16, 16))
    hist = ae.train(net, tb, train.shape[0], iters=1500, batch=16, lr=2e-3, log=print)
    tb.free()
    print(f"training loss: first 10 iterations {np.mean(hist[:10]):.3e}, last 100 {np.mean(hist[-100:]):.3e}",
          flush=True)
    assert np.isfinite(hist).all() and np.mean(hist[-100:]) < np.mean(hist[:10])
    hb = blk[3][live[3]]
    hd = DeviceArray.from_host(hb.reshape(-1, 6, 16, 16, 16))
    dec3, _ = ae.reconstruct(net, hd, hb.shape[0])
    hd.free()
    e_ae = ae.block_errors(hb, dec3, sc[3][live[3]])
    e_lin = ae.block_errors(hb, ae.linear_baseline(hb), sc[3][live[3]])
    print(f"held-out source: block error median autoencoder {np.median(e_ae):.3f}, 4^3 averages "
          f"{np.median(e_lin):.3f}; best tenth {np.quantile(e_ae, 0.1):.3f} / {np.quantile(e_lin, 0.1):.3f}",
          flush=True)
    for thr in (1e-3, 1e-2, 1e-1, 0.3):
        print(f"  held-out live blocks within {thr:g}: autoencoder {np.mean(e_ae <= thr):.3f}, "
              f"4^3 averages {np.mean(e_lin <= thr):.3f}", flush=True)

    # decoded blocks of every source once; the sweep picks per threshold
    comp = {}
    for q in range(4):
        stats, dec, true, scales = ae.compress(net, rec[q], subc, threshold=1e-3)
        comp[q] = (stats.errors, dec, true, scales)
    pairs = [(0, 1), (2, 3), (1, 3)]
    sl = tuple(slice(a - 1, a - 1 + cc * t, t) for a, cc, t in zip((4, 4, 4), subc, (2, 2, 2)))
    lam = np.ascontiguousarray(base.lam[sl].ravel(order="F"), np.float32)
    mu = np.ascontiguousarray(base.mu[sl].ravel(order="F"), np.float32)
    dl, dm = DeviceArray.from_host(lam), DeviceArray.from_host(mu)
    npts = int(np.prod(subc))
    knt = rec[0].shape[0]
    dt = 10 * base.stept
    t = dt * np.arange(1, knt + 1)
    b_, a_ = signal.butter(2, 0.4)
    energy = {q: np.einsum("rcp,rcp->p", rec[q].astype(np.float64), rec[q].astype(np.float64)) for q in rec}
    prng = np.random.default_rng(16)
    pts, ref = {}, {}
    for n, (qa, qb) in enumerate(pairs):
        both = np.nonzero((energy[qa] > 1e-6 * energy[qa].max()) & (energy[qb] > 1e-6 * energy[qb].max()))[0]
        pts[n] = np.sort(prng.choice(both, size=min(2048, len(both)), replace=False))
        T = lambda x: np.ascontiguousarray(x[:, :, pts[n]].transpose(1, 0, 2))  # noqa: E731
        V = np.sin(2 * np.pi * 0.3 * t + n).astype(np.float32)
        U = np.cos(2 * np.pi * 0.3 * t + n).astype(np.float32)
        ref[n] = (V, U, oracle_pipeline(64, T(rec[qa]), T(rec[qb]), lam[pts[n]], mu[pts[n]], dt, V, U, b_, a_, 2,
                                        knt - 1, 1e16))
    prev = -1.0
    for thr in (1e-3, 1e-2, 1e-1, 0.3, 1.0):
        recs, acc, stored, full = {}, [], 0, 0
        for q in range(4):
            err, dec, true, scales = comp[q]
            zero = scales.max(axis=1) == 0
            a = (err <= thr) | zero
            mixed = np.where(a[:, None], dec, true)
            recs[q] = ae.records_from_blocks(mixed, scales, subc, rec[q].shape[0])
            acc.append(a[~zero].mean() if (~zero).any() else 1.0)
            stored += int(a.sum()) * (net.latent * 2 + 6 * 4) + int((~a).sum()) * 6 * ae.BLOCK * 4
            full += a.size * 6 * ae.BLOCK * 4
        worst = 0.0
        for n, (qa, qb) in enumerate(pairs):
            V, U, r = ref[n]
            dA = DeviceArray.from_host(np.ascontiguousarray(recs[qa], np.float32))
            dB = DeviceArray.from_host(np.ascontiguousarray(recs[qb], np.float32))
            RA = kernel.snapshot_series(dA.ptr, knt, 6, npts)
            SB = kernel.snapshot_series(dB.ptr, knt, 6, npts)
            got = [o.to_host()[pts[n]] for o in kernel.pair_kernel(None, None, RA, SB, dl, dm, dt, V, U, 2, knt - 1,
                                                                      1e16, 2, filt=("P2", b_, a_))]
            worst = max(worst, max(rel(g, rr) for g, rr in zip(got, r)))
            dA.free()
            dB.free()
        frac = float(np.mean(acc))
        print(f"threshold {thr:g}: live blocks as latents {frac:.3f}, stored {stored / full:.3f} of FP32, "
              f"kernel rel-L2 vs FP64 oracle {worst:.2e}", flush=True)
        assert frac >= prev
        prev = frac
        if thr == 1e-3:
            for q in range(4):
                e = [rel(recs[q][r, cc], rec[q][r, cc]) for r in range(knt) for cc in range(6) if np.any(rec[q][r, cc])]
                assert max(e) <= 1e-3
            assert worst < 1e-2


def test_blocks_of_tiny_components_are_zero(gpu_lib):
    """A block component whose largest magnitude is below 2^-100 (the
    numerical precursors ahead of a wavefront reach subnormal values) is a
    zero block with scale 0, not NaN (1 / max would overflow)."""
    from fwant_amd import ae
    subc = (32, 16, 16)
    rec = np.zeros((1, 6, int(np.prod(subc))), np.float32)
    v = rec.reshape(1, 6, 16, 16, 32)
    v[0, 0, :, :, :16] = 1e-40
    v[0, 1, :, :, :16] = 3e-30
    v[0, 1, 0, 0, 0] = 0.0
    b, sc = ae.blocks_of(rec, subc)
    h = b.to_host().reshape(2, 6, -1)
    assert np.isfinite(h).all()
    assert sc[0, 0] == 0.0 and np.all(h[0, 0] == 0.0)
    assert sc[0, 1] == np.float32(3e-30) and np.abs(h[0, 1]).max() == 1.0
