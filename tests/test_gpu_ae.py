"""The SGT block autoencoder on the GPU:
  - convolution forward / backward-data / backward-weight kernels (implicit
    GEMMs on the matrix cores) equal the numpy oracle (FP64) to FP32
    accuracy, for both layer kinds used and the edge cases of their tiling;
  - the whole network's loss and gradients (BF16-rounded activations) equal
    the oracle's;
  - trained on SGT blocks the solver produces, the network reconstructs
    held-out blocks better than the method without it (4^3 block averages,
    fewer numbers than the 512 latents) and, with the per-block FP32
    fallback at the threshold, the station-pair kernels stay within 1% of
    the FP32 ones.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


# (K, S, P, Di, N, Ci, Co): the network's two layer kinds at small and full
# size, more output channels than one 16-row matrix tile, an odd volume with
# stride 2 and no padding; the three launch shapes (waves splitting the
# reduction of few columns, 16-column and 64-column wave tiles); reductions
# longer than one LDS table pass of 1024 entries (forward 1280, backward
# data 2560)
@pytest.mark.parametrize("K,S,P,Di,N,Ci,Co", [(3, 1, 1, 8, 3, 5, 7), (4, 2, 1, 8, 3, 5, 7), (4, 2, 1, 16, 3, 16, 16),
                                             (3, 1, 1, 6, 2, 6, 20), (3, 2, 0, 7, 2, 4, 3), (3, 1, 1, 16, 16, 16, 6),
                                             (3, 1, 1, 16, 64, 6, 16), (4, 2, 1, 8, 2, 20, 5), (4, 1, 1, 6, 2, 3, 40)])
def test_conv_primitives(gpu_lib, K, S, P, Di, N, Ci, Co):
    from fwant_amd._lib import DeviceArray, lib
    from oracle import ae_ref
    rng = np.random.default_rng(K + Di + N)
    print(f"conv K={K} S={S} P={P} Di={Di} N={N} Ci={Ci} Co={Co}", flush=True)
    Do = (Di + 2 * P - K) // S + 1
    x = rng.standard_normal((N, Ci, Di, Di, Di)).astype(np.float32)
    w = rng.standard_normal((Co, Ci, K ** 3)).astype(np.float32)
    b = rng.standard_normal(Co).astype(np.float32)
    dy = rng.standard_normal((N, Co, Do, Do, Do)).astype(np.float32)
    L = lib()
    d = DeviceArray.from_host
    y, dx = DeviceArray((N, Co, Do, Do, Do)), DeviceArray((N, Ci, Di, Di, Di))
    dwb, part = DeviceArray((Co * Ci * K ** 3 + Co,)), DeviceArray((16 * (Co * Ci * K ** 3 + Co),))
    P_ = lambda a: C.c_void_p(a.ptr)  # noqa: E731
    xd, wd, bd, dyd = d(x), d(w), d(b), d(dy)
    assert L.fwae_conv_fwd(N, Ci, Co, Di, Do, K, S, P, P_(xd), P_(wd), P_(bd), P_(y), None) == 0
    assert L.fwae_conv_bwd_data(N, Ci, Co, Di, Do, K, S, P, P_(dyd), P_(wd), None, P_(dx), None) == 0
    assert L.fwae_conv_bwd_weight(N, Ci, Co, Di, Do, K, S, P, P_(xd), P_(dyd), 16, P_(part), P_(dwb), None) == 0
    W = w.reshape(Co, Ci, K, K, K).astype(np.float64)
    assert rel(y.to_host(), ae_ref.conv_fwd(x.astype(np.float64), W, b.astype(np.float64), S, P)) < 1e-6
    assert rel(dx.to_host(), ae_ref.conv_bwd_data(dy.astype(np.float64), W, S, P, Di)) < 1e-6
    # the fused activation: the plain convolution then fwae_act_fwd, bit for bit
    for slope in (0.1, 1.0):
        ya, yp = DeviceArray((N, Co, Do, Do, Do)), DeviceArray((N, Co, Do, Do, Do))
        assert L.fwae_conv_fwd_act(N, Ci, Co, Di, Do, K, S, P, P_(xd), P_(wd), P_(bd), slope, P_(ya), None) == 0
        assert L.fwae_conv_fwd(N, Ci, Co, Di, Do, K, S, P, P_(xd), P_(wd), P_(bd), P_(yp), None) == 0
        assert L.fwae_act_fwd(yp.numel(), slope, P_(yp), None) == 0
        np.testing.assert_array_equal(ya.to_host(), yp.to_host())
        dxa, dxp = DeviceArray((N, Ci, Di, Di, Di)), DeviceArray((N, Ci, Di, Di, Di))
        bi = d(rng.standard_normal(Ci).astype(np.float32))
        assert L.fwae_conv_bwd_data_act(N, Ci, Co, Di, Do, K, S, P, P_(dyd), P_(wd), P_(bi), slope, P_(dxa), None) == 0
        assert L.fwae_conv_bwd_data(N, Ci, Co, Di, Do, K, S, P, P_(dyd), P_(wd), P_(bi), P_(dxp), None) == 0
        assert L.fwae_act_fwd(dxp.numel(), slope, P_(dxp), None) == 0
        np.testing.assert_array_equal(dxa.to_host(), dxp.to_host())
    gw, gb = ae_ref.conv_bwd_weight(x.astype(np.float64), dy.astype(np.float64), S, P, K)
    h = dwb.to_host()
    # each of the 16 partial sums adds N Do^3 / 16 products in FP32: the
    # rounding error of such a sum grows like the square root of its length
    terms = N * Do ** 3 / 16
    tol = 1e-6 * max(1.0, np.sqrt(terms / 1024))
    ew, eb = rel(h[:-Co], gw.ravel()), rel(h[-Co:], gb)
    print(f"dW rel {ew:.2e}, db rel {eb:.2e} ({terms:.0f} terms per partial sum, bound {tol:.1e})")
    assert ew < tol and eb < tol


def test_network_gradients(gpu_lib):
    from fwant_amd import ae
    from fwant_amd._lib import DeviceArray
    from oracle import ae_ref
    rng = np.random.default_rng(9)
    w = ae.init_weights(3)
    w = [(a, (rng.standard_normal(b.shape) * 0.05).astype(np.float32)) for a, b in w]
    x = (rng.uniform(-1, 1, (2, 6, 16, 16, 16))).astype(np.float32)
    net = ae.AE(weights=w)
    xd = DeviceArray.from_host(x)
    out = net.forward(xd, 2)
    loss, grad = net.loss_and_grad(out, xd, 2)
    net.backward(grad, 2)
    got = net.gradients()
    oloss, ograds = ae_ref.loss_and_grads(ae.LAYERS, w, x.astype(np.float64), round_bf16=True)
    print(f"loss GPU {loss:.6e} oracle {oloss:.6e}")
    assert abs(loss - oloss) < 1e-3 * oloss
    for q, ((gw, gb), (ow, ob)) in enumerate(zip(got, ograds)):
        ew, eb = rel(gw, ow), rel(gb, ob)
        print(f"layer {q}: dW rel {ew:.2e}, db rel {eb:.2e}")
        assert ew < 2e-2 and eb < 2e-2


def sgt_records():
    from fwant_amd import models
    from fwant_amd.solver import Solver
    m = models.small_test(ni=64, nj=64, nk=32, nt=120, seed=91, moment=False)
    snap = dict(subs=(4, 4, 4), subc=(64, 64, 32), subt=(1, 1, 1), tinv=3,
                fields=("Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz"), max_records=41)
    s = Solver(m, snapshots=[snap])
    s.run(120)
    rec = s.snapshot(0)
    s.close()
    return rec[8:], snap["subc"]  # skip the first records, before the waves fill many blocks


def test_training_beats_the_method_without_it(gpu_lib):
    from fwant_amd import ae
    rec, subc = sgt_records()
    blocks, scales = ae.blocks_of(rec, subc)
    n = blocks.shape[0]
    live = scales.max(axis=1) > 0
    host = blocks.to_host().reshape(n, -1)
    rng = np.random.default_rng(0)
    idx = rng.permutation(np.nonzero(live)[0])
    ntest = len(idx) // 5
    test_i, train_i = idx[:ntest], idx[ntest:]
    from fwant_amd._lib import DeviceArray
    train_blocks = DeviceArray.from_host(host[train_i].reshape(-1, 6, 16, 16, 16))
    net = ae.AE(seed=1)
    print(f"{len(train_i)} training blocks, {ntest} held out", flush=True)
    hist = ae.train(net, train_blocks, len(train_i), iters=1500, batch=16, lr=2e-3, log=print)
    assert np.mean(hist[-100:]) < 0.5 * np.mean(hist[:10])
    test_blocks = DeviceArray.from_host(host[test_i].reshape(-1, 6, 16, 16, 16))
    dec, _ = ae.reconstruct(net, test_blocks, ntest)
    e_ae = ae.block_errors(host[test_i], dec, scales[test_i])
    e_lin = ae.block_errors(host[test_i], ae.linear_baseline(host[test_i]), scales[test_i])
    print(f"held-out block error: autoencoder median {np.median(e_ae):.3e}, 4^3 averages median {np.median(e_lin):.3e}")
    for t in (1e-3, 1e-2, 5e-2, 1e-1):
        print(f"  accepted at {t:g}: autoencoder {np.mean(e_ae <= t):.3f}, averages {np.mean(e_lin <= t):.3f}")
    assert np.median(e_ae) < np.median(e_lin)
    # the compressed store with the per-block fallback: every block within the threshold
    out, stats = ae.compressed_records(net, rec, subc, threshold=1e-3)
    nb = stats.accepted.size
    print(f"store at 1e-3: {stats.accepted.mean():.3f} of {nb} blocks as latents, "
          f"{stats.stored_bytes / stats.fp32_bytes:.3f} of the FP32 bytes")
    errs = [rel(out[r, c], rec[r, c]) for r in range(rec.shape[0]) for c in range(6) if np.any(rec[r, c])]
    assert max(errs) <= 1e-3
    # records round trip through the block layout exactly
    b, sc = ae.blocks_of(rec, subc)
    np.testing.assert_allclose(ae.records_from_blocks(b.to_host(), sc, subc, rec.shape[0]), rec, rtol=1e-6, atol=0)

