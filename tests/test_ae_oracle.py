"""Pins of the numpy autoencoder restatement (oracle/ae_ref.py):
  - BF16 rounding equals torch's bfloat16 conversion;
  - convolution forward equals a direct sum at random positions;
  - backward data / weights are the adjoints of forward (<A x, y> = <x, A^T y>);
  - the whole network's gradients equal central finite differences (no
    BF16 rounding, FP64)."""
import numpy as np
import pytest

from oracle import ae_ref


def test_bf16_round_matches_torch():
    import torch
    x = (np.random.default_rng(0).standard_normal(10000) * 10.0 ** np.random.default_rng(1).uniform(-5, 5, 10000))
    x = x.astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(ae_ref.bf16_round(x), want)


@pytest.mark.parametrize("K,S,P,D", [(3, 1, 1, 6), (4, 2, 1, 8)])
def test_conv_forward_and_adjoints(K, S, P, D):
    rng = np.random.default_rng(K)
    N, Ci, Co = 2, 3, 4
    Do = (D + 2 * P - K) // S + 1
    x = rng.standard_normal((N, Ci, D, D, D))
    w = rng.standard_normal((Co, Ci, K, K, K))
    y = ae_ref.conv_fwd(x, w, None, S, P)
    for _ in range(10):
        n, o, z, yy, xx = (rng.integers(0, v) for v in (N, Co, Do, Do, Do))
        s = 0.0
        for c in range(Ci):
            for kz in range(K):
                for ky in range(K):
                    for kx in range(K):
                        iz, iy, ix = z * S - P + kz, yy * S - P + ky, xx * S - P + kx
                        if 0 <= iz < D and 0 <= iy < D and 0 <= ix < D:
                            s += w[o, c, kz, ky, kx] * x[n, c, iz, iy, ix]
        assert abs(y[n, o, z, yy, xx] - s) < 1e-10
    dy = rng.standard_normal(y.shape)
    dx = ae_ref.conv_bwd_data(dy, w, S, P, D)
    assert abs((y * dy).sum() - (x * dx).sum()) < 1e-8 * abs((y * dy).sum())
    dw, _ = ae_ref.conv_bwd_weight(x, dy, S, P, K)
    assert abs((dw * w).sum() - (y * dy).sum()) < 1e-8 * abs((y * dy).sum())


def test_network_gradients_finite_differences():
    from fwant_amd.ae import LAYERS, init_weights
    rng = np.random.default_rng(3)
    w = [(a.astype(np.float64), b.astype(np.float64) + rng.standard_normal(b.shape) * 0.01) for a, b in init_weights(5)]
    x = rng.standard_normal((1, 6, 16, 16, 16)) * 0.5
    loss, grads = ae_ref.loss_and_grads(LAYERS, w, x, round_bf16=False)
    h = 1e-6
    for q in (0, 3, 5, 7):
        for which in (0, 1):
            arr = w[q][which]
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = ae_ref.loss_and_grads(LAYERS, w, x, round_bf16=False)
                arr[idx] = old - h
                lm, _ = ae_ref.loss_and_grads(LAYERS, w, x, round_bf16=False)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                g = grads[q][which][idx]
                assert abs(fd - g) <= 1e-5 * max(abs(fd), 1e-8) + 1e-10, (q, which, fd, g)
