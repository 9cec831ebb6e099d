"""TEST INFRASTRUCTURE ONLY: a numpy (FP64) restatement of the SGT block
autoencoder of fwant_amd/ae.py -- convolution forward, backward data and
backward weights, leaky ReLU with BF16 rounding of stored activations, MSE
-- used by tests/ to check the GPU kernels and gradients.  Its backward
pass is pinned by central finite differences (tests/test_ae_oracle.py)."""
import numpy as np


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(f), f, out)


def conv_fwd(x, w, b, S, P):
    """x [N,Ci,D,D,D], w [Co,Ci,K,K,K] -> y [N,Co,Do,Do,Do]."""
    N, Ci, D = x.shape[0], x.shape[1], x.shape[2]
    Co, K = w.shape[0], w.shape[2]
    Do = (D + 2 * P - K) // S + 1
    xp = np.zeros((N, Ci, D + 2 * P + S, D + 2 * P + S, D + 2 * P + S))
    xp[:, :, P:P + D, P:P + D, P:P + D] = x
    y = np.zeros((N, Co, Do, Do, Do))
    for kz in range(K):
        for ky in range(K):
            for kx in range(K):
                sl = xp[:, :, kz:kz + S * Do:S, ky:ky + S * Do:S, kx:kx + S * Do:S]
                y += np.einsum("nczyx,oc->nozyx", sl, w[:, :, kz, ky, kx])
    if b is not None:
        y += b[None, :, None, None, None]
    return y


def conv_bwd_data(dy, w, S, P, D):
    N, Co, Do = dy.shape[0], dy.shape[1], dy.shape[2]
    Ci, K = w.shape[1], w.shape[2]
    dxp = np.zeros((N, Ci, D + 2 * P + S, D + 2 * P + S, D + 2 * P + S))
    for kz in range(K):
        for ky in range(K):
            for kx in range(K):
                dxp[:, :, kz:kz + S * Do:S, ky:ky + S * Do:S, kx:kx + S * Do:S] += \
                    np.einsum("nozyx,oc->nczyx", dy, w[:, :, kz, ky, kx])
    return dxp[:, :, P:P + D, P:P + D, P:P + D]


def conv_bwd_weight(x, dy, S, P, K):
    N, Ci, D = x.shape[0], x.shape[1], x.shape[2]
    Co, Do = dy.shape[1], dy.shape[2]
    xp = np.zeros((N, Ci, D + 2 * P + S, D + 2 * P + S, D + 2 * P + S))
    xp[:, :, P:P + D, P:P + D, P:P + D] = x
    dw = np.zeros((Co, Ci, K, K, K))
    for kz in range(K):
        for ky in range(K):
            for kx in range(K):
                sl = xp[:, :, kz:kz + S * Do:S, ky:ky + S * Do:S, kx:kx + S * Do:S]
                dw[:, :, kz, ky, kx] = np.einsum("nozyx,nczyx->oc", dy, sl)
    return dw, dy.sum(axis=(0, 2, 3, 4))


def forward(layers, weights, x, round_bf16=True):
    """Returns the per-layer (input, pre-activation, output)."""
    acts = []
    cur = x
    for (kind, ci, co, di, do, k, s, p, a), (w, b) in zip(layers, weights):
        W = w.reshape(w.shape[0], w.shape[1], k, k, k).astype(np.float64)
        if kind == "conv":
            z = conv_fwd(cur, W, b.astype(np.float64), s, p)
        else:
            z = conv_bwd_data(cur, W, s, p, do) + b.astype(np.float64)[None, :, None, None, None]
        y = np.where(z > 0, z, a * z)
        if round_bf16:
            y = bf16_round(y.astype(np.float32)).astype(np.float64)
        acts.append((cur, z, y))
        cur = y
    return acts


def loss_and_grads(layers, weights, x, round_bf16=True):
    """MSE(out, x) and the gradients (dw in each layer's storage layout, db)."""
    acts = forward(layers, weights, x, round_bf16)
    out = acts[-1][2]
    n = out.size
    loss = ((out - x) ** 2).sum() / n
    g = 2 * (out - x) / n
    grads = [None] * len(layers)
    for q in reversed(range(len(layers))):
        kind, ci, co, di, do, k, s, p, a = layers[q]
        w = weights[q][0].reshape(weights[q][0].shape[0], weights[q][0].shape[1], k, k, k).astype(np.float64)
        inp, z, y = acts[q]
        g = np.where(z > 0, g, a * g)
        if kind == "conv":
            dw, db = conv_bwd_weight(inp, g, s, p, k)
            dx = conv_bwd_data(g, w, s, p, di)
        else:
            dw, _ = conv_bwd_weight(g, inp, s, p, k)  # [Cin][Cout] layout of the transposed weight
            db = g.sum(axis=(0, 2, 3, 4))
            dx = conv_fwd(g, w, None, s, p)
        grads[q] = (dw.reshape(weights[q][0].shape), db)
        g = dx
    return loss, grads
