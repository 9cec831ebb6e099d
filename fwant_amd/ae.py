"""Learned SGT compression: a small 3-D convolutional autoencoder on 16^3
blocks of the 6 stress components, trained and run on the GPU by the
library's own kernels (fwae_*), with a per-block FP32 fallback.

Network (FP32 weights, activations rounded to BF16 at every layer):
  encoder  conv 6->16 k3 s1  | conv 16->16 k4 s2 (8^3) | conv 16->16 k4 s2 (4^3) | conv 16->8 k3 s1 (latent 8x4^3)
  decoder  conv 8->16 k3 s1  | convT 16->16 k4 s2 (8^3) | convT 16->16 k4 s2 (16^3) | conv 16->6 k3 s1
leaky ReLU (slope 0.1) between layers, linear latent and output.  Blocks
are scaled per block and component to [-1, 1] (fwae_blocks); the loss is
the mean squared error in that scaled space.

compress(records, subc, threshold): a block is kept as its 512 latent
values (BF16) plus 6 FP32 scales when the decoded block's relative L2
error is at most `threshold`, otherwise it is kept uncompressed (FP32).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DeviceArray, kcheck

SLOPE = 0.1
# (kind, Cin, Cout, size in, size out, K, S, P, activation slope; 1.0 = linear)
LAYERS = (
    ("conv", 6, 16, 16, 16, 3, 1, 1, SLOPE),
    ("conv", 16, 16, 16, 8, 4, 2, 1, SLOPE),
    ("conv", 16, 16, 8, 4, 4, 2, 1, SLOPE),
    ("conv", 16, 8, 4, 4, 3, 1, 1, 1.0),
    ("conv", 8, 16, 4, 4, 3, 1, 1, SLOPE),
    ("convT", 16, 16, 4, 8, 4, 2, 1, SLOPE),
    ("convT", 16, 16, 8, 16, 4, 2, 1, SLOPE),
    ("conv", 16, 6, 16, 16, 3, 1, 1, 1.0),
)
N_ENC = 4
BLOCK = 16 ** 3
NCOMP = 6
LATENT = 8 * 4 ** 3


def _p(a):
    return C.c_void_p(a if isinstance(a, int) else a.ptr)


def init_weights(seed=0, layers=LAYERS):
    """He-scaled normal weights, zero biases: list of (w, b) host arrays.
    Conv weight [Cout][Cin][K^3]; transposed conv weight [Cin][Cout][K^3]."""
    rng = np.random.default_rng(seed)
    out = []
    for kind, ci, co, di, do, k, s, p, a in layers:
        fan_in = ci * k ** 3 if kind == "conv" else ci * k ** 3 // (s ** 3)
        std = np.sqrt(2.0 / fan_in)
        shape = (co, ci, k ** 3) if kind == "conv" else (ci, co, k ** 3)
        out.append(((rng.standard_normal(shape) * std).astype(np.float32), np.zeros(co, np.float32)))
    return out


class AE:
    """Weights and Adam state in HBM; forward / backward / step on batches.
    layers: the network (default LAYERS), n_enc: its encoder layers."""

    def __init__(self, weights=None, seed=0, nchunk=256, layers=LAYERS, n_enc=N_ENC):
        self.L = _lib.lib()
        self.layers, self.n_enc = tuple(layers), n_enc
        _, _, co, _, do, _, _, _, _ = self.layers[n_enc - 1]
        self.latent = co * do ** 3
        self.host = weights if weights is not None else init_weights(seed, self.layers)
        self.w = [DeviceArray.from_host(w) for w, _ in self.host]
        self.b = [DeviceArray.from_host(b) for _, b in self.host]
        # gradient buffer per layer: dw then db (the layout fwae_conv_bwd_weight writes)
        self.g = [DeviceArray((w.numel() + max(L[1], L[2]),), np.float32) for w, L in zip(self.w, self.layers)]
        self.m = [DeviceArray.from_host(np.zeros(g.numel(), np.float32)) for g in self.g]
        self.v = [DeviceArray.from_host(np.zeros(g.numel(), np.float32)) for g in self.g]
        self.nchunk = nchunk
        self.part = DeviceArray((nchunk * max(g.numel() for g in self.g),), np.float32)
        self.step_no = 0
        self._acts = {}

    # -- buffers -------------------------------------------------------------
    def _buf(self, key, shape):
        a = self._acts.get(key)
        if a is None or a.shape != shape:
            a = DeviceArray(shape, np.float32)
            self._acts[key] = a
        return a

    # -- forward ----------------------------------------------------------------
    def _layer_fwd(self, q, x, n, y):
        # the activation (leaky ReLU, BF16 rounding) is applied as the
        # convolution stores its output
        kind, ci, co, di, do, k, s, p, a = self.layers[q]
        if kind == "conv":
            kcheck(self.L.fwae_conv_fwd_act(n, ci, co, di, do, k, s, p, _p(x), _p(self.w[q]), _p(self.b[q]), a, _p(y),
                                            None), "fwae_conv_fwd_act")
        else:  # transposed: the conv's backward data with the roles of the channels swapped
            kcheck(self.L.fwae_conv_bwd_data_act(n, co, ci, do, di, k, s, p, _p(x), _p(self.w[q]), _p(self.b[q]), a,
                                                 _p(y), None), "fwae_conv_bwd_data_act")

    def forward(self, x, n, layers=None, tag="t"):
        """x: DeviceArray of n inputs of the first layer in `layers` (default:
        all); returns the last layer's output (activations kept for backward)."""
        cur = x
        self._inputs = {}
        for q in (range(len(self.layers)) if layers is None else layers):
            kind, ci, co, di, do, k, s, p, a = self.layers[q]
            y = self._buf((tag, q), (n, co, do, do, do))
            self._inputs[q] = cur
            self._layer_fwd(q, cur, n, y)
            cur = y
        return cur

    def encode(self, x, n):
        return self.forward(x, n, range(self.n_enc), tag="e")

    def decode(self, z, n):
        return self.forward(z, n, range(self.n_enc, len(self.layers)), tag="d")

    # -- backward ---------------------------------------------------------------
    def backward(self, dy, n):
        """dy: gradient of the loss w.r.t. the network output (overwritten)."""
        cur = dy
        for q in reversed(range(len(self.layers))):
            kind, ci, co, di, do, k, s, p, a = self.layers[q]
            y = self._acts[("t", q)]
            x = self._inputs[q]
            kcheck(self.L.fwae_act_bwd(y.numel(), a, _p(y), _p(cur), None), "fwae_act_bwd")
            if kind == "conv":
                kcheck(self.L.fwae_conv_bwd_weight(n, ci, co, di, do, k, s, p, _p(x), _p(cur), self.nchunk,
                                                   _p(self.part), _p(self.g[q]), None), "fwae_conv_bwd_weight")
            else:
                # dW of the transposed layer: the conv's weight gradient with input = this layer's
                # output gradient and output gradient = this layer's input; its "bias" part is
                # replaced below by the sum of the output gradient per output channel
                kcheck(self.L.fwae_conv_bwd_weight(n, co, ci, do, di, k, s, p, _p(cur), _p(x), self.nchunk,
                                                   _p(self.part), _p(self.g[q]), None), "fwae_conv_bwd_weight")
                self._bias_grad_T(q, cur, n)
            if q == 0:
                break
            dx = self._buf(("g", q), (n, ci, di, di, di))
            if kind == "conv":
                kcheck(self.L.fwae_conv_bwd_data(n, ci, co, di, do, k, s, p, _p(cur), _p(self.w[q]), None, _p(dx),
                                                 None), "fwae_conv_bwd_data")
            else:
                kcheck(self.L.fwae_conv_fwd(n, co, ci, do, di, k, s, p, _p(cur), _p(self.w[q]), None, _p(dx), None),
                       "fwae_conv_fwd")
            cur = dx

    def _bias_grad_T(self, q, dz, n):
        """Bias gradient of a transposed layer: sum of dz per output channel,
        computed as the weight gradient of a 1x1x1 convolution from an
        all-ones input channel."""
        kind, ci, co, di, do, k, s, p, a = self.layers[q]
        ones = self._buf(("ones", q), (n, 1, do, do, do))
        if not getattr(ones, "_filled", False):
            kKK = np.ones((n, 1, do, do, do), np.float32)
            _lib.lib().fw_memcpy_h2d(C.c_void_p(ones.ptr), kKK.ctypes.data_as(C.c_void_p), ones.nbytes)
            ones._filled = True
        tmp = self._buf(("bT", q), (co * 1 * 1 + co,))
        kcheck(self.L.fwae_conv_bwd_weight(n, 1, co, do, do, 1, 1, 0, _p(ones), _p(dz), self.nchunk, _p(self.part),
                                           _p(tmp), None), "fwae_conv_bwd_weight")
        nw = self.w[q].numel()
        kcheck(_lib.lib().fw_memcpy_d2d(C.c_void_p(self.g[q].ptr + 4 * nw), C.c_void_p(tmp.ptr), 4 * co),
               "fw_memcpy_d2d")

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_no += 1
        for q in range(len(self.layers)):
            nw = self.w[q].numel()
            kcheck(self.L.fwae_adam(nw, _p(self.w[q]), _p(self.g[q]), _p(self.m[q]), _p(self.v[q]), lr, b1, b2, eps,
                                    self.step_no, None), "fwae_adam")
            kcheck(self.L.fwae_adam(self.b[q].numel(), _p(self.b[q]), C.c_void_p(self.g[q].ptr + 4 * nw),
                                    C.c_void_p(self.m[q].ptr + 4 * nw), C.c_void_p(self.v[q].ptr + 4 * nw), lr, b1,
                                    b2, eps, self.step_no, None), "fwae_adam")

    def loss_and_grad(self, out, target, n):
        grad = self._buf("grad", out.shape)
        loss = DeviceArray((n,), np.float64)
        kcheck(self.L.fwae_mse(n, out.numel() // n, _p(out), _p(target), _p(grad), _p(loss), None), "fwae_mse")
        return float(loss.to_host().sum() / out.numel()), grad

    def weights(self):
        return [(w.to_host(), b.to_host()) for w, b in zip(self.w, self.b)]

    def gradients(self):
        out = []
        for q, g in enumerate(self.g):
            h = g.to_host()
            nw = self.w[q].numel()
            out.append((h[:nw].reshape(self.w[q].shape), h[nw:nw + self.b[q].numel()]))
        return out


def blocks_of(records, subc):
    """Snapshot records [nrec][6][npts] (host) -> scaled blocks
    [nrec*nblk][6][16][16][16] (DeviceArray) and scales [nrec*nblk][6] (host)."""
    records = np.ascontiguousarray(records, np.float32)
    nrec, nc, npts = records.shape
    nb = [(c + 15) // 16 for c in subc]
    nblk = nb[0] * nb[1] * nb[2]
    rec = DeviceArray.from_host(records)
    out = DeviceArray((nrec * nblk, nc, 16, 16, 16), np.float32)
    sc = DeviceArray((nrec * nblk, nc), np.float32)
    kcheck(_lib.lib().fwae_blocks(_p(rec), nrec, nc, int(subc[0]), int(subc[1]), int(subc[2]), _p(out), _p(sc), None),
           "fwae_blocks")
    return out, sc.to_host()


def train(ae, blocks, n_blocks, iters=500, batch=32, lr=1e-3, seed=0, log=None):
    """Adam on random batches of the scaled blocks (learning rate cut by 3x
    at 60% and 85% of the iterations); returns the loss history."""
    rng = np.random.default_rng(seed)
    host = blocks.to_host().reshape(n_blocks, -1)
    hist = []
    x = DeviceArray((batch, NCOMP, 16, 16, 16), np.float32)
    for it in range(iters):
        idx = rng.integers(0, n_blocks, batch)
        xb = np.ascontiguousarray(host[idx])
        _lib.lib().fw_memcpy_h2d(C.c_void_p(x.ptr), xb.ctypes.data_as(C.c_void_p), x.nbytes)
        out = ae.forward(x, batch)
        loss, grad = ae.loss_and_grad(out, x, batch)
        ae.backward(grad, batch)
        ae.step(lr * (1.0 if it < 0.6 * iters else (1 / 3 if it < 0.85 * iters else 1 / 9)))
        hist.append(loss)
        if log and (it % 100 == 0 or it == iters - 1):
            log(f"iter {it}: loss {loss:.3e}")
    return hist


def reconstruct(ae, blocks, n, batch=64):
    """Decoded scaled blocks (host) and the latents (host, BF16-rounded by the network)."""
    host = blocks.to_host().reshape(n, -1)
    out = np.zeros_like(host)
    lat = np.zeros((n, ae.latent), np.float32)
    x = DeviceArray((batch, NCOMP, 16, 16, 16), np.float32)
    for a in range(0, n, batch):
        m = min(batch, n - a)
        xb = np.zeros((batch, host.shape[1]), np.float32)
        xb[:m] = host[a:a + m]
        _lib.lib().fw_memcpy_h2d(C.c_void_p(x.ptr), xb.ctypes.data_as(C.c_void_p), x.nbytes)
        z = ae.encode(x, batch)
        y = ae.decode(z, batch)
        out[a:a + m] = y.to_host().reshape(batch, -1)[:m]
        lat[a:a + m] = z.to_host().reshape(batch, -1)[:m]
    return out, lat


@dataclass
class Compressed:
    accepted: np.ndarray    # per block: True = latent, False = FP32 fallback
    errors: np.ndarray      # decoded relative L2 error of every block
    stored_bytes: int
    fp32_bytes: int


def block_errors(blocks_true, blocks_dec, scales):
    """Relative L2 error per block in physical units (scales per component)."""
    n = blocks_true.shape[0]
    t = blocks_true.reshape(n, NCOMP, -1) * scales[:, :, None]
    d = blocks_dec.reshape(n, NCOMP, -1) * scales[:, :, None]
    num = np.sqrt(((d.astype(np.float64) - t) ** 2).sum(axis=(1, 2)))
    den = np.sqrt((t.astype(np.float64) ** 2).sum(axis=(1, 2)))
    return np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)


def compress(ae, records, subc, threshold=1e-3):
    """Per-block choice between the latent and the FP32 block; returns the
    decoded records (accepted blocks decoded, the rest exact) and the stats."""
    blocks, scales = blocks_of(records, subc)
    n = blocks.shape[0]
    dec, _ = reconstruct(ae, blocks, n)
    true = blocks.to_host().reshape(n, -1)
    err = block_errors(true, dec, scales)
    zero = scales.max(axis=1) == 0
    acc = (err <= threshold) | zero
    stored = int(acc.sum()) * (ae.latent * 2 + NCOMP * 4) + int((~acc).sum()) * NCOMP * BLOCK * 4
    return Compressed(acc, err, stored, n * NCOMP * BLOCK * 4), dec, true, scales


def linear_baseline(blocks_host):
    """The method without a network at the same latent size: 16^3 -> 4^3
    averages per component (8 of 6x64 values kept as 6 channels of 4^3,
    i.e. 384 values, less than the 512 latents) and trilinear-free
    piecewise-constant upsampling."""
    n = blocks_host.shape[0]
    b = blocks_host.reshape(n, NCOMP, 4, 4, 4, 4, 4, 4)
    avg = b.mean(axis=(3, 5, 7))
    up = np.repeat(np.repeat(np.repeat(avg, 4, axis=2), 4, axis=3), 4, axis=4)
    return up.reshape(n, -1)


def records_from_blocks(blocks_scaled, scales, subc, nrec):
    """Inverse of blocks_of on the host: [nrec*nblk][6][16^3] scaled blocks
    and their scales -> records [nrec][6][npts] (padding dropped)."""
    nb = [(c + 15) // 16 for c in subc]
    nblk = nb[0] * nb[1] * nb[2]
    b = (blocks_scaled.reshape(nrec, nblk, NCOMP, BLOCK) * scales.reshape(nrec, nblk, NCOMP, 1)).astype(np.float32)
    b = b.reshape(nrec, nb[2], nb[1], nb[0], NCOMP, 16, 16, 16)
    vol = b.transpose(0, 4, 1, 5, 2, 6, 3, 7).reshape(nrec, NCOMP, nb[2] * 16, nb[1] * 16, nb[0] * 16)
    vol = vol[:, :, : subc[2], : subc[1], : subc[0]]
    return np.ascontiguousarray(vol.reshape(nrec, NCOMP, -1))


def compressed_records(ae, records, subc, threshold=1e-3):
    """Records as the compressed store gives them back: accepted blocks
    decoded, the others exact; with the storage statistics."""
    stats, dec, true, scales = compress(ae, records, subc, threshold)
    mixed = np.where(stats.accepted[:, None], dec, true)
    return records_from_blocks(mixed, scales, subc, records.shape[0]), stats

