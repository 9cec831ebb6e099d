"""Pins of the oracle's reduced-precision SGT codec (oracle/fw_oracle.c
sgt_encode / sgt_decode):

  - FP32 -> FP16 rounding (nearest even, subnormals, overflow) against
    numpy's float16 conversion, bit for bit; FP16 -> FP32 for every code;
  - FP32 -> BF16 rounding against torch's bfloat16 conversion, bit for bit;
  - block scales against numpy maxima over the 16^3 blocks of a snapshot
    whose extents are not multiples of 16;
  - restore error within half a unit in the last place of the format.
"""
import numpy as np
import pytest

from oracle.pyoracle import Lib, sgt_decode, sgt_encode


def sample_values(n, seed):
    rng = np.random.default_rng(seed)
    mags = 10.0 ** rng.uniform(-12, 5.5, n)
    x = (mags * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    # exact ties of the FP16 grid (normal and subnormal) and boundary values
    k = rng.integers(1, 2047, 2000)
    e = rng.integers(-24, 15, 2000)
    ties = ((k + 0.5) * 2.0 ** (e - 10)).astype(np.float32)
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 2 ** -24, 2 ** -25, 3 * 2 ** -26, 2 ** -14,
                        1.0, 32768.0, 6.0e-8], np.float32)
    return np.concatenate([x, ties, -ties, special])


def test_fp16_rounding_matches_numpy():
    L = Lib(32)
    x = sample_values(50000, 1)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    got = np.array([L.f32_to_fp16(float(v)) for v in x], np.uint16)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


def test_fp16_decode_every_code():
    L = Lib(32)
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    want = codes.view(np.float16).astype(np.float32)
    got = np.array([L.fp16_to_f32(int(c)) for c in codes], np.float32)
    finite = np.isfinite(want)
    np.testing.assert_array_equal(got[finite], want[finite])
    assert np.all(np.isinf(got[np.isinf(want)]))


def test_bf16_rounding_matches_torch():
    import torch
    L = Lib(32)
    x = sample_values(50000, 2)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([L.f32_to_bf16(float(v)) for v in x], np.uint16)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_block_scales_and_restore_error(fmt):
    rng = np.random.default_rng(3)
    subc = (37, 21, 18)
    npts = int(np.prod(subc))
    x = (rng.standard_normal((2, 3, npts)) * 10.0 ** rng.uniform(-3, 3, (2, 3, 1))).astype(np.float32)
    x[1, 2, :] = 0.0                     # an all-zero slab
    codes, scales = sgt_encode(fmt, x, subc)
    nb = [(c + 15) // 16 for c in subc]
    assert scales.shape == (2, 3, nb[0] * nb[1] * nb[2])
    v = x.reshape(2, 3, subc[2], subc[1], subc[0])
    for b2 in range(nb[2]):
        for b1 in range(nb[1]):
            for b0 in range(nb[0]):
                blk = b0 + nb[0] * (b1 + nb[1] * b2)
                amax = np.abs(v[:, :, 16 * b2:16 * b2 + 16, 16 * b1:16 * b1 + 16, 16 * b0:16 * b0 + 16]).max(axis=(2, 3, 4))
                want = amax if fmt == "bf16" else amax * np.float32(1 / 32768)
                np.testing.assert_array_equal(scales[:, :, blk], want.astype(np.float32))
    y = sgt_decode(fmt, codes, scales, subc)
    assert np.all(y[1, 2] == 0.0)
    # half an ulp of the 8-bit (BF16) / 11-bit (FP16) significand, two FP32
    # roundings (x/amax and code*scale), and FP16's subnormal spacing
    ulp = 2.0 ** -7 if fmt == "bf16" else 2.0 ** -10
    err = np.abs(y.astype(np.float64) - x)
    bound = (0.5 * ulp + 2.0 ** -22) * np.abs(x) + 2.0 ** -39 * np.abs(x).max(axis=-1, keepdims=True)
    assert np.all(err <= bound)
    rel = np.linalg.norm(y - x) / np.linalg.norm(x)
    print(f"{fmt}: restore rel-L2 {rel:.2e}")
    assert rel < (2e-3 if fmt == "bf16" else 3e-4)


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_tiny_blocks_store_zeros(fmt):
    """Blocks whose largest magnitude is below 2^-100 (the numerical
    precursors ahead of a wavefront reach subnormal values) are stored as
    zeros with scale 0: 32768 / amax (FP16) or 1 / amax (BF16) would
    overflow and turn zeros into NaN codes.  Blocks just above keep their
    restore bound."""
    subc = (32, 16, 16)  # two blocks
    npts = int(np.prod(subc))
    x = np.zeros((4, npts), np.float32)
    v = x.reshape(4, 16, 16, 32)
    rng = np.random.default_rng(9)
    for s, amax in enumerate((1e-38, 1e-34, 2e-31, 1e-29)):
        blk = rng.standard_normal((16, 16, 16)).astype(np.float32)
        blk[0, 0, 0] = 0.0
        v[s, :, :, :16] = blk / np.abs(blk).max() * np.float32(amax)
    codes, scales = sgt_encode(fmt, x, subc)
    y = sgt_decode(fmt, codes, scales, subc)
    assert np.isfinite(y).all()
    tiny = np.float32(2.0 ** -100)
    for s in range(4):
        a = np.abs(v[s, :, :, :16]).max()
        if a < tiny:
            assert scales[s, 0] == 0.0 and np.all(y[s] == 0.0)
        else:
            err = np.abs(y[s].astype(np.float64) - x[s])
            assert err.max() <= 2.0 ** -7 * a
    assert np.all(scales[:, 1] == 0.0)  # the all-zero block
