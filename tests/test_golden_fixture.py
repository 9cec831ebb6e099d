"""The committed golden fixture is well formed: 1008 records of 16
receivers x 9 fields in FP32 and FP64, finite, with the free-surface
traction condition visible (Tzz, Txz, Tyz vanish at surface receivers)."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "two_layer_100x100x50.npz")


def test_fixture_shape_and_physics():
    d = np.load(GOLD)
    r32, r64 = d["rec32"], d["rec64"]
    assert r32.shape == r64.shape == (1008, 16, 9)
    assert np.isfinite(r32).all() and np.isfinite(r64).all()
    stress = np.abs(r64[..., 3:5]).max()
    assert np.abs(r64[..., 5:6]).max() < 1e-12 * stress and np.abs(r64[..., 7:9]).max() < 1e-12 * stress
    # FP32 reference arithmetic vs FP64: ~2e-4 (FP32 coordinates in cell volumes)
    e = np.linalg.norm(r32.astype(float) - r64) / np.linalg.norm(r64)
    assert 1e-6 < e < 1e-3
