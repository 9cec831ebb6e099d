"""fwant_amd: MI355X-native FWANT elastic-wave solver and sensitivity-kernel
integral (HIP kernels for gfx950 behind the C ABI of include/fwant.h)."""
from . import models  # noqa: F401

__all__ = ["models", "solver", "kernel"]
