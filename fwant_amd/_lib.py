"""ctypes binding of libfwant.so (the C ABI declared in include/fwant.h).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback in the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfwant.so")

FW_OK = 0
NFIELD = 9
FIELDS = ("Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")

# every symbol declared in include/fwant.h
SYMBOLS = (
    "fw_create", "fw_destroy", "fw_last_error", "fw_dims", "fw_set_grid", "fw_build_metric", "fw_build_metric3d",
    "fw_set_media", "fw_set_force", "fw_set_moment", "fw_set_receivers", "fw_add_snapshot",
    "fw_init", "fw_get_pml_profiles", "fw_step", "fw_stage", "fw_run", "fw_reference_steps", "fw_sync",
    "fw_get_receivers", "fw_get_snapshot", "fw_snapshot_device", "fw_add_snapshot_packed",
    "fw_get_snapshot_packed", "fw_snapshot_packed_device", "fwk_sgt_decode", "fw_set_batch",
    "fw_select_member", "fw_set_batch_width", "fw_use_graphs", "fw_autotune", "fw_set_cycle_origin",
    "fw_checkpoint_size", "fw_checkpoint_save", "fw_checkpoint_load", "fw_get_wavefield",
    "fw_set_wavefield", "fw_enable_timing", "fw_get_timing",
    "fwk_stress2strain", "fwk_scatter", "fwk_scatter_param", "fwk_vscatter", "fwk_eoshift", "fwk_reparam",
    "fwk_filter", "fwk_scatter2kernel", "fwk_scatter2kernel_noalloc", "fwk_rotate",
    "fwk_last_error", "fwae_conv_fwd", "fwae_conv_bwd_data", "fwae_conv_bwd_weight", "fwae_conv_fwd_act",
    "fwae_conv_bwd_data_act", "fwae_act_fwd", "fwae_act_bwd",
    "fwae_mse", "fwae_adam", "fwae_blocks", "fwk_assemble", "fwl_create", "fwl_solve", "fwl_destroy", "fwl_last_error", "fw_device_alloc", "fw_device_free", "fw_memcpy_h2d", "fw_memcpy_d2h",
    "fw_memcpy_d2d", "fw_device_sync", "fw_hbm_probe",
)


SGT_FORMATS = {"fp32": 0, "bf16": 1, "fp16": 2}


class FwConfig(C.Structure):
    _fields_ = [
        ("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
        ("stept", C.c_float),
        ("freenode", C.c_int),
        ("abs_number", C.c_int * 6),
        ("abs_velocity", C.c_float * 6),
        ("cfs_bmax", C.c_float * 6),
        ("cfs_amax", C.c_float * 6),
    ]


_lib = None

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p
H = C.c_void_p


def _sig(lib):
    S = {
        "fw_create": (C.c_int, [C.POINTER(FwConfig), C.POINTER(H)]),
        "fw_destroy": (C.c_int, [H]),
        "fw_last_error": (C.c_char_p, [H]),
        "fw_dims": (C.c_int, [H, ip, ip, ip, ip]),
        "fw_set_grid": (C.c_int, [H, fp, fp, fp, fp, fp, fp, fp, fp]),
        "fw_build_metric": (C.c_int, [H, dp, dp, dp, fp, fp, fp, fp, fp]),
        "fw_build_metric3d": (C.c_int, [H, dp, dp, dp, fp, fp, fp, fp]),
        "fw_set_media": (C.c_int, [H, fp, fp, fp]),
        "fw_set_force": (C.c_int, [H, C.c_int, C.c_int, C.c_int, ip, fp, fp, fp, fp, fp, fp, fp]),
        "fw_set_moment": (C.c_int, [H, C.c_int, C.c_int, C.c_int, ip, fp, fp, fp, fp, fp]),
        "fw_set_receivers": (C.c_int, [H, C.c_int, C.c_int, ip, C.c_int]),
        "fw_add_snapshot": (C.c_int, [H, ip, ip, ip, C.c_int, C.c_uint, C.c_int, ip]),
        "fw_init": (C.c_int, [H]),
        "fw_get_pml_profiles": (C.c_int, [H, fp, fp, fp]),
        "fw_step": (C.c_int, [H, C.c_int]),
        "fw_stage": (C.c_int, [H, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "fw_run": (C.c_int, [H, C.c_int, C.c_int]),
        "fw_reference_steps": (C.c_int, [C.c_int]),
        "fw_sync": (C.c_int, [H]),
        "fw_get_receivers": (C.c_int, [H, fp, ip]),
        "fw_get_snapshot": (C.c_int, [H, C.c_int, fp, ip]),
        "fw_snapshot_device": (C.c_int, [H, C.c_int, C.POINTER(vp), ip, C.POINTER(C.c_long)]),
        "fw_add_snapshot_packed": (C.c_int, [H, ip, ip, ip, C.c_int, C.c_uint, C.c_int, C.c_int, ip]),
        "fw_get_snapshot_packed": (C.c_int, [H, C.c_int, vp, fp, ip, C.POINTER(C.c_long)]),
        "fw_snapshot_packed_device": (C.c_int, [H, C.c_int, C.POINTER(vp), C.POINTER(vp), ip, C.POINTER(C.c_long),
                                                C.POINTER(C.c_long)]),
        "fwk_sgt_decode": (C.c_int, [C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_long, vp, vp]),
        "fw_set_batch": (C.c_int, [H, C.c_int, ip, ip, ip]),
        "fw_select_member": (C.c_int, [H, C.c_int]),
        "fw_set_batch_width": (C.c_int, [H, C.c_int]),
        "fw_use_graphs": (C.c_int, [H, C.c_int]),
        "fw_autotune": (C.c_int, [H, C.c_int, ip, ip, fp]),
        "fw_set_cycle_origin": (C.c_int, [H, C.c_int]),
        "fw_checkpoint_size": (C.c_int, [H, C.POINTER(C.c_size_t)]),
        "fw_checkpoint_save": (C.c_int, [H, vp]),
        "fw_checkpoint_load": (C.c_int, [H, vp]),
        "fw_get_wavefield": (C.c_int, [H, C.c_int, fp]),
        "fw_set_wavefield": (C.c_int, [H, C.c_int, fp]),
        "fw_enable_timing": (C.c_int, [H, C.c_int]),
        "fw_get_timing": (C.c_int, [H, dp, dp, ip]),
        "fwk_stress2strain": (C.c_int, [vp, C.c_long, C.c_long, C.c_long, C.c_int, vp, vp, vp]),
        "fwk_scatter": (C.c_int, [vp, C.c_long, C.c_long, vp, C.c_long, C.c_long, C.c_long, C.c_int,
                                  C.c_float, vp, vp, vp, vp, vp]),
        "fwk_scatter_param": (C.c_int, [C.c_int, vp, C.c_long, C.c_long, vp, C.c_long, C.c_long, C.c_long, C.c_int,
                                        C.c_float, vp, vp, vp, vp, vp]),
        "fwk_vscatter": (C.c_int, [vp, C.c_long, C.c_long, vp, C.c_long, C.c_long, C.c_long, C.c_int, C.c_float,
                                   vp, vp, vp]),
        "fwk_eoshift": (C.c_int, [vp, C.c_long, C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, vp]),
        "fwk_reparam": (C.c_int, [C.c_int, C.c_long, vp, vp, vp, vp, vp, vp, vp]),
        "fwk_filter": (C.c_int, [C.c_int, C.c_long, C.c_int, C.c_int, dp, dp, vp, vp, vp]),
        "fwk_scatter2kernel": (C.c_int, [C.c_long, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, fp, fp,
                                         vp, vp, vp, vp, vp, vp, vp]),
        "fwk_scatter2kernel_noalloc": (C.c_int, [C.c_long, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, fp,
                                                 fp, vp, vp, vp, vp, vp, vp, vp]),
        "fwk_rotate": (C.c_int, [C.c_long, vp, vp, vp, fp, vp]),
        "fwk_last_error": (C.c_char_p, []),
        "fwae_conv_fwd": (C.c_int, [C.c_int] * 8 + [vp, vp, vp, vp, vp]),
        "fwae_conv_bwd_data": (C.c_int, [C.c_int] * 8 + [vp, vp, vp, vp, vp]),
        "fwae_conv_bwd_weight": (C.c_int, [C.c_int] * 8 + [vp, vp, C.c_int, vp, vp, vp]),
        "fwae_conv_fwd_act": (C.c_int, [C.c_int] * 8 + [vp, vp, vp, C.c_float, vp, vp]),
        "fwae_conv_bwd_data_act": (C.c_int, [C.c_int] * 8 + [vp, vp, vp, C.c_float, vp, vp]),
        "fwae_act_fwd": (C.c_int, [C.c_long, C.c_float, vp, vp]),
        "fwae_act_bwd": (C.c_int, [C.c_long, C.c_float, vp, vp, vp]),
        "fwae_mse": (C.c_int, [C.c_int, C.c_long, vp, vp, vp, vp, vp]),
        "fwae_adam": (C.c_int, [C.c_long, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, vp]),
        "fwae_blocks": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp]),
        "fwk_assemble": (C.c_int, [C.c_long, vp, C.c_long, vp, vp, vp, vp, vp]),
        "fwl_create": (C.c_int, [C.c_long, C.c_long, C.c_long, vp, vp, vp, vp, C.POINTER(vp)]),
        "fwl_solve": (C.c_int, [vp, C.c_double, C.c_double, C.c_long, vp, vp, C.POINTER(C.c_long),
                                C.POINTER(C.c_double)]),
        "fwl_destroy": (C.c_int, [vp]),
        "fwl_last_error": (C.c_char_p, [vp]),
        "fw_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
        "fw_device_free": (C.c_int, [vp]),
        "fw_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
        "fw_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
        "fw_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
        "fw_device_sync": (C.c_int, []),
        "fw_hbm_probe": (C.c_int, [C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    }
    for name, (res, args) in S.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Return the loaded libfwant (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libfwant.so not found at {LIB_PATH}: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        _lib = C.CDLL(LIB_PATH)
        _sig(_lib)
    return _lib


class FwError(RuntimeError):
    pass


def check(rc, handle=None, what=""):
    if rc != FW_OK:
        L = lib()
        msg = L.fw_last_error(handle) if handle is not None else L.fw_last_error(None)
        raise FwError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def kcheck(rc, what=""):
    if rc != FW_OK:
        msg = lib().fwk_last_error()
        raise FwError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def fptr(a):
    return a.ctypes.data_as(fp)


def dptr(a):
    return a.ctypes.data_as(dp)


def iptr(a):
    return a.ctypes.data_as(ip)


class DeviceArray:
    """A float32/float64 array in HBM owned by this object (allocated through
    the library's HIP runtime)."""

    def __init__(self, shape, dtype=np.float32, _ptr=None, strides=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        # element strides of a (component, sample, point) series; None: C order
        self.strides = strides
        self.owner = _ptr is None
        if _ptr is None:
            p = C.c_void_p()
            kcheck(lib().fw_device_alloc(self.nbytes, C.byref(p)), "fw_device_alloc")
            self.ptr = p.value
        else:
            self.ptr = int(_ptr)

    @classmethod
    def wrap(cls, ptr, shape, dtype=np.float32, strides=None):
        """A view of device memory owned elsewhere (e.g. snapshot storage):
        `strides` (component, sample) element strides of a 3-D series whose
        points are contiguous; the view covers prod(shape) elements."""
        return cls(shape, dtype, _ptr=ptr, strides=strides)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        kcheck(lib().fw_memcpy_h2d(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes), "fw_memcpy_h2d")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        kcheck(lib().fw_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes), "fw_memcpy_d2h")
        return out

    def copy(self):
        d = DeviceArray(self.shape, self.dtype)
        d.strides = self.strides
        kcheck(lib().fw_memcpy_d2d(C.c_void_p(d.ptr), C.c_void_p(self.ptr), self.nbytes), "fw_memcpy_d2d")
        return d

    def numel(self):
        return int(np.prod(self.shape))

    def free(self):
        if getattr(self, "ptr", None) and getattr(self, "owner", False):
            lib().fw_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    kcheck(lib().fw_device_sync(), "fw_device_sync")
