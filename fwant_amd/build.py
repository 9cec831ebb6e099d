"""Build libfwant.so (HIP, gfx950) in-tree.

The library is compiled from fwant_amd/csrc/*.hip with hipcc for gfx950 only.
Packed-FP32 instruction selection is disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfwant.so")
SOURCES = ["fw_solver.hip", "fw_kernel.hip", "fw_inv.hip", "fw_ae.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "fw_drive")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = True) -> str:
    """Fortran driver over the ISO_C_BINDING module (fortran/fwant_c.f90),
    linked against the in-tree libfwant.so (rpath $ORIGIN/..)."""
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, "fwant_c.o")
    steps = [
        [FLANG, "-O2", "-module-dir", bdir, "-c", os.path.join(FORTRAN, "fwant_c.f90"), "-o", obj],
        [FLANG, "-O2", "-module-dir", bdir, os.path.join(FORTRAN, "fw_drive.f90"), obj, "-L" + HERE, "-lfwant",
         "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed building the Fortran driver")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfwant.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    build_fortran()
