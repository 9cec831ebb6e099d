"""Build libfwant.so (HIP, gfx950) in-tree.

The library is compiled from fwant_amd/csrc/*.hip with hipcc for gfx950 only,
one object per translation unit, in parallel: the stage kernels
(fw_stage_inst.hip) are compiled once per (grid kind, tile shape, half of
the F/B patterns).  Objects go
to fwant_amd/_build/, the library to fwant_amd/libfwant.so.
Packed-FP32 instruction selection is disabled (-packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfwant.so")
SOURCES = ["fw_solver.hip", "fw_kernel.hip", "fw_inv.hip", "fw_ae.hip"]
STAGE_SRC = "fw_stage_inst.hip"
NTILE = 4
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # the stage kernel reads its 624-byte argument struct through several
    # hundred field accesses; above InstCombine's default limit of users the
    # struct is copied to scratch instead of read from the kernel-argument
    # segment
    "-mllvm", "-instcombine-max-copied-from-constant-users=4000",
]


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "fw_drive")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def build_fortran(verbose: bool = True) -> str:
    """Fortran driver over the ISO_C_BINDING module (fortran/fwant_c.f90)
    and the drop-in loop modules (fortran/fwant_shim.f90), linked against the
    in-tree libfwant.so (rpath $ORIGIN/..)."""
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, "fwant_c.o")
    shim = os.path.join(bdir, "fwant_shim.o")
    steps = [
        [FLANG, "-O2", "-module-dir", bdir, "-c", os.path.join(FORTRAN, "fwant_c.f90"), "-o", obj],
        [FLANG, "-O2", "-module-dir", bdir, "-c", os.path.join(FORTRAN, "fwant_shim.f90"), "-o", shim],
        [FLANG, "-O2", "-module-dir", bdir, os.path.join(FORTRAN, "fw_drive.f90"), obj, shim, "-L" + HERE, "-lfwant",
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


def _units():
    """(object path, source path, extra flags) of every translation unit."""
    u = [(os.path.join(BUILD, s.replace(".hip", ".o")), os.path.join(CSRC, s), []) for s in SOURCES]
    for curv in range(2):
        for t in range(NTILE):
            for half in range(2):
                tag = "c" if curv else ""
                u.append((os.path.join(BUILD, f"fw_stage{tag}_{t}_{half}.o"), os.path.join(CSRC, STAGE_SRC),
                          [f"-DFW_TILE={t}", f"-DFW_PHALF={half}", f"-DFW_CURV={curv}"]))
    return u


def _stale(obj, src):
    """True when obj is older than its source or any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps.append(os.path.join(HERE, "..", "include", "fwant.h"))
    return any(os.path.exists(d) and os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = True, force: bool = True, jobs: int = None) -> str:
    """Compile every translation unit (all of them when force, else the stale
    ones) in parallel and link libfwant.so."""
    os.makedirs(BUILD, exist_ok=True)
    units = [u for u in _units() if force or _stale(u[0], u[1])]
    jobs = jobs or max(1, min(len(units) or 1, os.cpu_count() or 1))

    def compile_one(u):
        obj, src, extra = u
        cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", obj + ".tmp", src]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            return cmd, r.stdout + r.stderr
        os.replace(obj + ".tmp", obj)
        return cmd, None

    with cf.ThreadPoolExecutor(jobs) as ex:
        for cmd, err in ex.map(compile_one, units):
            if verbose:
                print(" ".join(cmd), flush=True)
            if err is not None:
                sys.stderr.write(err)
                raise RuntimeError("hipcc failed building " + cmd[-1])
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[u[0] for u in _units()]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libfwant.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--changed" not in sys.argv)
    build_fortran()
