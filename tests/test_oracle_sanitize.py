"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle's C restatement (oracle/fw_oracle.c) is built a second time with
-fsanitize=address,undefined (oracle/Makefile, target san) and the fast
oracle tests run again against that build in a child process that preloads
libasan: every sweep pattern, mode C (metric, surface solve, traction
imaging), sources, PML profiles and a curved grid with PML, the SGT codec,
the kernel routines and LSQR.  Any out-of-bounds access, use after free or
undefined operation ends the child with a report, and the test fails with it.
Leak checking is off (the Python interpreter's own allocations would be
reported).
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FILES = [
    "tests/test_oracle_ops.py",
    "tests/test_oracle_rhs.py",
    "tests/test_oracle_curved.py",
    "tests/test_oracle_src.py",
    "tests/test_oracle_src_injection.py",
    "tests/test_oracle_pml.py",
    "tests/test_oracle_topo.py",
    "tests/test_oracle_sgt_codec.py",
    "tests/test_oracle_kernel.py",
    "tests/test_oracle_kernel_variants.py",
    "tests/test_oracle_inversion.py",
]
# long runs whose code paths the cases above already cover
SLOW = ["rayleigh_speed_along", "flat_over_the_pml", "long_run_under", "pml_absorbs"]


def _libasan():
    cc = shutil.which("gcc")
    if cc is None:
        return None
    p = subprocess.run([cc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_clean_under_asan_and_ubsan():
    asan = _libasan()
    if asan is None:
        pytest.skip("gcc's libasan is not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ORACLE_SANITIZE"] = "1"
    env["ORACLE_THREADS"] = "2"
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
           "-k", " and ".join(f"not {s}" for s in SLOW)] + FILES
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert " passed" in out
