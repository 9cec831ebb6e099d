"""The TomoKernel.conf runner end to end on the GPU against the FP32 oracle
pipeline (SI_ker_sta.F90:135-821): the case of tests/kernel_case.py with
station and event SGTs from the GPU solver.  Tolerance: 1e-5 relative L2
per kernel (the GPU sums in FP64, the oracle in FP32); points off the
kernel grid stay zero; a second run with first_run off and block_from
2 1 1 appends to the existing files and leaves them as they were (1e-6).
"""
import os

import pytest

import kernel_case as kc

pytestmark = pytest.mark.gpu


def test_kernel_stage_files_match_oracle_pipeline(gpu_lib, tmp_path):
    from fwant_amd import kernel_run as kr
    root = str(tmp_path)
    m = kc.model()
    rec = kc.solver_records(m)
    syn, T = kc.write_case(root, m, rec)
    print("kernel runner", flush=True)
    st = kr.run(os.path.join(root, "TomoKernel.conf"), "sta.list")
    want = kc.oracle_kernels(m, rec, syn, T)
    saved = kc.check_run(root, st, want)
    kc.check_append_run(root, saved)
