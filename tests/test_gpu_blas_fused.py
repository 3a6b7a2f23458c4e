"""The fused re-orthogonalisation passes of the Krylov solvers
(vmultisub_dot_dev, bmultisub_dots_dev) against the two-kernel forms they
replace, bit for bit, for FP64/FP32, paired/unpaired loads, single and batched
vectors, and basis sizes around every kernel template and LDS-staging boundary
(tests/cpp/test_fused_blas.hip, built with the library)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "grid-mg-bench_amd", "bin", "test_fused_blas")


def test_fused_reorthogonalisation_bit_equal(gpu):
    assert os.path.exists(BIN), "build the package first (make -C grid-mg-bench_amd)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
