"""The C++ host API (grid-mg-bench_amd/include/mgc/mgc.hpp) driven from a plain
g++ program over the C ABI: Create, CoarsenOperator, the multigrid runChecks
and an FGMRES + multigrid solve (tests/cpp/test_cpp_api.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "grid-mg-bench_amd", "bin", "test_cpp_api")


def test_cpp_host_api(gpu):
    assert os.path.exists(BIN), "build the package first (make -C grid-mg-bench_amd)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
