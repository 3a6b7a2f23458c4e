import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "grid-mg-bench_amd", "python"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # build only when nothing is built yet (the driver builds before testing)
    need = [os.path.join(ROOT, "grid-mg-bench_amd", "libmgc.so"),
            os.path.join(ROOT, "oracle", "liboracle.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device; GPU tests fail loudly without one."""
    import mgc
    n = mgc.device_count()
    assert n >= 1, "no HIP device visible"
    print(f"[gpu] {mgc.device_name()}", flush=True)
    return mgc
