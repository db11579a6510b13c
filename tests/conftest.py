import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def built():
    """Build the oracle and the package libraries once (compile only, no GPU)."""
    import oracle

    oracle.lib()
    from acctricnt_amd import _ffi

    bins = [os.path.join(ROOT, "acctricnt_amd", "bin", b) for b in ("main_scanxp", "main_triangle_count")]
    if not all(os.path.exists(p) for p in [_ffi.GRAPH_LIB_PATH, _ffi.CN_LIB_PATH, *bins]):
        import subprocess

        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "acctricnt_amd", "csrc")], check=True)
    return True
