import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session", autouse=True)
def _built_artifacts():
    """Build the oracle (and, in a container with hipcc, the library) when missing."""
    lib = ROOT / "crabsaxs_amd" / "libcrabsaxs_hip.so"
    if not lib.exists() and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run([sys.executable, "-m", "crabsaxs_amd.build"], cwd=ROOT, check=True)
    suite = ROOT / "tests" / "cpp" / "reference_suite"
    if lib.exists() and not suite.exists():
        subprocess.run([sys.executable, "-c",
                        "from crabsaxs_amd.build import build_cpp_suite; build_cpp_suite()"],
                       cwd=ROOT, check=True)
    oracle = ROOT / "oracle" / "liboracle.so"
    if not oracle.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    yield


@pytest.fixture(scope="session")
def gpu_ctx():
    """The default context with the result cache off, so that every call
    runs its kernels and comparisons between calls (launch shapes, batch
    against single, run against run) compare two computations.  The cache
    has its own tests, which turn it on."""
    import crabsaxs_amd as csx
    ctx = csx.default_context()
    ctx.set_cache_limit(0)
    return ctx
