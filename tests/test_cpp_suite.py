"""The C++ API (include/crabsaxs.hpp) through tests/cpp/reference_suite: the
reference's own unit tests for the path restated in C++."""
import subprocess
from pathlib import Path

import pytest

SUITE = Path(__file__).resolve().parent / "cpp" / "reference_suite"


def run(mode):
    if not SUITE.exists():
        pytest.fail(f"{SUITE} is not built: run `python -m crabsaxs_amd.build`")
    p = subprocess.run([str(SUITE), mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    print(p.stdout)
    assert p.returncode == 0, p.stdout


def test_cpp_suite_host_tests():
    run("--cpu")


@pytest.mark.gpu
def test_cpp_suite_gpu_tests():
    run("--gpu")
