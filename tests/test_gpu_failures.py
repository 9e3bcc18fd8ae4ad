"""Failure handling on a device: an injected allocation failure
(CSX_DEVICE_MEMORY_LIMIT caps the library's device memory) is reported as
OutOfMemory and leaves the context usable.  Runs in a child process, since
the limit is read once per process."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import numpy as np
import crabsaxs_amd as csx
small = csx.generate_globule(100, 8.0, 1.0, seed=1)
q = np.linspace(0.0, 0.5, 20)
a = csx.debye_intensity(small, q)
print("small ok", flush=True)
big = csx.generate_globule(300000, 86.0, 1.0, seed=2)
try:
    csx.debye_intensity(big, q)
    print("big computed", flush=True)
except csx.OutOfMemory as e:
    print("out of memory:", e, flush=True)
b = csx.debye_intensity(small, q)
assert np.array_equal(a, b)
print("small again ok", flush=True)
"""


def run_child(limit):
    env = dict(os.environ)
    if limit is None:
        env.pop("CSX_DEVICE_MEMORY_LIMIT", None)
    else:
        env["CSX_DEVICE_MEMORY_LIMIT"] = str(limit)
    p = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=240)
    print(p.stdout, p.stderr[-2000:])
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stdout


def test_injected_allocation_failure_is_reported_and_recoverable():
    # 16 MiB: the small structure and its temporaries fit; the 300,000-atom
    # structure's device copies (about 20 MB) do not
    out = run_child(16 << 20)
    assert "small ok" in out and "out of memory:" in out and "small again ok" in out
    out = run_child(None)
    assert "big computed" in out and "small again ok" in out
