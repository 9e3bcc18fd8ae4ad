"""Build the HIP/C++ library libcrabsaxs_hip.so in-tree for gfx950.

Usage: python -m crabsaxs_amd.build  (or build() from __graft_entry__).
Compiles every source under csrc/ with hipcc; host-only physics code is built
with -ffp-contract=off so no multiply-add is fused, matching the FP64
operation order of the reference.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libcrabsaxs_hip.so"
ARCH = os.environ.get("CRABSAXS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]
# Packed FP32 VALU instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_ONLY = ["-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]

SOURCES = [
    ("physics.cpp", "host"),
    ("capi.cpp", "host"),
    ("multipole.cpp", "host"),
    ("fit.cpp", "host"),
    ("adjoint.cpp", "host"),
    ("rng.cpp", "host"),
    ("pr.cpp", "host"),
    ("grid.cpp", "host"),
    ("ingest.cpp", "host"),
    ("debye.hip", "device"),
    ("multipole.hip", "device"),
    ("adjoint.hip", "device"),
    ("ensemble.hip", "device-exact"),
    ("pr.hip", "device-exact"),
    ("grid.hip", "device-exact"),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    with tempfile.TemporaryDirectory(prefix="csx_build_") as tmp:
        objs = []
        for name, kind in SOURCES:
            src = CSRC / name
            obj = Path(tmp) / (src.name + ".o")
            if kind.startswith("device"):
                # device-exact: no FMA contraction, so FP64 results match the
                # host restatement bit for bit
                extra = ["-ffp-contract=off"] if kind == "device-exact" else []
                cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, *extra, "-c", str(src), "-o", str(obj)]
            else:
                cmd = [HIPCC, "-x", "c++", *COMMON, *HOST_ONLY, "-c", str(src), "-o", str(obj)]
            if verbose:
                print(" ".join(cmd))
            _run(cmd)
            objs.append(str(obj))
        tmp_lib = Path(tmp) / LIB.name
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp_lib)])
        shutil.copyfile(tmp_lib, str(LIB) + ".tmp")
        os.replace(str(LIB) + ".tmp", LIB)
    return LIB


def build_cpp_suite() -> Path:
    """The C++ API test program tests/cpp/reference_suite (host compiler,
    header-only API over the library just built)."""
    root = PKG.parent
    src = root / "tests" / "cpp" / "reference_suite.cpp"
    out = src.with_suffix("")
    cxx = shutil.which("g++") or "c++"
    tmp = str(out) + ".tmp"
    _run([cxx, "-std=c++17", "-O1", "-Wall", "-Wextra", f"-I{root / 'include'}", str(src),
          "-o", tmp, f"-L{PKG}", "-lcrabsaxs_hip", "-Wl,-rpath,$ORIGIN/../../crabsaxs_amd",
          f"-Wl,-rpath-link,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_suite())
