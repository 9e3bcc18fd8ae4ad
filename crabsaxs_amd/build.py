"""Build the HIP/C++ library libcrabsaxs_hip.so in-tree for gfx950.

Usage: python -m crabsaxs_amd.build  (or build() from __graft_entry__);
python -m crabsaxs_amd.build --host-asan DIR for the host-sanitized variant.
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
from concurrent.futures import ThreadPoolExecutor
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
# roctx ranges around the C-ABI calls
LINK = [f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"]

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


def build(verbose: bool = False, jobs: int = 0) -> Path:
    """Compile every source (in parallel, `jobs` at a time; 0: one per CPU,
    at most 8) and link the library."""
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="csx_build_") as tmp:
        objs, cmds = [], []
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
            cmds.append(cmd)
            objs.append(str(obj))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            for f in [pool.submit(_run, c) for c in cmds]:
                f.result()
        tmp_lib = Path(tmp) / LIB.name
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, *LINK, "-o", str(tmp_lib)])
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


def build_host_asan(out_dir: Path) -> Path:
    """The library and the C++ suite with the host code under AddressSanitizer
    (device code is not instrumented), into out_dir: out_dir/lib and
    out_dir/cpp/reference_suite, which finds the fixture files through
    out_dir/golden.  It is meant for the host-only paths, in a container
    without a GPU: `reference_suite --cpu` (typing, form factors, fitting,
    ingest, RNG, the no-device context path), which runs clean with
    ASAN_OPTIONS=detect_leaks=1.  With a GPU, the sanitizer runtime also
    tracks the HIP runtime's device allocations, and at process exit it can
    recycle them after the HIP runtime has unloaded (its internal check
    dev_runtime_unloaded_); the C++ API's default context is therefore
    never destroyed at exit, and the sanitized build is not run on a GPU."""
    out_dir = Path(out_dir).resolve()
    (out_dir / "lib").mkdir(parents=True, exist_ok=True)
    (out_dir / "cpp").mkdir(parents=True, exist_ok=True)
    golden = out_dir / "golden"
    if not golden.exists():
        golden.symlink_to(PKG.parent / "tests" / "golden")
    asan = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    lib = out_dir / "lib" / LIB.name
    with tempfile.TemporaryDirectory(prefix="csx_asan_") as tmp:
        objs = []
        for name, kind in SOURCES:
            src = CSRC / name
            obj = Path(tmp) / (src.name + ".o")
            if kind.startswith("device"):
                extra = ["-ffp-contract=off"] if kind == "device-exact" else []
                cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, *extra, *asan, "-c", str(src), "-o", str(obj)]
            else:
                cmd = [HIPCC, "-x", "c++", *COMMON, *HOST_ONLY, *asan, "-c", str(src), "-o", str(obj)]
            _run(cmd)
            objs.append(str(obj))
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *asan, *objs, *LINK, "-o", str(lib)])
    suite = out_dir / "cpp" / "reference_suite"
    _run([HIPCC, "-x", "c++", "-std=c++17", "-O1", "-g", *asan, "-D__HIP_PLATFORM_AMD__",
          f"-I{PKG.parent / 'include'}", str(PKG.parent / "tests" / "cpp" / "reference_suite.cpp"),
          "-o", str(suite), f"-L{lib.parent}", "-lcrabsaxs_hip", "-Wl,-rpath,$ORIGIN/../lib",
          f"-Wl,-rpath-link,{ROCM}/lib"])
    return suite


if __name__ == "__main__":
    if "--host-asan" in sys.argv:
        print(build_host_asan(Path(sys.argv[sys.argv.index("--host-asan") + 1])))
    else:
        print(build(verbose="-v" in sys.argv))
        print(build_cpp_suite())
