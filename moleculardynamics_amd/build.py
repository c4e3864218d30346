"""Build the gfx950 engine library (libmdx.so) and the oracle, in-tree.

The library is compiled from moleculardynamics_amd/csrc/*.hip with hipcc for
gfx950 only and linked into moleculardynamics_amd/libmdx.so; the host tools
(the cpumd-compatible driver) are built from moleculardynamics_amd/host/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libmdx.so"
HOSTLIB = PKG / "libmdx_host.so"
REFLIB = PKG / "libmdx_ref.so"       # the reference's free functions over the C-ABI
REFABI = PKG / "refabi"
CPUMD = PKG / "mdx_cpumd"
ORACLE_DIR = ROOT / "oracle"

ARCH = "gfx950"
# -ffp-contract=off: a*b+c stays two roundings, so the FP64 kernels reproduce
# the oracle's arithmetic; the FP32 paths request fused multiply-adds
# explicitly.  Packed FP32 instructions are disabled for all device code.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return h


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")


def _compile_one(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    deps = [src, CSRC / "mdx_internal.h", ROOT / "include" / "mdx.h"]
    deps += list(CSRC.glob("*.h"))
    _run([hipcc(), *HIPFLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_library(jobs: int | None = None) -> Path:
    """Compile every .hip source for gfx950 and link libmdx.so (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> None:
    """The C++ host layer (config parser, loader, results, cpumd driver)."""
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return
    lib_srcs = [s for s in srcs if s.name != "cpumd_main.cpp"]
    common = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{ROOT / 'include'}", f"-I{HOST}"]
    tmp = HOSTLIB.with_suffix(".so.tmp")
    _run(["g++", *common, "-shared", "-o", str(tmp), *map(str, lib_srcs),
          f"-L{PKG}", "-lmdx", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOSTLIB)
    main = HOST / "cpumd_main.cpp"
    if main.exists():
        tmpx = CPUMD.with_suffix(".tmp")
        _run(["g++", *common, "-o", str(tmpx), str(main), f"-L{PKG}", "-lmdx_host", "-lmdx",
              f"-Wl,-rpath,$ORIGIN"])
        os.replace(tmpx, CPUMD)


def build_refabi() -> None:
    """libmdx_ref.so: predict/correct/lennardJonesForces/... with the
    reference's C++ signatures (include/mdx_reference.h)."""
    srcs = sorted(REFABI.glob("*.cpp"))
    tmp = REFLIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}", "-shared", "-o", str(tmp),
          *map(str, srcs), f"-L{PKG}", "-lmdx", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, REFLIB)


def build_test_tools() -> None:
    """Test drivers under tests/ (they link the oracle: test infrastructure)."""
    out = ROOT / "tests" / "_bin"
    out.mkdir(exist_ok=True)
    exe = out / "ref_adapter_harness"
    tmp = exe.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", str(ROOT / "tests" / "ref_adapter_harness.cpp"), "-o", str(tmp),
          f"-L{PKG}", "-lmdx_ref", "-lmdx", f"-L{ORACLE_DIR}", "-loracle",
          "-Wl,-rpath,$ORIGIN/../../moleculardynamics_amd", "-Wl,-rpath,$ORIGIN/../../oracle"])
    os.replace(tmp, exe)


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE_DIR), "-B"])
    return ORACLE_DIR / "liboracle.so"


def build_all() -> None:
    build_library()
    build_host()
    build_refabi()
    build_oracle()
    build_test_tools()


if __name__ == "__main__":
    build_all()
    print("built", LIB, HOSTLIB if HOSTLIB.exists() else "", ORACLE_DIR / "liboracle.so")
