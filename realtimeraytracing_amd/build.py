"""Builds librtq.so (the gfx950 HIP library behind include/rtq/rtq.h) in-tree.

Each .hip file is compiled to an object in parallel, then linked into
realtimeraytracing_amd/librtq.so.  Flags that matter for results:
  -ffp-contract=off                         no implicit FMA: float expressions
                                            round exactly as written
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division / sqrt (HIP default,
                                            stated explicitly)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "librtq.so"
SOURCES = ["api.hip", "batch.hip", "build.hip", "learned.hip", "obj.cpp", "primitives.hip", "shade.hip",
           "trace.hip", "wide.hip"]
ARCH = os.environ.get("RTQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-fvisibility=hidden",
        "-ffp-contract=off",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{ROOT / 'include'}",
    ]


def host_flags() -> list[str]:
    """Host-only sources (.cpp): same language level and FP rules, no device code."""
    return ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off", "-Wall", f"-I{ROOT / 'include'}"]


def _compile(src: str) -> Path:
    obj = OBJ / (Path(src).stem + ".o")
    flags = host_flags() if src.endswith(".cpp") else compile_flags()
    cmd = [hipcc(), *flags, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build_library(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


PROBE_SRC = CSRC / "probe" / "vmem_rate.hip"
PROBE_LIB = PKG / "librtqprobe.so"


def build_probe(verbose: bool = False) -> Path:
    """The vector-memory rate probe bench.py measures the trace kernels'
    ceiling with (its own small library; librtq does not link it)."""
    tmp = PROBE_LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *compile_flags(), "-shared", "-o", str(tmp), str(PROBE_SRC)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"probe build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, PROBE_LIB)
    if verbose:
        print(f"built {PROBE_LIB}")
    return PROBE_LIB


def build_oracle(verbose: bool = False) -> Path:
    """The CPU checker (test infrastructure, never loaded by the library)."""
    res = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    out = ROOT / "oracle" / "build" / "liboracle.so"
    if verbose:
        print(f"built {out}")
    return out


def build_cpp_check(verbose: bool = False) -> Path:
    """The C++ wrapper check program run by the GPU tests (tests/cpp)."""
    src = ROOT / "tests" / "cpp" / "wrapper_check.cpp"
    out_dir = ROOT / "tests" / "cpp" / "build"
    out_dir.mkdir(exist_ok=True)
    exe = out_dir / "wrapper_check"
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
           f"-I{rocm / 'include'}", str(src), "-o", str(exe),
           f"-L{PKG}", "-lrtq", f"-L{rocm / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{PKG}",
           "-Wl,-rpath,$ORIGIN/../../../realtimeraytracing_amd", f"-Wl,-rpath,{rocm / 'lib'}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"wrapper_check build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build_library(verbose=True)
    build_probe(verbose=True)
    if "--with-oracle" in sys.argv:
        build_oracle(verbose=True)
        build_cpp_check(verbose=True)
