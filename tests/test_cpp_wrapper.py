"""The C++ mirror of cr::BVH and the batch / order-model wrappers
(include/rtq/bvh.hpp): compiles as C++17 here, and the prebuilt check
program passes on the GPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "build" / "wrapper_check"


def test_wrapper_header_compiles(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                    f"-I{ROOT / 'include'}", "-I/opt/rocm/include",
                    str(ROOT / "tests" / "cpp" / "wrapper_check.cpp")], check=True)


@pytest.mark.gpu
def test_wrapper_check_program():
    assert EXE.exists(), "built by __graft_entry__.build()"
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and "wrapper_check OK" in res.stdout
