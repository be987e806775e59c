"""Build recipe for the C++ test programs (test infrastructure):
tests/cpp/build/test_exec from test_exec.cpp, linked against the product's
libwadjet_exec.so and libwadjet_gpu.so (built first by wadjet_amd/build.py)
through a relative rpath, so the tree can move as a whole; and
tests/cpp/build/pq_fuzz, the Parquet staging fuzz driver under sanitizers."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIBDIR = ROOT / "wadjet_amd" / "lib"
BIN = HERE / "build" / "test_exec"


FUZZ = HERE / "build" / "pq_fuzz"


def build_fuzz() -> Path:
    """The Parquet staging fuzz driver, with the staging source compiled into
    it under AddressSanitizer and UBSan (host code only)."""
    FUZZ.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(HERE / "pq_fuzz.cpp"), str(ROOT / "wadjet_amd" / "csrc" / "wg_pq_stage.cpp"),
           f"-L{LIBDIR}", "-lwadjet_gpu", "-L/opt/rocm/lib", "-lamdhip64", "-lz", "-ldl",
           "-Wl,-rpath,$ORIGIN/../../../wadjet_amd/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", str(FUZZ)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pq_fuzz build failed:\n{r.stderr}")
    return FUZZ


def build() -> Path:
    build_fuzz()
    BIN.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-pthread", str(HERE / "test_exec.cpp"),
           f"-L{LIBDIR}", "-lwadjet_exec", "-lwadjet_gpu", "-Wl,-rpath,$ORIGIN/../../../wadjet_amd/lib",
           "-o", str(BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"test_exec build failed:\n{r.stderr}")
    return BIN


if __name__ == "__main__":
    build()
