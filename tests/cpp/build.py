"""Build recipe for the C++ operator-layer test program (test infrastructure):
tests/cpp/build/test_exec from test_exec.cpp, linked against the product's
libwadjet_exec.so and libwadjet_gpu.so (built first by wadjet_amd/build.py)
through a relative rpath, so the tree can move as a whole."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIBDIR = ROOT / "wadjet_amd" / "lib"
BIN = HERE / "build" / "test_exec"


def build() -> Path:
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
