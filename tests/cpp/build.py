"""Build recipe for the C++ test programs (test infrastructure), linked
against the product's libwadjet_exec.so / libwadjet_gpu.so (built first by
wadjet_amd/build.py) through a relative rpath, so the tree can move as a whole:
  build/test_exec        the operator-layer tests (test_exec.cpp)
  build/test_exec_asan   the same with the operator layer compiled in under
                         AddressSanitizer + UBSan (its `cpu` mode: host code)
  build/test_exec_tsan   the same under ThreadSanitizer
  build/pq_fuzz          the Parquet staging fuzz driver, the staging source
                         compiled in under AddressSanitizer + UBSan."""
from __future__ import annotations

import concurrent.futures as cf
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIBDIR = ROOT / "wadjet_amd" / "lib"
OUT = HERE / "build"
BIN = OUT / "test_exec"
FUZZ = OUT / "pq_fuzz"
RPATH = "-Wl,-rpath,$ORIGIN/../../../wadjet_amd/lib"
BASE = ["g++", "-std=c++17", "-pthread"]
ASAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
TSAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread"]
EXEC = str(ROOT / "wadjet_amd" / "host" / "exec.cpp")
TEST = str(HERE / "test_exec.cpp")

COMMANDS = {
    BIN: BASE + ["-O2", "-Wall", "-Wextra", TEST, f"-L{LIBDIR}", "-lwadjet_exec", "-lwadjet_gpu", RPATH],
    OUT / "test_exec_asan": BASE + ASAN + [TEST, EXEC, f"-L{LIBDIR}", "-lwadjet_gpu", RPATH],
    OUT / "test_exec_tsan": BASE + TSAN + [TEST, EXEC, f"-L{LIBDIR}", "-lwadjet_gpu", RPATH],
    FUZZ: BASE + ASAN + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(HERE / "pq_fuzz.cpp"),
                         str(ROOT / "wadjet_amd" / "csrc" / "wg_pq_stage.cpp"), f"-L{LIBDIR}", "-lwadjet_gpu",
                         "-L/opt/rocm/lib", "-lamdhip64", "-lz", "-ldl", RPATH, "-Wl,-rpath,/opt/rocm/lib"],
}


def _build(item) -> None:
    out, cmd = item
    r = subprocess.run(cmd + ["-o", str(out)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{out.name} build failed:\n{r.stderr}")


def build() -> Path:
    OUT.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_build, COMMANDS.items()))
    return BIN


if __name__ == "__main__":
    build()
