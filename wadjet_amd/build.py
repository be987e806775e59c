"""Build recipe for the HIP library (gfx950): wadjet_amd/lib/libwadjet_gpu.so.

Each ``csrc/*.hip`` translation unit is compiled with hipcc and each host-only
``csrc/*.cpp`` one with the host compiler, in parallel, then linked into one
shared library that exports the C-ABI declared in
``include/wadjet_gpu.h`` and ``include/wadjet_gen.h``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libwadjet_gpu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions in any kernel.
# -munsafe-fp-atomics: float64 atomic add maps to the hardware instruction
#  (device-local coarse-grained memory only, which is all this library uses).
DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# host-only sources: the HIP runtime API from its C headers, no device code
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
              "-I/opt/rocm/include"]
HOST_LIBS = ["-lz", "-ldl", "-pthread"]


def _compile(src: Path, obj: Path) -> None:
    if src.suffix == ".cpp":
        cmd = ["g++", *HOST_FLAGS, f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *DEVICE_FLAGS, f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed for {src.name}:\n{r.stderr}")


def build(verbose: bool = True) -> Path:
    """Compile every HIP source and link the shared library. Always rebuilds."""
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = [objdir / (s.name.replace(".", "_") + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(lambda p: _compile(*p), zip(srcs, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *HOST_LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_layer()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


EXEC_LIB = LIBDIR / "libwadjet_exec.so"


def build_host_layer() -> None:
    """The C++ operator layer (host/exec.cpp) as libwadjet_exec.so, linked
    against the library with an $ORIGIN rpath so the pair travels together."""
    tmp = EXEC_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-pthread", "-shared", "-fPIC",
           str(PKG / "host" / "exec.cpp"), f"-L{LIBDIR}", "-lwadjet_gpu", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host layer build failed:\n{r.stderr}")
    os.replace(tmp, EXEC_LIB)


if __name__ == "__main__":
    build()
