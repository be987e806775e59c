"""Runs the C++ operator-layer tests (tests/cpp/test_exec.cpp): the host-only
pieces on the CPU, the GPU operators (GpuHashAggregate, GpuHashJoin with the
bounded probe, Workers > 1 with MergeSink, filter pushdown) on the device."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "build" / "test_exec"


def run(mode):
    assert BIN.exists(), f"{BIN} missing: run __graft_entry__.build()"
    r = subprocess.run([str(BIN), mode], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def test_host_operator_layer_cpu():
    run("cpu")


@pytest.mark.parametrize("sanitizer", ["asan", "tsan"])
def test_host_operator_layer_under_sanitizers(sanitizer):
    """The `cpu` tests (KernelFilter, Pipeline with 4 workers, views) with the
    operator layer built under AddressSanitizer + UBSan, and under
    ThreadSanitizer."""
    b = BIN.parent / f"test_exec_{sanitizer}"
    assert b.exists(), f"{b} missing: run __graft_entry__.build()"
    r = subprocess.run([str(b), "cpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_host_operator_layer_gpu():
    run("gpu")
