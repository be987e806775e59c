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


@pytest.mark.gpu
def test_host_operator_layer_gpu():
    run("gpu")
