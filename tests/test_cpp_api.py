"""Runs the C++ restatement of the reference's unit tests (tests/cpp)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_fluid_api")


def _run(arg):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    r = subprocess.run([BIN] + arg, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_reference_unit_tests_cpu():
    _run(["--cpu"])


@pytest.mark.gpu
def test_cpp_reference_unit_tests_gpu():
    _run([])
