"""The drop-in C++ interface (include/hptt/hptt.h) on the GPU: a C++ program
using hptt::create_plan as the reference's README does, checked against the
oracle (tests/cpp/dropin_test.cpp; built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "dropin_test")


def test_cpp_dropin_interface(device):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_test is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout, r.stdout + r.stderr
