"""The C++ facade (include/golap/golap.hpp) driven like the reference's plan code, in its own
process (tests/cpp/facade_q1.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_facade_q1_and_select():
    exe = os.path.join(HERE, "cpp", "facade_q1")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK")
