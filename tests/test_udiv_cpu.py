"""The run-time-divisor division used for computed group keys (golap_public_amd/csrc/udiv.hpp,
Granlund-Montgomery) against the hardware division, compiled for the host from the same header
the kernels include."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_udiv32_matches_hardware_division(tmp_path):
    exe = tmp_path / "udiv_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", str(exe), os.path.join(HERE, "cpp", "udiv_check.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "checks passed" in out.stdout
