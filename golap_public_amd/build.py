"""Builds the gfx950 library golap_public_amd/libgolap_hip.so in-tree with hipcc.

Usage: python -m golap_public_amd.build  (or build() from Python).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgolap_hip.so")

HIP_SOURCES = ["kernels.hip", "star.hip", "scan.hip", "keys.hip", "chunkmeta.hip", "capi_scan.cpp", "capi.cpp", "capi_star.cpp", "capi_batch.cpp", "capi_server.cpp", "encode.cpp", "synth.cpp", "storage.cpp", "outofcore.cpp", "encode_dev.hip", "host_join.cpp"]

# -packed-fp32-ops: no v_pk_*_f32 generation (nothing on this path is floating point, the flag
# keeps it that way if code is added).
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compile every source for gfx950 and link the shared library (always recompiles)."""
    out_tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-shared", "-o", out_tmp, "-x", "hip"]
    cmd += [os.path.join(CSRC, s) for s in HIP_SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host compilation pass does not know the gfx950 feature flag and says so; harmless
    lines = [l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l]
    if lines:
        print("\n".join(lines), flush=True)
    if proc.returncode != 0:
        raise subprocess.CalledProcessError(proc.returncode, cmd)
    os.replace(out_tmp, LIB)
    return LIB


def build_cpp_tests(verbose=False):
    """The C++ facade test program (host code only: the facade calls the C ABI)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_q1.cpp")
    exe = os.path.join(root, "tests", "cpp", "facade_q1")
    cmd = ["g++", "-std=c++17", "-O2", "-o", exe, src, "-I" + os.path.join(root, "include"),
           "-I/opt/rocm/include", "-L" + HERE, "-lgolap_hip", "-L" + os.path.join(root, "oracle"), "-loracle",
           "-Wl,-rpath," + HERE, "-Wl,-rpath," + os.path.join(root, "oracle"), "-pthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return exe


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
