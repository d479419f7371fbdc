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

HIP_SOURCES = ["kernels.hip", "decode.hip", "q1wg.hip", "star.hip", "scan.hip", "keys.hip", "chunkmeta.hip", "capi_scan.cpp", "capi.cpp", "capi_star.cpp", "capi_batch.cpp", "capi_server.cpp", "encode.cpp", "synth.cpp", "storage.cpp", "outofcore.cpp", "encode_dev.hip", "host_join.cpp", "capi_noalloc.hip", "capi_stream.cpp"]

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


def _compile(src, obj, verbose):
    cmd = [hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj, "-x", "hip", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host compilation pass does not know the gfx950 feature flag and says so; harmless
    lines = [l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l]
    return proc.returncode, cmd, "\n".join(lines)


def build(verbose=False, jobs=None):
    """Compile every source for gfx950 (one object per source, in parallel) and link the shared
    library. Always recompiles."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(os.path.dirname(HERE), "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or min(len(HIP_SOURCES), max(1, min(16, os.cpu_count() or 1)))
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda so: _compile(os.path.join(CSRC, so[0]), so[1], verbose),
                              zip(HIP_SOURCES, objs)))
    failed = None
    for rc, cmd, out in results:
        if out:
            print(out, flush=True)
        if rc != 0 and failed is None:
            failed = (rc, cmd)
    if failed:
        raise subprocess.CalledProcessError(*failed)
    out_tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out_tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
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
