"""Build recipe for libhmsbeagle_amd.so (gfx950).

hipcc compiles the kernels for gfx950 only and links them with the host
front end into one shared library that exports exactly the beagle.h entry
points (everything else has hidden visibility and binds locally).
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhmsbeagle_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-fvisibility=hidden", "--offload-arch=gfx950",
          "-I" + os.path.join(REPO, "include"), "-I" + CSRC]
# packed FP32 instructions (v_pk_*_f32) stay off: the compiler would form them
# from ordinary float code, and this library's kernels do not rely on them
DEVICE = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["kernels.hip", "nodes.hip", "mfma.hip", "noalloc.hip", "instance.cpp", "checkpoint.cpp",
           "frontend.cpp"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src),
                                                 "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]), flush=True)
    # one compiler process per source (the two kernel files dominate)
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as pool:
        list(pool.map(run, cmds))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,-Bsymbolic", "-Wl,--no-undefined"]
    run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
