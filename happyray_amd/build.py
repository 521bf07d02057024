"""Compile the gfx950 HIP library (libhappyray.so) in-tree.

    python -m happyray_amd.build          # or happyray_amd.build.build()

Every translation unit is built with -ffp-contract=off (no fused
multiply-add, so kernel arithmetic equals the IEEE result of the same
expression on the CPU) and with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhappyray.so")
CLI = os.path.join(HERE, "happyray")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["hr_sort.hip", "hr_build.hip", "hr_hier.hip", "hr_trace.hip", "hr_pathmk.hip", "hr_persist.hip", "hr_engine.hip",
           "hr_batch.hip", "hr_server.hip"]
CXX_SOURCES = ["hr_scene.cpp", "hr_image.cpp"]   # host-only C++ (scene files, images)
ARCH = os.environ.get("HR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-Werror=uninitialized", "-Werror=sometimes-uninitialized",
    "-I", INCLUDE, "-I", CSRC,
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Werror=uninitialized", "-I", INCLUDE,
             "-I", CSRC]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++"), "-c", os.path.join(CSRC, src), "-o", out] + CXX_FLAGS
    else:
        cmd = [_hipcc(), "-c", "-x", "hip", os.path.join(CSRC, src), "-o", out] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(_compile, SOURCES + CXX_SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the headless renderer over the C ABI (Main.cpp's -nw path)
    cli_tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), os.path.join(CSRC, "hr_cli.cpp"), "-o", cli_tmp] + CXX_FLAGS + \
          ["-L", HERE, "-lhappyray", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the happyray executable failed:\n{r.stdout}\n{r.stderr}")
    os.replace(cli_tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
