"""Build libpoggers_amd.so in-tree: hipcc, gfx950, one object per .hip source.

Run as ``python -m poggers_amd.build`` or through ``__graft_entry__.build()``.
The shared library lands next to this file and is what the ctypes binding,
the tests and bench.py load.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpoggers_amd.so")
ARCH = os.environ.get("PG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["pg_capi.hip", "pg_tcf.hip", "pg_kv.hip", "pg_util.hip", "pg_batch.hip", "pg_serve.hip"] + [
    f"pg_tcf_p{p}.hip" for p in (1, 2, 4, 8, 16, 32, 64)]

# -packed-fp32-ops: no v_pk_*_f32 generation (none of this code is float, the
# flag keeps it that way).  The host pass warns that it does not know the
# feature; that is harmless.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), int(os.environ.get('MAX_JOBS', '8')))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_cpp_tests(verbose: bool = False) -> str:
    """Compile the C++ front-end test program (tests/cpp) against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_frontend.cpp")
    out = os.path.join(root, "tests", "cpp", "test_frontend")
    cmd = [hipcc(), "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I", os.path.join(root, "include"),
           src, "-o", out, "-L", HERE, "-lpoggers_amd", "-Wl,-rpath,$ORIGIN/../../poggers_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ front-end test build failed:\n{r.stdout}\n{r.stderr}")
    ring_src = os.path.join(root, "tests", "cpp", "test_ring.cpp")
    ring_out = os.path.join(root, "tests", "cpp", "test_ring")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", ring_src, "-o", ring_out], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ring test build failed:\n{r.stdout}\n{r.stderr}")
    plan_src = os.path.join(root, "tests", "cpp", "test_plan.cpp")
    plan_out = os.path.join(root, "tests", "cpp", "test_plan")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", plan_src, "-o", plan_out], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"planner test build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {out}, {ring_out} and {plan_out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_tests(verbose=True)
    sys.exit(0)
