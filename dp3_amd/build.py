"""Builds the HIP library dp3_amd/libddecal_hip.so for gfx950 with hipcc.

Usage: python -m dp3_amd.build   (or dp3_amd.build.build() from Python)
Objects go to dp3_amd/_build/ and are rebuilt when a source or header is
newer; the shared library is always relinked.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libddecal_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["solver_kernels.hip", "solver_host.hip", "predict.hip", "calibrate.hip", "mlp.hip",
           "subtract.hip", "ingest.hip"]
HEADERS = ["common.h", "solver_kernels.h", "warm.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DP3_AMD_ARCH", "gfx950")
FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    # Packed FP32 (v_pk_*_f32) code generation stays off.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _newer(src, obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + deps)


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    deps = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "ddecal_hip.h")]
    if not _newer(path, obj, deps):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=True, force=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if force:
        for s in srcs:
            o = os.path.join(BUILD, s.replace(".hip", ".o"))
            if os.path.exists(o):
                os.remove(o)
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_cpp(verbose)
    return OUT


CPP = os.path.join(HERE, "cpp")
CPP_OUT = os.path.join(HERE, "libddecal_hip_solver.so")
CXX = os.environ.get("CXX", "g++")


def build_cpp(verbose=True):
    """The C++ host layer (dp3_amd/cpp/: SolverBase-compatible
    IterativeDiagonalSolverHip over the C-ABI) as libddecal_hip_solver.so,
    linked against libddecal_hip.so (found next to it at run time)."""
    tmp = CPP_OUT + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
           os.path.join(CPP, "iterative_diagonal_solver_hip.cc"), "-L" + HERE, "-lddecal_hip",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ host layer build failed:\n{r.stderr}")
    os.replace(tmp, CPP_OUT)
    if verbose:
        print(f"built {CPP_OUT}")
    return CPP_OUT


if __name__ == "__main__":
    build(force="--force" in sys.argv)
