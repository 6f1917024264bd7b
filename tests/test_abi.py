"""The C-ABI boundary without a GPU: every function declared in include/*.h
(and in the oracle's header) is exported by the built shared library, and the
library loads. No function is called, so no GPU is needed."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"\b((?:ddh|orc)_[a-z0-9_]+)\s*\(")


def declared(header):
    text = open(header).read()
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


def exported(lib_path, names):
    lib = ctypes.CDLL(lib_path)
    return [n for n in names if not hasattr(lib, n)]


@pytest.mark.parametrize("header", sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
def test_library_exports_every_declared_entry_point(header):
    names = declared(header)
    assert len(names) > 10, names
    lib_path = os.path.join(ROOT, "dp3_amd", "libddecal_hip.so")
    if not os.path.exists(lib_path):
        pytest.fail("dp3_amd/libddecal_hip.so is not built (python -m dp3_amd.build)")
    missing = exported(lib_path, names)
    assert not missing, f"declared in {os.path.basename(header)} but not exported: {missing}"


def test_oracle_exports_every_declared_function():
    names = declared(os.path.join(ROOT, "oracle", "oracle.h"))
    missing = exported(os.path.join(ROOT, "oracle", "liboracle.so"), names)
    assert not missing, missing


def test_cpp_host_layer_links():
    """The C++ host layer and its test binary resolve every shared library
    they need (nothing is run)."""
    import subprocess
    for path in (os.path.join(ROOT, "dp3_amd", "libddecal_hip_solver.so"),
                 os.path.join(ROOT, "tests", "cpp", "test_iterative_diagonal_solver_hip")):
        if not os.path.exists(path):
            pytest.fail(f"{path} is not built (python __graft_entry__.py)")
        out = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
        assert "not found" not in out, out
    lib = ctypes.CDLL(os.path.join(ROOT, "dp3_amd", "libddecal_hip_solver.so"))
    assert lib is not None
