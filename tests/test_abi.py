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


def declared_arity(header):
    """Parameter count of every ddh_ function declared in `header`."""
    text = open(header).read()
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(ddh_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_python_binding_matches_declared_arity():
    """dp3_amd.hip's ctypes signatures take as many arguments as the C
    declarations (a mismatch only shows at call time on a GPU)."""
    from dp3_amd import hip as H
    arity = declared_arity(os.path.join(ROOT, "include", "ddecal_hip.h"))
    wrong = {name: (len(args), arity[name]) for name, (args, _) in H.SIGNATURES.items()
             if name in arity and len(args) != arity[name]}
    assert not wrong, f"ctypes argument count differs from the header: {wrong}"
    unknown = [n for n in H.SIGNATURES if n not in arity]
    assert not unknown, f"bound but not declared: {unknown}"
