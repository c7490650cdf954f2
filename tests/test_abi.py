"""CPU checks of the C-ABI boundary: libmpchol.so loads and exports every
function declared in include/mpchol.h (no compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpchol.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["mpc_handle_create", "mpc_factor", "mpc_pcg", "mpc_precond_cg", "mpc_matgen",
                 "mpc_quantize_panel", "mpc_syrk", "mpc_rbf_assemble"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from cholesky_gpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built (python -m cholesky_gpu_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (mpc_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_python_binding_signatures_cover_header():
    from cholesky_gpu_amd import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-fsyntax-only", "-x", "c", "-std=c99", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_error_string_without_gpu():
    from cholesky_gpu_amd import _lib
    lib = _lib.load()
    assert lib.mpc_version_string().startswith(b"mpchol")
    assert isinstance(lib.mpc_last_error_string(), bytes)
