"""CPU checks of the C ABI boundary: the library loads and exports every declared entry point."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nlf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nlf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["nlf_ctx_create", "nlf_mesh_create", "nlf_find_neighbours", "nlf_assemble", "nlf_spmv",
                     "nlf_precond_create", "nlf_cg_solve"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "nonlocal_amd", "libnlf.so")
    if not os.path.exists(path):
        pytest.fail("libnlf.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"symbols declared in include/nlf.h but not exported: {missing}"


def test_python_binding_covers_header():
    from nonlocal_amd import _lib
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, f"no ctypes signature for: {missing}"


def test_error_path_without_gpu_calls():
    from nonlocal_amd import _lib
    L = _lib.load()
    # null handles are rejected before any device work
    assert L.nlf_mesh_destroy(None) == 0
    assert L.nlf_matrix_info(None, None, None, None, None) != 0
    assert L.nlf_cg_solve(None, None, None, None, 1e-10, 10, 0, None) != 0
