"""The C++ mirror of the reference API (nonlocal_amd/include/nonlocal/nonlocal_amd.hpp): it compiles
against include/nlf.h and links libnlf.so (CPU); its heat solver driver reproduces the oracle's solution
and flux, and its elastic driver meets saint_venant.cpp's bars (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from nonlocal_amd import configs
from tests.helpers import direct_solution, oracle_assembly, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _compile(out, source="heat_q4.cpp"):
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT}/include",
           f"-I{ROOT}/nonlocal_amd/include", "-isystem", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(CPP, source), "-o", out, f"-L{ROOT}/nonlocal_amd", "-lnlf", "-L/opt/rocm/lib",
           "-lamdhip64", f"-Wl,-rpath,{ROOT}/nonlocal_amd", "-Wl,-rpath,/opt/rocm/lib"]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.parametrize("source", ["heat_q4.cpp", "saint_venant.cpp", "transient_q4.cpp"])
def test_mirror_compiles_and_links(tmp_path, source):
    res = _compile(str(tmp_path / "prog"), source)
    assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_cpp_heat_solver_matches_oracle(tmp_path):
    exe = os.path.join(CPP, "heat_q4")
    if not os.path.exists(exe):
        res = _compile(exe)
        assert res.returncode == 0, res.stderr
    out = str(tmp_path / "T.bin")
    print("[cpp] heat_q4 40 0.5 0.1", flush=True)
    res = subprocess.run([exe, "40", "0.5", "0.1", out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)
    T, fx, fy = np.fromfile(out, dtype=np.float64).reshape(3, -1)
    p = configs.cfg0()
    om, K, f, xD, inner = oracle_assembly(p)
    assert rel_l2(T, direct_solution(K, f)) < 1e-10
    # the solver computed the flux (calc_flux, as the reference's solver does) of its temperature
    n_ref, _ = om.heat_flux(p.models, p.materials, T)
    assert np.abs(np.stack([fx, fy]) - n_ref).max() <= 1e-12 * np.abs(n_ref).max()


@pytest.mark.gpu
def test_cpp_saint_venant_known_answers():
    """saint_venant.cpp's centre-line checks through the C++ mirror (equilibrium_equation with
    tractions given as functions of the point, then calc_strain_and_stress)."""
    exe = os.path.join(CPP, "saint_venant")
    if not os.path.exists(exe):
        res = _compile(exe, "saint_venant.cpp")
        assert res.returncode == 0, res.stderr
    print("[cpp] saint_venant", flush=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)
    fields = res.stdout.split()
    ux = float(fields[fields.index("ux_centre") + 1])
    sxx = float(fields[fields.index("sxx_centre") + 1])
    assert ux <= 2e-14 and sxx <= 1.3e-9


@pytest.mark.gpu
def test_cpp_transient_matches_oracle(tmp_path):
    """nonstationary_heat_equation_solver_2d of the C++ mirror: 5 steps against the oracle's implicit
    Euler on the same problem."""
    from nonlocal_amd import _abi as A
    from nonlocal_amd import meshgen
    from oracle import oracle as O
    exe = os.path.join(CPP, "transient_q4")
    if not os.path.exists(exe):
        res = _compile(exe, "transient_q4.cpp")
        assert res.returncode == 0, res.stderr
    out = str(tmp_path / "T.bin")
    print("[cpp] transient_q4 16 0.5 0.2 0.01 5", flush=True)
    res = subprocess.run([exe, "16", "0.5", "0.2", "0.01", "5", out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)
    T = np.fromfile(out, dtype=np.float64)
    mesh = meshgen.structured_q4(16)
    bcs = [A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(A.NLF_BC_FIRST_KIND, 0.0), A.boundary(), A.boundary()]
    ref = O.transient(mesh, [configs.bell(0.5, 0.2)], [A.thermal_material()], [2.0], bcs, 0.01,
                      np.sin(np.pi * mesh.xy[:, 0]), 5, source=lambda x, y: 1.0 + y)
    assert rel_l2(T, ref[-1]) < 1e-9
