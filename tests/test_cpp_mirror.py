"""The C++ mirror of the reference API (nonlocal_amd/include/nonlocal/nonlocal_amd.hpp): its drivers
compile against include/nlf.h and link libnlf.so, and its SU2 reader gives the Python reader's arrays
(CPU); written the way the reference's tests are, its heat and elastic drivers meet the reference's
bars (isotropic_solid_ring.cpp, saint_venant.cpp) and its solvers reproduce the oracle (GPU)."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from nonlocal_amd import _abi as A
from nonlocal_amd import configs, su2
from tests.helpers import MESH_DIR, direct_solution, oracle_assembly, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
SOURCES = ["heat_q4.cpp", "saint_venant.cpp", "transient_q4.cpp", "isotropic_solid_ring.cpp", "su2_dump.cpp"]


def _compile(out, source="heat_q4.cpp"):
    cmd = ["g++", "-O1", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-Wno-missing-field-initializers",
           f"-I{ROOT}/include", f"-I{ROOT}/nonlocal_amd/include", "-isystem", "/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", f'-DNLF_MESH_DIR="{MESH_DIR}"',
           os.path.join(CPP, source), "-o", out, f"-L{ROOT}/nonlocal_amd", "-lnlf", "-L/opt/rocm/lib",
           "-lamdhip64", f"-Wl,-rpath,{ROOT}/nonlocal_amd", "-Wl,-rpath,/opt/rocm/lib"]
    return subprocess.run(cmd, capture_output=True, text=True)


def _program(name):
    """The driver built by the build step, or compiled here when it is missing."""
    exe = os.path.join(CPP, name)
    if not os.path.exists(exe):
        res = _compile(exe, name + ".cpp")
        assert res.returncode == 0, res.stderr
    return exe


@pytest.mark.parametrize("source", SOURCES)
def test_mirror_compiles_and_links(tmp_path, source):
    res = _compile(str(tmp_path / "prog"), source)
    assert res.returncode == 0, res.stderr


@pytest.mark.parametrize("name", ["solid_ring", "composite_ring", "plate_10x1_h0.125", "square_1x1_16el_uniform_mesh"])
def test_cpp_su2_reader_matches_python(tmp_path, name):
    """mesh::read_su2 of the mirror (su2_parser.hpp:165-195) against nonlocal_amd.su2 on the reference's
    meshes: nodes, elements in the reference local order, groups, boundary edges, names."""
    exe = _program("su2_dump")
    out = str(tmp_path / "mesh.bin")
    res = subprocess.run([exe, os.path.join(MESH_DIR, name + ".su2"), out], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    m = su2.read(os.path.join(MESH_DIR, name + ".su2"))
    raw = open(out, "rb").read()
    N, E, nn, G, B, nb, GB = np.frombuffer(raw[:56], dtype=np.int64)
    pos = 56

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + count * np.dtype(dtype).itemsize], dtype=dtype)
        pos += count * np.dtype(dtype).itemsize
        return a
    xy = take(np.float64, 2 * N).reshape(N, 2)
    elems = take(np.int32, E * nn).reshape(E, nn)
    gbeg = take(np.int64, G + 1)
    bedges = take(np.int32, B * nb).reshape(B, nb)
    bbeg = take(np.int64, GB + 1)
    names = [ln.split(" ", 1) for ln in res.stdout.splitlines()]
    assert np.array_equal(xy, m.xy) and nn == m.nn
    assert np.array_equal(elems, m.elems) and np.array_equal(gbeg, m.group_begin)
    assert np.array_equal(bedges, m.bedges) and np.array_equal(bbeg, m.bgroup_begin)
    assert [n for k, n in names if k == "group"] == list(m.group_names)
    assert [n for k, n in names if k == "boundary"] == list(m.bgroup_names)


CPP_DISTANCES = {"l2": dict(n=2), "n3": dict(n=3), "n1.5": dict(n=1.5), "inf": dict(n="inf"),
                 "rotation": dict(distance=A.NLF_DISTANCE_ELLIPSE_WITH_ROTATION)}


@pytest.mark.gpu
@pytest.mark.parametrize("distance", list(CPP_DISTANCES))
def test_cpp_heat_solver_matches_oracle(tmp_path, distance):
    exe = _program("heat_q4")
    out = str(tmp_path / "T.bin")
    print(f"[cpp] heat_q4 40 0.5 0.1 {distance}", flush=True)
    res = subprocess.run([exe, "40", "0.5", "0.1", out, distance], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)
    T, fx, fy = np.fromfile(out, dtype=np.float64).reshape(3, -1)
    p = configs.cfg0()
    if distance == "rotation":
        p.mesh = dataclasses.replace(p.mesh, xy=p.mesh.xy - 0.5)
        p.models = [A.model(0.5, (0.1, 0.7 * 0.1), p=2.0, q=2.0, **CPP_DISTANCES[distance])]
    else:
        p.models = [A.model(0.5, 0.1, p=2.0, q=2.0, **CPP_DISTANCES[distance])]
    om, K, f, xD, inner = oracle_assembly(p)
    assert rel_l2(T, direct_solution(K, f)) < 1e-10
    # the solver computed the flux (calc_flux, as the reference's solver does) of its temperature
    n_ref, _ = om.heat_flux(p.models, p.materials, T)
    assert np.abs(np.stack([fx, fy]) - n_ref).max() <= 1e-12 * np.abs(n_ref).max()


def _run_checks(name):
    """A driver written like one of the reference's tests: it prints PASS/FAIL per check and exits 0
    when all pass."""
    exe = _program(name)
    print(f"[cpp] {name}", flush=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout.strip(), flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


@pytest.mark.gpu
def test_cpp_saint_venant_known_answers():
    """saint_venant.cpp of the reference (SU2 plate, raw_mechanical_parameters, displacement_2d /
    pressure_2d, equilibrium_equation<I>) through the mirror, at its four bars."""
    out = _run_checks("saint_venant")
    assert out.count("PASS") == 4


@pytest.mark.gpu
def test_cpp_isotropic_solid_ring_known_answers():
    """isotropic_solid_ring.cpp of the reference (mesh_2d from an SU2 stream, parameters_2d with
    designated initializers, temperature_2d, stationary_heat_equation_solver_2d<I>(..., {})) through the
    mirror, at its bars: temperature 1.4e-3, flux components 1.7e-2."""
    out = _run_checks("isotropic_solid_ring")
    assert out.count("PASS") == 3


@pytest.mark.gpu
def test_cpp_transient_matches_oracle(tmp_path):
    """nonstationary_heat_equation_solver_2d of the C++ mirror: 5 steps against the oracle's implicit
    Euler on the same problem."""
    from nonlocal_amd import _abi as A
    from nonlocal_amd import meshgen
    from oracle import oracle as O
    exe = _program("transient_q4")
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
