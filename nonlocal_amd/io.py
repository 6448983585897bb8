"""Result files: legacy-VTK and CSV writers of NonLocFEM (libraries/mesh/mesh_2d/
mesh_container_2d_utils.hpp:134-240) and the field sets the CLI writes (NonLocFEM/save_results.hpp:
10-71).  Host-side text I/O of device results (they are copied back first).

* save_as_vtk: "# vtk DataFile Version 4.2" unstructured grid, points with z = 0, one cell per 2D
  element -- serendipity quads written corners first (local order 0 2 4 6 1 3 5 7, the inverse of
  the SU2 reader's remap) -- and cell types 9 (bilinear) / 23 (serendipity);
* POINT_DATA blocks: SCALARS (temperature), VECTORS (flux, displacement; z = 0) and TENSORS
  (strain, stress as [[xx, xy, 0], [xy, yy, 0], [0, 0, 0]]);
* save_as_csv: header "x,y,<names>" and one row per node, max_digits10 (17) significant digits
  unless a precision is given.
"""
from __future__ import annotations

import numpy as np

_VTK_CELL = {4: 9, 8: 23}
_VTK_ORDER = {4: (0, 1, 2, 3), 8: (0, 2, 4, 6, 1, 3, 5, 7)}


def _fmt(v, precision):
    return repr(float(v)) if precision is None else f"{float(v):.{precision}g}"


def _host(v):
    if hasattr(v, "detach"):
        v = v.detach().cpu().numpy()
    return np.asarray(v, dtype=np.float64).ravel()


def save_as_vtk(path, mesh, precision=None, scalars=(), vectors=(), tensors=()):
    """mesh: a Mesh2D.  scalars: [(name, values[N])]; vectors: [(name, (x[N], y[N]))]; tensors:
    [(name, (xx[N], yy[N], xy[N]))].  Point data are written after the grid under POINT_DATA."""
    n = mesh.n_nodes
    lines = ["# vtk DataFile Version 4.2", "Data", "ASCII", "DATASET UNSTRUCTURED_GRID",
             f"POINTS {n} double"]
    lines += [f"{_fmt(x, precision)} {_fmt(y, precision)} 0" for x, y in mesh.xy]
    order = _VTK_ORDER[mesh.nn]
    lines.append(f"CELLS {mesh.n_elems} {mesh.n_elems * (mesh.nn + 1)}")
    lines += [f"{mesh.nn} " + " ".join(str(int(e[k])) for k in order) for e in mesh.elems]
    lines.append(f"CELL_TYPES {mesh.n_elems}")
    lines += [str(_VTK_CELL[mesh.nn])] * mesh.n_elems
    if scalars or vectors or tensors:
        lines.append(f"POINT_DATA {n}")
    for name, v in scalars:                               # save_scalars_to_vtk
        v = _host(v)
        lines += [f"SCALARS {name} double 1", "LOOKUP_TABLE default"] + [_fmt(a, precision) for a in v]
    for name, (vx, vy) in vectors:                        # save_vectors_to_vtk
        vx, vy = _host(vx), _host(vy)
        lines.append(f"VECTORS {name} double")
        lines += [f"{_fmt(a, precision)} {_fmt(b, precision)} 0" for a, b in zip(vx, vy)]
    for name, (txx, tyy, txy) in tensors:                 # save_tensors_to_vtk
        txx, tyy, txy = _host(txx), _host(tyy), _host(txy)
        lines.append(f"TENSORS {name} double")
        for a, b, c in zip(txx, tyy, txy):
            lines += [f"{_fmt(a, precision)} {_fmt(c, precision)} 0", f"{_fmt(c, precision)} {_fmt(b, precision)} 0",
                      "0 0 0", ""]
    with open(path, "w", encoding="ascii") as f:
        f.write("\n".join(lines) + "\n")


def save_as_csv(path, mesh, data, precision=None):
    """data: [(name, values[N])]; raises ValueError when a vector's length is not the node count
    (the reference's logic_error)."""
    cols = []
    for name, v in data:
        v = _host(v)
        if v.size != mesh.n_nodes:
            raise ValueError(f'The result cannot be saved because the mesh nodes number and elements in the vector '
                             f'"{name}" do not match.')
        cols.append(v)
    digits = 17 if precision is None else int(precision)
    with open(path, "w", encoding="ascii") as f:
        f.write(",".join(["x", "y"] + [name for name, _ in data]) + "\n")
        for i in range(mesh.n_nodes):
            row = [mesh.xy[i, 0], mesh.xy[i, 1]] + [c[i] for c in cols]
            f.write(",".join(f"{float(v):.{digits}g}" for v in row) + "\n")


def solution_fields(thermal=None, mechanical=None):
    """The fields save_results.hpp writes: (csv columns, vtk scalars, vtk vectors, vtk tensors)."""
    csv, scalars, vectors, tensors = [], [], [], []
    if thermal is not None:
        csv.append(("temperature", thermal.temperature()))
        scalars.append(("temperature", thermal.temperature()))
        if thermal.is_flux_calculated():
            fx, fy = thermal.flux()
            csv += [("flux_x", fx), ("flux_y", fy)]
            vectors.append(("flux", (fx, fy)))
    if mechanical is not None:
        ux, uy = mechanical.displacement()
        csv += [("displacement_x", ux), ("displacement_y", uy)]
        vectors.append(("displacement", (ux, uy)))
        if mechanical.is_strain_and_stress_calculated():
            e, s = mechanical.strain(), mechanical.stress()
            csv += [("strain_11", e[0]), ("strain_22", e[1]), ("strain_12", e[2]),
                    ("stress_11", s[0]), ("stress_22", s[1]), ("stress_12", s[2])]
            tensors += [("strain", e), ("stress", s)]
    if not csv:
        raise ValueError("Nothing to save.")
    return csv, scalars, vectors, tensors


def save_results(prefix, mesh, thermal=None, mechanical=None, precision=None):
    """<prefix>.csv and <prefix>.vtk of a thermal and/or mechanical solution (save_results.hpp:10-71)."""
    csv, scalars, vectors, tensors = solution_fields(thermal, mechanical)
    save_as_csv(f"{prefix}.csv", mesh, csv, precision)
    save_as_vtk(f"{prefix}.vtk", mesh, precision, scalars, vectors, tensors)
