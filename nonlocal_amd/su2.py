"""SU2 text meshes: mesh_parser<T, I, mesh_format::SU2> (libraries/mesh/mesh_2d/su2_parser.hpp:165-195).

Sections read: NELEM (2D elements, each line "type n0 n1 ... index"), NPOIN (nodes "x y index") and
NMARK (named markers, each line "type n0 n1 ..." without an index).  As in the reference:

* element lines are identified by their text (type and node ids, trailing blanks trimmed), so a
  line repeated in a section counts once (su2_parser.hpp:74-88);
* a marker whose first element is two-dimensional is an element group; its elements are taken out
  of the NELEM set, and whatever remains forms the group "DEFAULT" (:140-149, :172-178), a name a
  file may not use (:174-175);
* one-dimensional markers are boundary groups;
* node orders are remapped per VTK type (:49-72): bilinear quads (9) keep their order, serendipity
  quads (23) go from corners-then-midsides to c0 m01 c1 m12 c2 m23 c3 m30, quadratic edges (21)
  from (a, b, mid) to (a, mid, b).

Element numbering: the reference keeps each group's element lines in an std::unordered_set and the
groups in an std::unordered_map, so its element ids follow the standard library's hash order.  Here
element groups are numbered in the order their markers appear (DEFAULT last) and the elements of
a group in file order.  Element ids only change the order of floating-point sums in the assembly
and the order of neighbour lists, not the problem.

Only quadrilateral meshes are supported (bilinear Q4 with linear edges, serendipity Q8 with
quadratic edges): triangles (5, 22) and the 9-node Lagrange quad (28) raise an error.
"""
from __future__ import annotations

import os

import numpy as np

from .meshgen import Mesh2D

DEFAULT_GROUP = "DEFAULT"                      # Default_Group_Name (mesh_container_2d.hpp:16)

# VTK element codes (vtk_elements_set.hpp:7-15) -> (dimension, file-order -> local-order permutation)
_LINEAR, _QUADRATIC, _TRIANGLE, _QUADRATIC_TRIANGLE = 3, 21, 5, 22
_BILINEAR, _SERENDIPITY, _LAGRANGE = 9, 23, 28
_ORDER = {
    _LINEAR: (1, (0, 1)),
    _QUADRATIC: (1, (0, 2, 1)),                 # read_element<0, 2, 1>: stored [a, mid, b]
    _BILINEAR: (2, (0, 1, 2, 3)),
    _SERENDIPITY: (2, (0, 2, 4, 6, 1, 3, 5, 7)),  # read_element<0, 2, 4, 6, 1, 3, 5, 7>
}
_UNSUPPORTED = {_TRIANGLE: "triangle", _QUADRATIC_TRIANGLE: "quadratic triangle",
                _LAGRANGE: "9-node Lagrange quad"}


def _element(line: str):
    """(type, local-order node ids) of one element line (su2_parser.hpp:49-72)."""
    tokens = line.split()
    kind = int(tokens[0])
    if kind in _UNSUPPORTED:
        raise ValueError(f"SU2 element type {kind} ({_UNSUPPORTED[kind]}) is not supported")
    if kind not in _ORDER:
        raise ValueError("Unknown element type.")
    _, perm = _ORDER[kind]
    ids = [int(t) for t in tokens[1:1 + len(perm)]]
    if len(ids) != len(perm):
        raise ValueError(f"element line with too few node ids: {line!r}")
    local = [0] * len(perm)
    for k, p in enumerate(perm):               # stream >> element[K]... : the k-th id goes to slot perm[k]
        local[p] = ids[k]
    return kind, local


def _is_2d(kind: int) -> bool:
    return _ORDER[kind][0] == 2


class _Lines:
    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def next(self) -> str:
        while self.pos < len(self.lines):
            line = self.lines[self.pos]
            self.pos += 1
            if line.strip():
                return line.rstrip()
        raise ValueError("unexpected end of SU2 file")

    def keyword(self, name: str) -> str:
        line = self.next()
        key, _, value = line.partition("=")
        if key.strip() != name:
            raise ValueError(f"expected {name}=, found {line!r}")
        return value.strip()


def _unique(lines):
    """Element lines in first-appearance order, each text once."""
    seen = set()
    out = []
    for line in lines:
        if line not in seen:
            seen.add(line)
            out.append(line)
    return out


def parse(text: str) -> Mesh2D:
    """Mesh2D of an SU2 text (mesh_parser<SU2>::parse, su2_parser.hpp:165-195)."""
    src = _Lines(text)
    src.keyword("NDIME")
    n_el = int(src.keyword("NELEM"))
    elements_2d = []
    for _ in range(n_el):
        line = src.next().rstrip()
        elements_2d.append(line[:line.rfind(" ")].rstrip() if " " in line else line)   # drop the index
    elements_2d = _unique(elements_2d)
    n_nodes = int(src.keyword("NPOIN").split()[0])
    xy = np.empty((n_nodes, 2), dtype=np.float64)
    for k in range(n_nodes):
        t = src.next().split()
        xy[k, 0] = float(t[0])
        xy[k, 1] = float(t[1])
    n_markers = int(src.keyword("NMARK"))
    markers = []
    for _ in range(n_markers):
        name = src.keyword("MARKER_TAG")
        count = int(src.keyword("MARKER_ELEMS"))
        lines = _unique([src.next().strip() for _ in range(count)])
        markers.append((name, lines))
    names = [m[0] for m in markers]
    if DEFAULT_GROUP in names:
        raise ValueError(f'The group name cannot be named "{DEFAULT_GROUP}", as it is a reserved group name.')
    if len(set(names)) != len(names):
        raise ValueError("repeated marker name")

    groups_2d, groups_1d = [], []
    for name, lines in markers:
        if not lines:
            continue
        (groups_2d if _is_2d(_element(lines[0])[0]) else groups_1d).append((name, lines))
    in_groups = set()
    for _, lines in groups_2d:
        in_groups.update(lines)
    rest = [line for line in elements_2d if line not in in_groups]
    if rest:
        groups_2d.append((DEFAULT_GROUP, rest))
    if not groups_2d:
        raise ValueError("the mesh has no 2D elements")

    elems, kinds, group_begin = [], set(), [0]
    for _, lines in groups_2d:
        for line in lines:
            kind, ids = _element(line)
            if not _is_2d(kind):
                raise ValueError(f"a 2D group holds the 1D element {line!r}")
            kinds.add(kind)
            elems.append(ids)
        group_begin.append(len(elems))
    if len(kinds) != 1:
        raise ValueError("meshes mixing element types are not supported")
    kind2d = kinds.pop()
    nn = 4 if kind2d == _BILINEAR else 8
    edge_kind = _LINEAR if nn == 4 else _QUADRATIC

    bedges, bgroup_begin = [], [0]
    for _, lines in groups_1d:
        for line in lines:
            kind, ids = _element(line)
            if kind != edge_kind:
                raise ValueError(f"boundary element {line!r} does not match the {nn}-node quads")
            bedges.append(ids)
        bgroup_begin.append(len(bedges))
    nb = 2 if nn == 4 else 3
    return Mesh2D(
        xy=xy,
        elems=np.asarray(elems, dtype=np.int32).reshape(-1, nn),
        nn=nn,
        group_begin=np.asarray(group_begin, dtype=np.int64),
        group_names=[g[0] for g in groups_2d],
        bedges=np.asarray(bedges, dtype=np.int32).reshape(-1, nb),
        bgroup_begin=np.asarray(bgroup_begin, dtype=np.int64),
        bgroup_names=[g[0] for g in groups_1d],
    )


def read(path: str | os.PathLike) -> Mesh2D:
    with open(path, "r", encoding="ascii") as f:
        return parse(f.read())
