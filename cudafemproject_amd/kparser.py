"""LS-DYNA keyword (.k) mesh -> prepared-mesh directory.

Same interface and output as the reference's mesh parser
(mesh_parser_lib/fem_parser.py:4-22 run_parser, FileParser in
mesh_parser_lib/parser_src/parser_lib.py:14-379):

* ``nodes.txt``: node count, then the coordinate tokens of every ``*NODE``
  line as written in the .k file (node index = order of appearance);
* ``elements.txt``: element count, then the first dim+1 node ids of every
  ``*ELEMENT_SOLID`` / ``*ELEMENT_SHELL`` line, as node indices;
* ``constraints.txt``: for every ``*BOUNDARY_SPC_SET`` the nodes of its
  ``*SET_NODE_LIST`` with the code of its (dofx, dofy[, dofz]) flags
  (2D: x=1, y=2, xy=3; 3D: x=1, y=2, z=3, xy=4, xz=5, yz=6, xyz=7);
* ``loads.txt``: for every ``*LOAD_NODE_SET`` its nodes with the ordinate of
  the last point of its ``*DEFINE_CURVE`` along the given dof, or ``0``;
* ``stress.txt``: for every ``*LOAD_SEGMENT`` "n_1 .. n_dim n_interior
  adj_elem normal p": the segment's nodes, the remaining node and the index of
  the first element containing the segment, the un-normalised segment normal
  turned away from that node (2D: (y2-y1, -(x2-x1)); 3D: (x2-x1) x (x3-x1)),
  and the ordinate of the last point of the segment's curve.

Differences from the reference, each where the reference produces a file its
own reader cannot use: a constraint with all dof flags zero is skipped (the
reference writes the node id without a code or newline); a segment with no
matching curve is skipped and not counted (the reference counts it but
writes no line); a segment with no containing element raises an error (the
reference fails with an IndexError).  Sections are also read to EOF safely.
"""
from __future__ import annotations

import os
import re
from typing import Dict, List

import numpy as np

_CODES_2D = {(1, 0): 1, (0, 1): 2, (1, 1): 3}
_CODES_3D = {(1, 0, 0): 1, (0, 1, 0): 2, (0, 0, 1): 3, (1, 1, 0): 4, (1, 0, 1): 5,
             (0, 1, 1): 6, (1, 1, 1): 7}


def _tokens(line: str) -> List[str]:
    return [t for t in re.split(r"\s+", line.strip()) if t]


class KMesh:
    """The sections of a .k file that the prepared mesh needs."""

    def __init__(self, path: str, dim: int):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        with open(path) as f:
            lines = f.read().splitlines()
        self.nodes: List[List[str]] = []           # [id, x, y(, z)] as text
        self.elements: List[List[str]] = []        # node id tokens (dim+1)
        self.sets: Dict[str, List[str]] = {}       # set header -> node ids
        self.spc: List[List[str]] = []             # *BOUNDARY_SPC_SET cards
        self.load_sets: List[List[str]] = []       # *LOAD_NODE_SET cards
        self.segments: List[List[str]] = []        # *LOAD_SEGMENT cards
        self.curves: Dict[str, str] = {}           # curve header -> last ordinate
        n = len(lines)
        i = 0
        while i < n:
            key = lines[i].strip()
            i += 1
            if key == "*NODE":
                i += 1                              # the column-title line
                while i < n and lines[i].strip() != "$" and not lines[i].startswith("*"):
                    t = _tokens(lines[i])
                    if t:
                        self.nodes.append(t[:dim + 1])
                    i += 1
            elif key in ("*ELEMENT_SOLID", "*ELEMENT_SHELL"):
                while i < n and not lines[i].strip().startswith("*"):
                    t = _tokens(lines[i])
                    if t:
                        self.elements.append(t[2:2 + dim + 1])
                    i += 1
            elif key == "*SET_NODE_LIST":
                head = lines[i].strip() if i < n else ""
                i += 1
                ids: List[str] = []
                while i < n and not lines[i].strip().startswith("*"):
                    ids.extend(_tokens(lines[i]))
                    i += 1
                self.sets[head] = ids
            elif key == "*BOUNDARY_SPC_SET":
                self.spc.append(_tokens(lines[i]))
                i += 1
            elif key == "*LOAD_NODE_SET":
                self.load_sets.append(_tokens(lines[i]))
                i += 1
            elif key == "*LOAD_SEGMENT":
                self.segments.append(_tokens(lines[i]))
                i += 1
            elif key == "*DEFINE_CURVE":
                head = lines[i].strip() if i < n else ""
                i += 1
                last = None
                while i < n and not lines[i].strip().startswith("*"):
                    t = _tokens(lines[i])
                    if len(t) > 1:
                        last = t[1]
                    i += 1
                if last is not None:
                    self.curves[head] = last
        self.index = {t[0]: k for k, t in enumerate(self.nodes)}

    def coords(self, node_id: str) -> np.ndarray:
        return np.array([float(v) for v in self.nodes[self.index[node_id]][1:]])

    def _segment_element(self):
        """First element (file order) containing each segment's dim nodes."""
        by_face: Dict[tuple, int] = {}
        d = self.dim
        for e, el in enumerate(self.elements):
            for skip in range(d + 1):
                face = tuple(sorted(el[:skip] + el[skip + 1:]))
                by_face.setdefault(face, e)
        out = []
        for s in self.segments:
            face = tuple(sorted(s[3:3 + d]))
            if face not in by_face:
                raise ValueError(f"load segment {s[3:3 + d]} lies on no element")
            e = by_face[face]
            interior = [v for v in self.elements[e] if v not in face]
            out.append((e, interior[0]))
        return out

    def write(self, out_dir: str):
        d = self.dim
        os.makedirs(out_dir, exist_ok=True)
        idx = self.index
        with open(os.path.join(out_dir, "nodes.txt"), "w") as f:
            f.write(f"{len(self.nodes)}\n")
            for t in self.nodes:
                f.write(" ".join(t[1:d + 1]) + "\n")
        with open(os.path.join(out_dir, "elements.txt"), "w") as f:
            f.write(f"{len(self.elements)}\n")
            for el in self.elements:
                f.write(" ".join(str(idx[v]) for v in el) + "\n")
        rows = []
        for c in self.spc:
            ids = self.sets.get(c[0])
            if ids is None:
                continue
            flags = tuple(int(float(v)) for v in c[2:2 + d])
            code = (_CODES_2D if d == 2 else _CODES_3D).get(flags)
            if code is None:
                continue
            rows.extend(f"{idx[v]} {code}" for v in ids)
        with open(os.path.join(out_dir, "constraints.txt"), "w") as f:
            f.write(f"{len(rows)}\n")
            for r in rows:
                f.write(r + "\n")
        with open(os.path.join(out_dir, "loads.txt"), "w") as f:
            if not self.load_sets:
                f.write("0")
            for ls in self.load_sets:
                ids = self.sets.get(ls[0])
                if ids is None:
                    continue
                f.write(f"{len(ids)}\n")
                val = self.curves.get(ls[2])
                for v in ids:
                    line = f"{idx[v]} "
                    if val is not None:
                        comp = ["0.0"] * d
                        k = int(ls[1]) - 1
                        if 0 <= k < d:
                            comp[k] = str(float(val))
                            line += " ".join(comp)
                    f.write(line + "\n")
        lines = []
        for s, (e, inner) in zip(self.segments, self._segment_element()):
            val = self.curves.get(s[0])
            if val is None:
                continue
            xs = [self.coords(v) for v in s[3:3 + d]]
            xb = self.coords(inner)
            if d == 2:
                n = np.array([xs[1][1] - xs[0][1], -(xs[1][0] - xs[0][0])])
            else:
                n = np.cross(xs[1] - xs[0], xs[2] - xs[0])
            if float(np.dot(n, xb - xs[0])) > 0:
                n = -n
            ids = " ".join(str(idx[v]) for v in s[3:3 + d])
            lines.append(f"{ids} {idx[inner]} {e} " + " ".join(str(float(c)) for c in n)
                         + f" {float(val)}\n")
        with open(os.path.join(out_dir, "stress.txt"), "w") as f:
            f.write(f"{len(lines)}\n")
            f.writelines(lines)


def run_parser(proj_dir: str, prepared_mesh_dir: str, raw_mesh_dir: str, filename: str,
               dim) -> KMesh:
    """Reference signature (mesh_parser_lib/fem_parser.py:4): parse the .k file
    at raw_mesh_dir and write the prepared mesh into prepared_mesh_dir."""
    m = KMesh(raw_mesh_dir, int(dim))
    m.write(prepared_mesh_dir)
    return m
