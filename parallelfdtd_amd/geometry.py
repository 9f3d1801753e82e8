"""Room geometry input for the voxelizer (``pfdtd_voxelize``).

* ``load_json`` -- the reference's JSON room format as ``python/testBench.py:42-73`` reads it:
  ``vertices`` (flat x,y,z list, metres), ``indices`` (flat, 3 per triangle, from 0),
  ``layers_of_triangles`` (layer name per triangle) and ``layer_names``.
* ``read_vtk`` -- legacy-VTK polydata as ``FileReader::readVTK`` (``src/io/FileReader.cpp:41-89``)
  reads it: within its first 20 getline calls, ``POINTS n`` is followed by 3n coordinates in inches,
  converted to metres and rounded to ``truncate_vertices_to`` (0.01 m) in float arithmetic, and
  ``POLYGONS n size`` by n rows ``3 i1 i2 i3``.
* ``layer_materials`` -- per-triangle material indices from a layer -> material map (what the
  reference scripts build before ``addSurfaceMaterials``, ``python/testBench.py:86-105``).
* ``box_mesh``, ``mesh_volume`` -- an axis-aligned closed box (12 outward triangles) and the
  signed volume of a closed mesh.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


@dataclass
class Mesh:
    vertices: np.ndarray  # [n][3] float32, metres
    indices: np.ndarray   # [m][3] uint32
    layers: dict[str, list[int]] = field(default_factory=dict)  # layer name -> triangle ids

    @property
    def n_triangles(self) -> int:
        return len(self.indices)


def load_json(path: str | Path) -> Mesh:
    with open(path) as f:
        m = json.load(f)
    v = np.asarray(m["vertices"], np.float32).reshape(-1, 3)
    t = np.asarray(m["indices"], np.uint32).reshape(-1, 3)
    layers: dict[str, list[int]] = {}
    per_tri = m.get("layers_of_triangles", [])
    for name in m.get("layer_names", []):
        layers[name] = [i for i, n in enumerate(per_tri) if n == name]
    return Mesh(v, t, layers)


class _Stream:
    """The two std::istream operations readVTK uses: getline and >> of one token."""

    def __init__(self, text: str):
        self.t = text
        self.i = 0

    def getline(self) -> str | None:
        if self.i >= len(self.t):
            return None
        j = self.t.find("\n", self.i)
        j = len(self.t) if j < 0 else j
        line, self.i = self.t[self.i:j], j + 1
        return line

    def token(self) -> str:
        n = len(self.t)
        while self.i < n and self.t[self.i].isspace():
            self.i += 1
        j = self.i
        while j < n and not self.t[j].isspace():
            j += 1
        tok, self.i = self.t[self.i:j], j
        if not tok:
            raise ValueError("unexpected end of VTK data")
        return tok


def read_vtk(path: str | Path, truncate_vertices_to: float = 0.01) -> Mesh:
    st = _Stream(Path(path).read_text())
    v = np.zeros((0, 3), np.float32)
    t = np.zeros((0, 3), np.uint32)
    trunc = np.float32(truncate_vertices_to)
    for _ in range(20):  # 20 getline calls; data after a keyword is read token by token
        line = st.getline()
        if line is None:
            break
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "POINTS":
            n = int(parts[1])
            raw = np.asarray([st.token() for _ in range(3 * n)], np.float32)
            with np.errstate(all="ignore"):
                r = np.floor((raw * np.float32(0.0254)) / trunc + np.float32(0.5)).astype(np.float32)
            v = (r * trunc).astype(np.float32).reshape(-1, 3)
        elif parts[0] == "POLYGONS":
            n = int(parts[1])
            raw = np.asarray([st.token() for _ in range(4 * n)], np.int64)
            t = raw.reshape(-1, 4)[:, 1:].astype(np.uint32)
    return Mesh(v, t)


def write_vtk(path: str | Path, mesh: Mesh) -> None:
    """Legacy-VTK polydata in the reference's unit convention (coordinates in inches)."""
    v = np.asarray(mesh.vertices, np.float64) / 0.0254
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nroom\nASCII\nDATASET POLYDATA\n")
        f.write(f"POINTS {len(v)} float\n")
        for p in v:
            f.write(f"{p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        f.write(f"POLYGONS {len(mesh.indices)} {4 * len(mesh.indices)}\n")
        for a, b, c in np.asarray(mesh.indices):
            f.write(f"3 {a} {b} {c}\n")


def layer_materials(mesh: Mesh, materials: dict[str, int], default: int = 0) -> np.ndarray:
    """uint8 material index per triangle: `materials[layer]` for the triangles of each named
    layer, `default` elsewhere."""
    m = np.full(mesh.n_triangles, default, np.uint8)
    for name, idx in materials.items():
        m[mesh.layers[name]] = idx
    return m


# faces of the unit cube as corner bit patterns (x=1, y=2, z=4), each split into two
# triangles wound counter-clockwise seen from outside
_BOX_FACES = {
    "-x": (0, 4, 6, 2), "+x": (1, 3, 7, 5), "-y": (0, 1, 5, 4), "+y": (2, 6, 7, 3),
    "-z": (0, 2, 3, 1), "+z": (4, 5, 7, 6),
}


def box_mesh(lo, hi) -> Mesh:
    """Closed axis-aligned box [lo, hi] with outward normals; layers '-x', '+x', ... name the
    two triangles of each face."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    corners = np.array([[hi[0] if c & 1 else lo[0], hi[1] if c & 2 else lo[1],
                         hi[2] if c & 4 else lo[2]] for c in range(8)], np.float32)
    tris, layers = [], {}
    for name, (a, b, c, d) in _BOX_FACES.items():
        layers[name] = [len(tris), len(tris) + 1]
        tris += [(a, b, c), (a, c, d)]
    return Mesh(corners, np.asarray(tris, np.uint32), layers)


def merge(*meshes: Mesh) -> Mesh:
    """One mesh holding all triangles (layers prefixed by the part number)."""
    vs, ts, layers, off, toff = [], [], {}, 0, 0
    for k, m in enumerate(meshes):
        vs.append(np.asarray(m.vertices, np.float32))
        ts.append(np.asarray(m.indices, np.uint32) + off)
        for name, idx in m.layers.items():
            layers[f"{k}:{name}"] = [i + toff for i in idx]
        off += len(m.vertices)
        toff += m.n_triangles
    return Mesh(np.concatenate(vs), np.concatenate(ts), layers)


def mesh_volume(mesh: Mesh) -> float:
    """Signed enclosed volume (positive for outward winding)."""
    v = np.asarray(mesh.vertices, np.float64)
    t = np.asarray(mesh.indices, np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)
