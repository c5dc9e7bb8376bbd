"""Readers and writers for the reference's file formats.

  .tet : "<n> vertices", "<m> cells", n lines "x y z", m lines "4 a b c d" (0-based)
         (test_voronoi.cu:55-73)
  .vtk : 4 header lines, "POINTS n <type>", coordinates, "CELLS m size",
         m lines "4 a b c d" with 1-based indices (test_voronoi.cu:74-93)
  .xyz : count line, then "x y z" per site (test_voronoi.cu:135-147); written
         as "x  y  z" with 6 significant digits (test_voronoi.cu:123-133)
"""
from __future__ import annotations

import numpy as np


def write_tet(path: str, verts: np.ndarray, tets: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write(f"{len(verts)} vertices\n{len(tets)} cells\n")
        np.savetxt(f, verts.astype(np.float64), fmt="%.9g")
        np.savetxt(f, np.hstack([np.full((len(tets), 1), 4), tets]), fmt="%d")


def read_tet(path: str):
    if path.endswith(".vtk"):
        return read_vtk(path)
    with open(path) as f:
        nv = int(f.readline().split()[0])
        nt = int(f.readline().split()[0])
        data = np.fromstring(f.read(), sep=" ")
    verts = data[: 3 * nv].reshape(nv, 3).astype(np.float32)
    cells = data[3 * nv : 3 * nv + 5 * nt].reshape(nt, 5).astype(np.int64)
    assert np.all(cells[:, 0] == 4)
    return verts, cells[:, 1:].astype(np.int32)


def read_vtk(path: str):
    with open(path) as f:
        for _ in range(4):
            f.readline()
        tok = f.read().split()
    assert tok[0] == "POINTS"
    nv = int(tok[1])
    verts = np.array(tok[3 : 3 + 3 * nv], dtype=np.float64).reshape(nv, 3).astype(np.float32)
    p = 3 + 3 * nv
    assert tok[p] == "CELLS"
    nt = int(tok[p + 1])
    cells = np.array(tok[p + 3 : p + 3 + 5 * nt], dtype=np.int64).reshape(nt, 5)
    assert np.all(cells[:, 0] == 4)
    return verts, (cells[:, 1:] - 1).astype(np.int32)


def write_vtk(path: str, verts: np.ndarray, tets: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\ntet mesh\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {len(verts)} float\n")
        np.savetxt(f, verts.astype(np.float64), fmt="%.9g")
        f.write(f"CELLS {len(tets)} {5 * len(tets)}\n")
        np.savetxt(f, np.hstack([np.full((len(tets), 1), 4), tets + 1]), fmt="%d")
        f.write(f"CELL_TYPES {len(tets)}\n")
        np.savetxt(f, np.full(len(tets), 10), fmt="%d")


def write_xyz(path: str, sites: np.ndarray, sep: str = " ") -> None:
    with open(path, "w") as f:
        f.write(f"{len(sites)}\n")
        for x, y, z in sites.astype(np.float64):
            f.write(f"{x:.9g}{sep}{y:.9g}{sep}{z:.9g}\n")


def read_xyz(path: str) -> np.ndarray:
    with open(path) as f:
        n = int(f.readline().split()[0])
        data = np.fromstring(f.read(), sep=" ")
    return data[: 3 * n].reshape(n, 3).astype(np.float32)
