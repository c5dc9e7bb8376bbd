"""Writes tests/golden/qhull_cells.npz: exact (cell of site s) ∩ (tet t) volumes
and centroids for small seeded problems, computed with scipy's Qhull
(HalfspaceIntersection + ConvexHull) in FP64 from all other sites, i.e. an
independent computation of the quantity the reference's kernel A integrates.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np
from scipy.optimize import linprog
from scipy.spatial import ConvexHull, HalfspaceIntersection

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "3d-voronoi-gpu_amd"))
from voronoi_amd import meshgen  # noqa: E402

FACES = [(2, 1, 3), (0, 2, 3), (1, 0, 3), (0, 1, 2)]  # faces cbd, acd, bad, abc


def cell_tet(verts, tet, sites, s):
    """Volume and centroid of {x in tet : |x - s| <= |x - z| for all sites z}."""
    V = verts.astype(np.float64)[tet]
    S = sites.astype(np.float64)
    hs = []
    for f in FACES:
        v0, v1, v2 = V[f[0]], V[f[1]], V[f[2]]
        n = np.cross(v1 - v0, v2 - v0)  # inward for a positively oriented tet
        hs.append(np.r_[-n, n @ v0])  # scipy form: A x + b <= 0
    for z in range(len(S)):
        if z == s:
            continue
        d = S[s] - S[z]
        hs.append(np.r_[-d, d @ (S[s] + S[z]) / 2])
    hs = np.array(hs)
    A, b = hs[:, :3], -hs[:, 3]
    nrm = np.linalg.norm(A, axis=1)
    res = linprog([0, 0, 0, -1], A_ub=np.c_[A, nrm], b_ub=b, bounds=[(None, None)] * 3 + [(0, None)])
    if res.status != 0 or res.x[3] < 1e-7:
        return 0.0, np.zeros(3)
    pts = HalfspaceIntersection(hs, res.x[:3]).intersections
    hull = ConvexHull(pts)
    # centroid: decompose into tets from an interior point
    c0 = pts.mean(axis=0)
    vol, mom = 0.0, np.zeros(3)
    for simplex in hull.simplices:
        a, b_, c = pts[simplex]
        v = abs(np.dot(np.cross(b_ - a, c - a), c0 - a)) / 6
        vol += v
        mom += v * (a + b_ + c + c0) / 4
    return vol, mom / vol


def make(kind, n, n_sites, seed):
    verts, tets, sites = meshgen.make_problem(kind, n, n_sites, seed=seed)
    vol = np.zeros((len(tets), n_sites))
    cen = np.zeros((len(tets), n_sites, 3))
    for t in range(len(tets)):
        for s in range(n_sites):
            vol[t, s], cen[t, s] = cell_tet(verts, tets[t], sites, s)
    return dict(verts=verts, tets=tets, sites=sites, vol=vol, cen=cen)


def main():
    out = {}
    for name, (kind, n, ns, seed) in {"cube3_s12": ("cube", 3, 12, 21), "cube4_s30": ("cube", 4, 30, 22)}.items():
        d = make(kind, n, ns, seed)
        for k, v in d.items():
            out[f"{name}/{k}"] = v
        print(name, "tets", len(d["tets"]), "sites", ns, "total volume", d["vol"].sum())
    np.savez_compressed(os.path.join(HERE, "qhull_cells.npz"), **out)


if __name__ == "__main__":
    main()
