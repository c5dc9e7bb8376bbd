import functools
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "3d-voronoi-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@functools.lru_cache(maxsize=None)
def problem(kind: str, n: int, n_sites: int, seed: int = 7):
    """Normalised (verts, tets, sites) of a seeded synthetic construction."""
    from voronoi_amd import meshgen

    return meshgen.make_problem(kind, n, n_sites, seed=seed)


@functools.lru_cache(maxsize=None)
def oracle_knn(kind: str, n: int, n_sites: int, k: int, tk: int, seed: int = 7):
    import oracle as O

    verts, tets, sites = problem(kind, n, n_sites, seed)
    kk = min(k, n_sites - 1)
    return O.site_knn(sites, kk), O.tet_knn(verts, tets, sites, min(tk, n_sites))


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def vdn():
    from voronoi_amd import native

    native.lib()
    return native
