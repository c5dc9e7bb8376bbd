"""The reference's drop-in surface on the GPU: the VolumeVoronoiGPU command
line (files written, record.csv columns, out.xyz equal to the engine's result)
and the C++ shims with the reference's kNN signatures."""
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

pytestmark = pytest.mark.gpu

BUILD = os.path.join(ROOT, "3d-voronoi-gpu_amd", "_build")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_cli_on_the_reference_fixture(vdn, tmp_path):
    from voronoi_amd import fileio, meshgen

    exe = os.path.join(BUILD, "VolumeVoronoiGPU")
    # the FP64 form: its site sums do not depend on the order of the atomics,
    # so the two runs below agree to the 6 printed digits
    env = dict(os.environ, VD_ACCUM="f64")
    r = subprocess.run([exe, os.path.join(GOLDEN, "joint.tet"), os.path.join(GOLDEN, "joint.xyz"), "2", "0"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    for f in ("out.xyz", "iter_0.xyz", "iter_1.xyz", "record.csv"):
        assert (tmp_path / f).exists(), f
    rec = (tmp_path / "record.csv").read_text().splitlines()
    assert rec[0].startswith("n_site, n_tet, TET_K, SITE_K, Iter, k-NN")
    assert [row.split(", ")[:5] for row in rec[1:]] == [["3000", "31349", "20", "90", "0"],
                                                       ["3000", "31349", "20", "90", "1"]]
    assert (tmp_path / "out.xyz").read_text() == (tmp_path / "iter_1.xyz").read_text()
    # the same two iterations through the C ABI
    verts, tets = fileio.read_tet(os.path.join(GOLDEN, "joint.tet"))
    sites = fileio.read_xyz(os.path.join(GOLDEN, "joint.xyz"))
    with vdn.Engine(accum="f64", form="cells") as e:
        e.set_mesh(meshgen.normalize(verts), tets)
        e.set_sites(sites)
        e.lloyd(2)
        want = e.sites()
    got = fileio.read_xyz(str(tmp_path / "out.xyz"))
    np.testing.assert_allclose(got, want, rtol=6e-6)  # out.xyz holds 6 significant digits


def test_knearests_and_knncuda_shims(O, tmp_path):
    rng = np.random.default_rng(4)
    n = 15000
    pts = (rng.random((n, 3)) * 1000).astype(np.float32)
    pts.tofile(tmp_path / "pts.f32")
    r = subprocess.run([os.path.join(BUILD, "compat_check"), str(tmp_path / "pts.f32"), str(n), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    perm = np.fromfile(tmp_path / "kn_perm.u32", dtype=np.uint32).astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "kn_points.f32", dtype=np.float32).reshape(n, 3), pts[perm])
    knn = np.fromfile(tmp_path / "kn_knn.u32", dtype=np.uint32).reshape(n, 90).astype(np.int64)
    want = O.site_knn(pts, 90)
    np.testing.assert_array_equal(perm[knn], want[perm])
    rows = np.fromfile(tmp_path / "dev_knn.i32", dtype=np.int32).reshape(91, n).T
    np.testing.assert_array_equal(rows, O.knn(pts, pts, 91, False))
