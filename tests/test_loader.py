"""The mesh loader of the command line and of the Python package (CPU only).

The reference maps a loaded mesh into [0,1000]^3 with
x' = 1000 * (x - lo) / side, lo = bbox min - 0.1 % of the largest side and
side = the grown box's largest side (test_voronoi.cu:14-38, 103-118).  For a
box [0,a]x[0,b]x[0,c] with a the largest side that is, in exact arithmetic,
x' = 1000 * (x + 0.001a) / (1.002a) on every axis; for the unit cube
1000 * (x + 0.001) / 1.002.  The C++ loader (csrc/cli/mesh_io.hpp, built here
into a small host program) and meshgen.normalize are checked against that
closed form, against each other bit for bit, and on the reference's own
fixture in both its .tet form and a 1-based .vtk conversion.
"""
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI_DIR = os.path.join(ROOT, "3d-voronoi-gpu_amd", "csrc", "cli")


@pytest.fixture(scope="module")
def loader(tmp_path_factory):
    exe = tmp_path_factory.mktemp("loader") / "loader_dump"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CLI_DIR, "-o", str(exe),
                    os.path.join(ROOT, "tests", "loader_dump.cpp")], check=True)

    def run(mesh, sites, out):
        os.makedirs(out, exist_ok=True)
        r = subprocess.run([str(exe), str(mesh), str(sites), str(out)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        v = np.fromfile(os.path.join(out, "verts.f32"), dtype=np.float32).reshape(-1, 3)
        t = np.fromfile(os.path.join(out, "tets.i32"), dtype=np.int32).reshape(-1, 4)
        s = np.fromfile(os.path.join(out, "sites.f32"), dtype=np.float32).reshape(-1, 3)
        txt = open(os.path.join(out, "sites.txt")).read()
        return v, t, s, txt

    return run


def _box_mesh(sides, n=4, seed=5):
    """A jittered cube mesh scaled to the box [0,a]x[0,b]x[0,c]: its boundary
    vertices lie exactly on the box faces."""
    from voronoi_amd import meshgen

    v, t = meshgen.jittered_cube(n, seed=seed)
    return (v * np.asarray(sides, np.float32)).astype(np.float32), t


@pytest.mark.parametrize("sides", [(1.0, 1.0, 1.0), (2.0, 1.0, 0.5), (0.25, 4.0, 1.0)])
def test_normalisation_closed_form(loader, tmp_path, sides):
    from voronoi_amd import fileio, meshgen

    v, t = _box_mesh(sides)
    assert v.min() == 0.0 and np.allclose(v.max(axis=0), sides)
    path = tmp_path / "box.tet"
    fileio.write_tet(str(path), v, t)
    fileio.write_xyz(str(tmp_path / "s.xyz"), v[:5])
    got, gt, _, _ = loader(path, tmp_path / "s.xyz", tmp_path / "out")
    # the file holds the FP32 vertices exactly (9 significant digits)
    vf, tf = fileio.read_tet(str(path))
    np.testing.assert_array_equal(vf, v)
    np.testing.assert_array_equal(gt, t)
    a = max(sides)
    want = 1000.0 * (v.astype(np.float64) + 0.001 * a) / (1.002 * a)
    # FP32 evaluation of the same map: a few ulp of 1000
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-4)
    # the two faces of the largest side land at 1000 * 0.001 / 1.002 and 1000 * 1.001 / 1.002
    ax = int(np.argmax(sides))
    assert abs(got[:, ax].min() - 1000 * 0.001 / 1.002) < 1e-4
    assert abs(got[:, ax].max() - 1000 * 1.001 / 1.002) < 1e-4
    # the Python restatement gives the same bits
    np.testing.assert_array_equal(meshgen.normalize(v), got)


def test_unit_cube_values():
    """A few values of the unit-cube map, x' = 1000 (x + 0.001) / 1.002."""
    from voronoi_amd import meshgen

    v = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.25, 0.75]], dtype=np.float32)
    got = meshgen.normalize(v)
    np.testing.assert_allclose(got[0], [0.998004] * 3, atol=1e-5)
    np.testing.assert_allclose(got[1], [999.002] * 3, atol=1e-3)
    np.testing.assert_allclose(got[2], [499.999, 250.499, 749.501], atol=1e-3)


def test_reference_fixture_tet_and_vtk(loader, tmp_path):
    """data/joint.tet through the C++ loader equals fileio + meshgen.normalize
    bit for bit, and its 1-based .vtk conversion loads to the same mesh."""
    from voronoi_amd import fileio, meshgen

    tet = os.path.join(GOLDEN, "joint.tet")
    xyz = os.path.join(GOLDEN, "joint.xyz")
    got_v, got_t, got_s, txt = loader(tet, xyz, tmp_path / "tet")
    v, t = fileio.read_tet(tet)
    s = fileio.read_xyz(xyz)
    assert got_v.shape == (len(v), 3) and got_t.shape == (31349, 4) and len(got_s) == 3000
    np.testing.assert_array_equal(got_t, t)
    np.testing.assert_array_equal(got_v, meshgen.normalize(v))
    np.testing.assert_array_equal(got_s, s)
    lines = txt.splitlines()
    assert lines[0] == "3000" and lines[1] == "%g  %g  %g" % tuple(float(x) for x in s[0])
    # every normalised vertex inside the grown frame
    assert got_v.min() > 0 and got_v.max() < 1000
    vtk = tmp_path / "joint.vtk"
    fileio.write_vtk(str(vtk), v, t)
    vv, vt, _, _ = loader(vtk, xyz, tmp_path / "vtk")
    np.testing.assert_array_equal(vt, got_t)
    np.testing.assert_array_equal(vv, got_v)
    # the .vtk file really is 1-based: its first cell line is the first tet plus one
    body = open(vtk).read().split("CELLS")[1].splitlines()
    assert [int(x) for x in body[1].split()] == [4] + (t[0] + 1).tolist()


def test_loader_rejects_bad_input(loader, tmp_path):
    bad = tmp_path / "bad.tet"
    bad.write_text("4 vertices\n1 cells\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n3 0 1 2\n")
    (tmp_path / "s.xyz").write_text("1\n0 0 0\n")
    exe_dir = tmp_path / "o"
    os.makedirs(exe_dir)
    with pytest.raises(AssertionError):
        loader(bad, tmp_path / "s.xyz", exe_dir)
    with pytest.raises(AssertionError):
        loader(tmp_path / "mesh.obj", tmp_path / "s.xyz", exe_dir)
