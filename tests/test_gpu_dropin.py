"""The drop-in path on the reference's own input, against the oracle.

data/joint.tet + data/joint.xyz (the reference's example, README.md:31-33) is a
non-convex joint of 31,349 tets with 3,000 sites.  The reference's out.xyz
comes from its cell -> tet formulation (kernels B+C, voronoi.cu:1090-1109); its
tet -> cell kernel A runs beside it (voronoi.cu:1034-1076).  Here:

  * per pair and per site, form "cells" against the oracle's clip_cells and
    form "site" against its clip_tets, at the reference's TET_K heuristic (20
    for this fixture, voronoi.cu:936) and at TET_K = 64, bit for bit in the
    FP64 accumulation mode (order-independent fixed-point site sums);
  * three Lloyd iterations through the engine, the whole-program driver
    (compute_clipped_voro_diagram, voronoi.cu:925-1188) and the
    VolumeVoronoiGPU command line, each iteration's sites against an oracle
    chain of site kNN -> candidate kNN -> clipping -> update: bit for bit in
    memory, and the iter_<i>.xyz / out.xyz files character for character
    (6 significant digits, test_voronoi.cu:123-133, voronoi.cu:1166-1173);
  * the reference's default FP32 atomics, to 1e-5 of the frame;
  * the record.csv rows and columns, and a 1-based .vtk copy of the mesh.

The mesh is normalised by meshgen.normalize, which equals the command line's
C++ loader bit for bit (tests/test_loader.py).
"""
import functools
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

pytestmark = pytest.mark.gpu

BUILD = os.path.join(ROOT, "3d-voronoi-gpu_amd", "_build")
GOLDEN = os.path.join(ROOT, "tests", "golden")
TET = os.path.join(GOLDEN, "joint.tet")
XYZ = os.path.join(GOLDEN, "joint.xyz")


@functools.lru_cache(maxsize=None)
def joint():
    from voronoi_amd import fileio, meshgen

    v, t = fileio.read_tet(TET)
    return meshgen.normalize(v), t, fileio.read_xyz(XYZ)


@functools.lru_cache(maxsize=None)
def oracle_chain(form, n_iter, tk=0):
    """Sites after each of n_iter oracle Lloyd iterations (FP32 clipping, FP64
    fixed-point site sums, FP64 update) of the given formulation."""
    import oracle as O

    v, t, s = joint()
    tk = O.tet_k(len(s), len(t), tk)
    clip = O.clip_cells if form == "cells" else O.clip_tets
    out = []
    for _ in range(n_iter):
        sk = O.site_knn(s, 90)
        tkn = O.tet_knn(v, t, s, tk)
        r = clip(s, sk, v, t, tkn, geom="f32", pair_acc="f64", site_mode=O.SITE_FIXED)
        s = O.lloyd_update(r["site_vals"], s, fp32=False)
        out.append(s)
    return out


def _fmt(sites):
    return "%d\n" % len(sites) + "".join("%g  %g  %g\n" % tuple(float(x) for x in p) for p in sites)


@pytest.mark.parametrize("tk", [0, 64])
@pytest.mark.parametrize("form", ["cells", "site"])
def test_joint_pairs_vs_oracle(vdn, O, form, tk):
    v, t, s = joint()
    print(f"joint form={form} tet_k={tk}", flush=True)
    with vdn.Engine(tet_k=tk, form=form, accum="f64", record_pairs=1) as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        K, TK, _, _ = e.dims()
        assert K == 90 and TK == (20 if tk == 0 else tk)
        e.knn_step()
        sk, tkn = e.site_knn(), e.tet_knn()
        e.clip_step()
        st, vals = e.pairs()
        mom = e.moments()
        hist = e.status_hist()
        e.update_step()
        new = e.sites()
    np.testing.assert_array_equal(sk, O.site_knn(s, 90))
    np.testing.assert_array_equal(tkn, O.tet_knn(v, t, s, TK))
    clip = O.clip_cells if form == "cells" else O.clip_tets
    r = clip(s, sk, v, t, tkn, geom="f32", pair_acc="f64", site_mode=O.SITE_FIXED)
    bad = np.nonzero(st.astype(np.int32) != r["pair_status"])[0]
    assert bad.size == 0, f"{bad.size} statuses differ, first {bad[:5]}"
    dv = np.nonzero(np.any(vals != r["pair_vals"], axis=1))[0]
    assert dv.size == 0, f"{dv.size} pair values differ, first {dv[:5]}"
    assert [hist[k] for k in vdn.STATUS_NAMES] == np.bincount(r["pair_status"], minlength=7).tolist()
    np.testing.assert_array_equal(mom, r["site_vals"])
    np.testing.assert_array_equal(new, O.lloyd_update(r["site_vals"], s, fp32=False))
    # the non-convex joint is covered: every site owns some volume
    assert np.all(mom[:, 0] > 0)


@pytest.mark.parametrize("form", ["cells", "site"])
def test_joint_lloyd_chain_engine(vdn, form):
    """Each of three iterations through the engine equals the oracle chain."""
    v, t, s = joint()
    want = oracle_chain(form, 3)
    with vdn.Engine(form=form, accum="f64") as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        for i in range(3):
            print(f"form={form} iteration {i}", flush=True)
            e.lloyd(1)
            np.testing.assert_array_equal(e.sites(), want[i])


def test_joint_lloyd_chain_driver(vdn, tmp_path):
    """compute_clipped_voro_diagram (the reference's driver signature, SoA
    sites as the command line passes fewer than 14,000): sites in memory and
    every iter_<i>.xyz equal the oracle chain."""
    v, t, s = joint()
    want = oracle_chain("cells", 3)
    soa = np.ascontiguousarray(s.T)
    got = vdn.compute_clipped_voro_diagram(v, t, soa, nb_iter=3, tet_k=0, out_dir=str(tmp_path), transposed=True,
                                           form="cells", accum="f64")
    np.testing.assert_array_equal(got.T, want[2])
    for i in range(3):
        assert (tmp_path / f"iter_{i}.xyz").read_text() == _fmt(want[i]), f"iter_{i}.xyz"


def _cli(args, cwd, **env):
    exe = os.path.join(BUILD, "VolumeVoronoiGPU")
    r = subprocess.run([exe] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, **env))
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0
    return r


def test_cli_on_the_reference_fixture(vdn, tmp_path):
    """VolumeVoronoiGPU joint.tet joint.xyz 3 0: files, record.csv rows and
    columns, and every iteration's file equal to the oracle chain."""
    want = oracle_chain("cells", 3)
    _cli([TET, XYZ, 3, 0], tmp_path, VD_ACCUM="f64")
    for i in range(3):
        assert (tmp_path / f"iter_{i}.xyz").read_text() == _fmt(want[i]), f"iter_{i}.xyz"
    assert (tmp_path / "out.xyz").read_text() == _fmt(want[2])
    rec = (tmp_path / "record.csv").read_text().splitlines()
    assert rec[0] == "n_site, n_tet, TET_K, SITE_K, Iter, k-NN, Clip_tet-cell, Clip_cell-tet, Compute_new_site, GPU2CPU"
    rows = [r.split(", ") for r in rec[1:]]
    assert [r[:5] for r in rows] == [["3000", "31349", "20", "90", str(i)] for i in range(3)]
    for r in rows:
        assert len(r) == 11 and r[10] == ""  # each value followed by ", " (voronoi.cu:993-1163)
        knn, tet_cell, cell_tet = (float(x) for x in r[5:8])
        # the cell -> tet formulation ran: its time in its own column (voronoi.cu:1102-1116)
        assert cell_tet > 0 and tet_cell == 0.0 and knn > 0
        assert all(len(x.split(".")[1]) == 5 for x in r[5:10])  # fixed, 5 decimals
    # the tet -> cell form puts its time in the other column
    sub = tmp_path / "site"
    sub.mkdir()
    _cli([TET, XYZ, 1, 0], sub, VD_ACCUM="f64", VD_FORM="site")
    r = (sub / "record.csv").read_text().splitlines()[1].split(", ")
    assert float(r[6]) > 0 and float(r[7]) == 0.0
    assert (sub / "out.xyz").read_text() == _fmt(oracle_chain("site", 1)[0])


def test_cli_vtk_equals_tet(vdn, tmp_path):
    """The same mesh as a 1-based .vtk file gives the .tet run's out.xyz."""
    from voronoi_amd import fileio

    v, t = fileio.read_tet(TET)
    fileio.write_vtk(str(tmp_path / "joint.vtk"), v, t)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _cli([TET, XYZ, 1], a, VD_ACCUM="f64")
    _cli([tmp_path / "joint.vtk", XYZ, 1], b, VD_ACCUM="f64")
    assert (b / "out.xyz").read_text() == (a / "out.xyz").read_text()
    assert (a / "out.xyz").read_text() == _fmt(oracle_chain("cells", 1)[0])


def test_cli_fp32_atomics_close_to_oracle(vdn, tmp_path):
    """The reference's default FP32 atomic accumulation, whose summation order
    varies from run to run.  After iteration 1 every site is within 1e-5 of
    the frame size of the FP64 chain's.  Iteration 2 starts from sites that
    differ at rounding level, so a pair next to one of the clipping's
    thresholds (the security radius, the 0.1 piece volume) may come out the
    other way and move its site by a small fraction of a cell (seen: one site
    of 3,000 by 0.069, a thousandth of its cell's size): every site within
    1e-3 of the frame, all but 0.1 % within 1e-5."""
    from voronoi_amd import fileio

    _cli([TET, XYZ, 2, 0], tmp_path)
    want = oracle_chain("cells", 2)
    first = fileio.read_xyz(str(tmp_path / "iter_0.xyz"))
    np.testing.assert_allclose(first, want[0], rtol=0, atol=1e-2)
    got = fileio.read_xyz(str(tmp_path / "out.xyz"))
    d = np.abs(got - want[1]).max(axis=1)
    assert d.max() <= 1.0
    assert (d > 1e-2).mean() <= 1e-3
