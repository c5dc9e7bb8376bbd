"""The C++ shims with the reference's kNN signatures (knearests.h,
knncuda.h) on the GPU, through a program built on them.  The command line and
the Lloyd driver are tested against the oracle in test_gpu_dropin.py."""
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

pytestmark = pytest.mark.gpu

BUILD = os.path.join(ROOT, "3d-voronoi-gpu_amd", "_build")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_knearests_and_knncuda_shims(O, tmp_path):
    rng = np.random.default_rng(4)
    n = 15000
    pts = (rng.random((n, 3)) * 1000).astype(np.float32)
    pts.tofile(tmp_path / "pts.f32")
    r = subprocess.run([os.path.join(BUILD, "compat_check"), str(tmp_path / "pts.f32"), str(n), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # kn_print_stats (knearests.cu:434-460): min/max/avg/total, then the histogram of points per cell
    stats = [ln for ln in r.stderr.splitlines() if ln.startswith("Grid:")][0]
    assert stats.endswith(f"total {n}")
    histo = dict(tuple(int(x) for x in ln.strip("[").split("] => ")) for ln in r.stderr.splitlines()
                 if ln.startswith("["))
    assert sum(c * k for c, k in histo.items()) == n and min(histo) == int(stats.split()[4])
    perm = np.fromfile(tmp_path / "kn_perm.u32", dtype=np.uint32).astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "kn_points.f32", dtype=np.float32).reshape(n, 3), pts[perm])
    knn = np.fromfile(tmp_path / "kn_knn.u32", dtype=np.uint32).reshape(n, 90).astype(np.int64)
    want = O.site_knn(pts, 90)
    np.testing.assert_array_equal(perm[knn], want[perm])
    rows = np.fromfile(tmp_path / "dev_knn.i32", dtype=np.int32).reshape(91, n).T
    np.testing.assert_array_equal(rows, O.knn(pts, pts, 91, False))
