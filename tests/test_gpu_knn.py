"""Exact kNN on the GPU against the oracle (bit-exact index lists).

The oracle's kNN is itself pinned against brute force in test_oracle.py.
"""
import numpy as np
import pytest
from conftest import oracle_knn, problem

pytestmark = pytest.mark.gpu


def _engine(vdn, verts, tets, sites, **cfg):
    e = vdn.Engine(**cfg)
    e.set_mesh(verts, tets)
    e.set_sites(sites)
    e.knn_step()
    return e


@pytest.mark.parametrize("case", [("cube", 8, 300), ("cube", 20, 2000), ("cube", 30, 16200), ("torus", 24, 3000)])
def test_site_and_tet_knn_bit_exact(vdn, case):
    kind, n, ns = case
    verts, tets, sites = problem(kind, n, ns)
    sk, tk = oracle_knn(kind, n, ns, 90, 64)
    with _engine(vdn, verts, tets, sites, tet_k=64) as e:
        k, t, _, _ = e.dims()
        assert (k, t) == (sk.shape[1], tk.shape[1])
        g_sk, g_tk = e.site_knn(), e.tet_knn()
    print(f"{case}: site knn mismatched rows {np.sum(np.any(g_sk != sk, axis=1))}, "
          f"tet knn mismatched rows {np.sum(np.any(g_tk != tk, axis=1))}", flush=True)
    np.testing.assert_array_equal(g_sk, sk)
    np.testing.assert_array_equal(g_tk, tk)


def test_knn_full_size_config2(vdn, O):
    """55^3 x 6 tets, 100k sites: the single-solve configuration."""
    verts, tets, sites = problem("cube", 55, 100000)
    sk = O.site_knn(sites, 90)
    with _engine(vdn, verts, tets, sites, tet_k=64) as e:
        g_sk, g_tk = e.site_knn(), e.tet_knn()
    np.testing.assert_array_equal(g_sk, sk)
    # candidate lists: exact on a sample of tets (the oracle's grid kNN on all
    # 998k tets is checked by the smaller cases)
    rng = np.random.default_rng(0)
    idx = rng.choice(len(tets), 20000, replace=False)
    cent = O.tet_centroids(verts, tets)[idx]
    ref = O.knn(sites, cent, 64, False)
    np.testing.assert_array_equal(g_tk[idx], ref)


def test_lattice_ties_break_by_index(vdn, O):
    """Sites on a lattice: many equal distances; ties go to the lower index."""
    g = np.arange(12, dtype=np.float32) * 40.0 + 50.0
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    sites = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    rng = np.random.default_rng(1)
    sites = sites[rng.permutation(len(sites))]
    verts, tets, _ = problem("cube", 6, 10)
    with _engine(vdn, verts, tets, sites, tet_k=27) as e:
        g_sk, g_tk = e.site_knn(), e.tet_knn()
    np.testing.assert_array_equal(g_sk, O.knn(sites, sites, 90, True, grid=False))
    cent = O.tet_centroids(verts, tets)
    np.testing.assert_array_equal(g_tk, O.knn(sites, cent, 27, False, grid=False))


def test_k_larger_than_sites(vdn, O):
    verts, tets, sites = problem("cube", 4, 5)
    with _engine(vdn, verts, tets, sites, tet_k=64) as e:
        k, tk, ns, _ = e.dims()
        assert k == 4 and tk == 5
        np.testing.assert_array_equal(e.site_knn(), O.knn(sites, sites, 4, True, grid=False))
        np.testing.assert_array_equal(e.tet_knn(), O.knn(sites, O.tet_centroids(verts, tets), 5, False, grid=False))


class _Hip:
    """Device buffers through the HIP runtime the library itself links (torch
    ships a different HIP runtime and cannot share the process)."""

    def __init__(self):
        import ctypes

        self.ct = ctypes
        self.L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def to_dev(self, a):
        p = self.ct.c_void_p()
        assert self.L.hipMalloc(self.ct.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(self.ct.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, a):
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(a.ctypes.data_as(self.ct.c_void_p), p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self, p):
        self.L.hipFree(p)


def test_knn_dev_reference_layout(vdn, O):
    """vd_knn_dev: pitched SoA rows in and out, like knn_cuda_global_dev."""
    rng = np.random.default_rng(3)
    ref = rng.random((3000, 3), dtype=np.float32) * 1000
    qry = rng.random((1000, 3), dtype=np.float32) * 1000
    k = 17
    pitch_r, pitch_q = 3072, 1024
    h_ref = np.zeros((3, pitch_r), dtype=np.float32)
    h_ref[:, :3000] = ref.T
    h_q = np.zeros((3, pitch_q), dtype=np.float32)
    h_q[:, :1000] = qry.T
    h_idx = np.full((k, pitch_q), -7, dtype=np.int32)
    hip = _Hip()
    dref, dq, didx = hip.to_dev(h_ref), hip.to_dev(h_q), hip.to_dev(h_idx)
    try:
        rc = vdn.lib().vd_knn_dev(dref, 3000, pitch_r, dq, 1000, pitch_q, 3, k, didx, pitch_q, None)
        assert rc == 0
        got = hip.to_host(didx, np.empty_like(h_idx))[:, :1000].T
        np.testing.assert_array_equal(got, O.knn(ref, qry, k, False, grid=False))
        assert vdn.lib().vd_knn_dev(dref, 3000, pitch_r, dq, 1000, pitch_q, 2, k, didx, pitch_q, None) != 0
    finally:
        for p in (dref, dq, didx):
            hip.free(p)


def test_kn_problem_sorted_ids(vdn, O):
    """kn_prepare/kn_solve: lists in grid order, entries are grid positions."""
    _, _, sites = problem("cube", 10, 20000)
    kn = vdn.KnProblem(sites, 90)
    kn.solve()
    perm = kn.permutation().astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(len(sites)))
    np.testing.assert_array_equal(kn.points(), sites[perm])
    lists = kn.knearests().astype(np.int64)
    want = O.site_knn(sites, 90)
    np.testing.assert_array_equal(perm[lists], want[perm])
    info = kn.info()
    assert info.allocated_points == len(sites) and info.dimx >= 1
    kn.close()


@pytest.mark.parametrize("kind,n,ns,tk", [("cube", 12, 3000, 64), ("torus", 20, 2500, 96)])
def test_knn_bounded_by_previous_rows(vdn, O, kind, n, ns, tk):
    """From the second kNN step on, each query starts from the k points of its
    previous row; the rows must still equal the oracle's exact kNN, after Lloyd
    moves, after the sites are replaced by an unrelated set of the same size
    (the old rows then bound loosely), and in a batch."""
    from voronoi_amd import meshgen

    verts, tets, sites = problem(kind, n, ns)
    other = meshgen.make_problem(kind, n, ns, seed=99)[2]
    with _engine(vdn, verts, tets, sites, tet_k=tk, accum="f64") as e:
        for step in range(4):
            if step == 3:
                e.set_sites(other)
            cur = e.sites()
            e.knn_step()
            print(f"{kind} step {step}", flush=True)
            np.testing.assert_array_equal(e.site_knn(), O.site_knn(cur, 90))
            np.testing.assert_array_equal(e.tet_knn(), O.tet_knn(verts, tets, cur, tk))
            e.clip_step()
            e.update_step()
    probs = [meshgen.make_problem("cube", 7, 400, seed=5), meshgen.make_problem(kind, n, ns, seed=6)]
    with vdn.Engine(tet_k=tk, accum="f64") as e:
        e.set_batch(probs)
        off = e.batch_sites
        for step in range(3):
            cur = e.sites()
            e.knn_step()
            sk, tkn = e.site_knn(), e.tet_knn()
            t0 = 0
            for i, (v, t, _) in enumerate(probs):
                s = cur[off[i]:off[i + 1]]
                np.testing.assert_array_equal(sk[off[i]:off[i + 1]] - off[i], O.site_knn(s, e.dims()[0]))
                np.testing.assert_array_equal(tkn[t0:t0 + len(t)] - off[i], O.tet_knn(v, t, s, tk))
                t0 += len(t)
            e.clip_step()
            e.update_step()
