"""vd_knn_dev_noalloc, the kNN entry point that allocates nothing, through the
C ABI: two calls in a row, the second with a smaller input, each equal bit for
bit to vd_knn_dev on the same device buffers (and the first to the oracle's
brute force); sizes past its scratch limits are refused before any work."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _soa(pts, pitch):
    a = np.zeros((3, pitch), dtype=np.float32)
    a[:, :len(pts)] = pts.T
    return a


def _case(vdn, hip, ref, qry, k, pitch_r, pitch_q):
    """(noalloc result, vd_knn_dev result) as (n_query, k) arrays."""
    h_ref, h_q = _soa(ref, pitch_r), _soa(qry, pitch_q)
    h_idx = np.full((k, pitch_q), -7, dtype=np.int32)
    dref, dq, da, db = hip.to_dev(h_ref), hip.to_dev(h_q), hip.to_dev(h_idx), hip.to_dev(h_idx)
    try:
        L = vdn.lib()
        rc = L.vd_knn_dev_noalloc(dref, len(ref), pitch_r, dq, len(qry), pitch_q, 3, k, da, pitch_q, None)
        assert rc == 0
        got = hip.to_host(da, np.empty_like(h_idx))[:, :len(qry)].T
        rc = L.vd_knn_dev(dref, len(ref), pitch_r, dq, len(qry), pitch_q, 3, k, db, pitch_q, None)
        assert rc == 0
        want = hip.to_host(db, np.empty_like(h_idx))[:, :len(qry)].T
        return got, want
    finally:
        for p in (dref, dq, da, db):
            hip.free(p)


def test_knn_dev_noalloc_equals_knn_dev(vdn, O):
    from test_gpu_knn import _Hip

    L = vdn.lib()
    hip = _Hip()
    rng = np.random.default_rng(12)
    # first call: 20,000 reference points, 50,000 queries, k = 40
    ref = rng.random((20000, 3), dtype=np.float32) * 1000
    qry = rng.random((50000, 3), dtype=np.float32) * 1000
    print("call 1: 20000 ref, 50000 queries, k 40", flush=True)
    got, want = _case(vdn, hip, ref, qry, 40, 20480, 50176)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, O.knn(ref, qry, 40, False, grid=True))
    # second call, smaller: 3,000 clustered reference points, 1,000 queries, k = 17
    ref2 = (rng.normal(500, 40, (3000, 3))).astype(np.float32)
    qry2 = rng.random((1000, 3), dtype=np.float32) * 1000
    print("call 2: 3000 ref, 1000 queries, k 17", flush=True)
    got2, want2 = _case(vdn, hip, ref2, qry2, 17, 3072, 1024)
    np.testing.assert_array_equal(got2, want2)
    np.testing.assert_array_equal(got2, O.knn(ref2, qry2, 17, False, grid=False))
    # past the limits: refused by the argument check, before any device work
    h = _soa(ref2, 3072)
    d = hip.to_dev(h)
    try:
        n_big = 262144 + 1
        assert L.vd_knn_dev_noalloc(d, n_big, n_big, d, 10, 3072, 3, 8, d, 3072, None) == -1
        assert L.vd_knn_dev_noalloc(d, 100, 3072, d, 1048577, 1048577, 3, 8, d, 1048577, None) == -1
    finally:
        hip.free(d)
