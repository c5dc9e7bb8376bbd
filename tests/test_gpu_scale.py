"""Batched and served workloads against the oracle, up to their full sizes.

  * A small batch (cube and torus instances): every instance's per-pair
    status and values, site sums and update equal the oracle run on that
    instance alone, bit for bit (FP64 fixed-point site sums), in both
    tet -> cell launch forms.
  * The batched workload at full size: 256 jittered-cube instances of 6n^3
    tets, n = round(logU[12,45]), tets/10 uniform sites, TK = 64 (3.85e7 tets,
    2.46e9 pairs in one context).  For every instance: its site kNN equals the
    oracle's; a block of its tets has the oracle's candidate rows and, pair by
    pair, the oracle's clipping results; its status counts add up; its cells
    cover the mesh (sum of cell volumes = mesh volume) and the sampled tets'
    pieces cover each tet.  Six instances across the size range are clipped
    whole by the oracle: site sums and new sites bit for bit.
  * The served workload's job mix (jittered cubes of that size law plus
    voxelised tori with Gaussian-mixture sites, 5..50 iterations each) through
    the job server under a budget that forces queueing: every job's final
    sites equal the job run alone, and two jobs (a torus and a cube) equal an
    oracle Lloyd chain of their length, bit for bit.
The moments are the robust decomposition, the accumulation FP64: the
configuration whose accuracy meets 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(accum="f64", moments="robust")


def _instances():
    from voronoi_amd import meshgen

    out = []
    for i, (n, ns) in enumerate([(6, 150), (9, 400), (7, 260), (11, 700), (5, 120)]):
        out.append(meshgen.make_problem("cube", n, ns, seed=100 + i))
    out.append(meshgen.make_problem("torus", 14, 300, seed=9))
    return out


def _mesh_volume(v, t):
    from voronoi_amd import meshgen

    return np.abs(meshgen.orientation(v, t)) / 6.0


@pytest.mark.parametrize("form", ["site", "tet"])
def test_batch_instances_vs_oracle(vdn, O, form):
    probs = _instances()
    with vdn.Engine(tet_k=32, form=form, record_pairs=1, **CFG) as e:
        e.set_batch(probs)
        k, tk, _, _ = e.dims()
        e.knn_step()
        sk, tkn = e.site_knn(), e.tet_knn()
        e.clip_step()
        st, vals = e.pairs()
        mom = e.moments()
        e.update_step()
        new = e.sites()
        off = e.batch_sites
    t0 = 0
    for i, (v, t, s) in enumerate(probs):
        a, b = off[i], off[i + 1]
        p0, p1 = t0 * tk, (t0 + len(t)) * tk
        osk = O.site_knn(s, k)
        otk = O.tet_knn(v, t, s, tk)
        np.testing.assert_array_equal(sk[a:b] - a, osk)
        np.testing.assert_array_equal(tkn[t0:t0 + len(t)] - a, otk)
        r = O.clip_tets(s, osk, v, t, otk, geom="f32", pair_acc="f64", site_mode=O.SITE_FIXED, moments="robust")
        np.testing.assert_array_equal(st[p0:p1].astype(np.int32), r["pair_status"])
        np.testing.assert_array_equal(vals[p0:p1], r["pair_vals"])
        np.testing.assert_array_equal(mom[a:b], r["site_vals"])
        np.testing.assert_array_equal(new[a:b], O.lloyd_update(r["site_vals"], s, fp32=False))
        t0 += len(t)


@pytest.mark.timeout(600)
def test_batched_workload_full_size(vdn, O):
    from voronoi_amd import meshgen

    probs = meshgen.make_batch(256)
    nts = np.array([len(p[1]) for p in probs])
    nss = np.array([len(p[2]) for p in probs])
    print(f"batch: {len(probs)} instances, {nts.sum()} tets, {nss.sum()} sites", flush=True)
    assert nts.min() == 6 * 12 ** 3 and nts.max() <= 6 * 45 ** 3 and nts.sum() * 64 > 2e9
    tb = np.concatenate([[0], np.cumsum(nts)])
    sb = np.concatenate([[0], np.cumsum(nss)])
    BLK = 8
    u = meshgen.uniform24(77, len(probs)).astype(np.float64)
    blk0 = [int(u[i] * (nts[i] - BLK)) for i in range(len(probs))]
    whole = np.argsort(nts)[np.linspace(0, len(probs) - 1, 6).astype(int)]
    with vdn.Engine(tet_k=64, record_pairs=1, **CFG) as e:
        e.set_batch(probs)
        K, TK, NS, NT = e.dims()
        assert (K, TK, NS, NT) == (90, 64, nss.sum(), nts.sum())
        e.knn_step()
        sk = e.site_knn()
        e.clip_step()
        print("batch: clipped", flush=True)
        hist = e.status_hist()
        st = e.pairs(vals=False)
        mom = e.moments()
        blocks = []
        for i in range(len(probs)):
            t0 = int(tb[i]) + blk0[i]
            rows = e.tet_knn(t0, BLK)
            pst, pv = e.pairs(t0 * TK, BLK * TK)
            blocks.append((rows, pst, pv))
        e.update_step()
        new = e.sites()
    assert sum(hist.values()) == NT * TK and hist["triangle_overflow"] == 0 and hist["vertex_overflow"] == 0
    np.testing.assert_array_equal(np.bincount(st, minlength=7), [hist[k] for k in vdn.STATUS_NAMES])
    for i, (v, t, s) in enumerate(probs):
        a, b = int(sb[i]), int(sb[i + 1])
        osk = O.site_knn(s, K)
        np.testing.assert_array_equal(sk[a:b] - a, osk, err_msg=f"instance {i}")
        # status counts of the instance
        h = np.bincount(st[tb[i] * TK:tb[i + 1] * TK], minlength=7)
        assert h.sum() == len(t) * TK and h[4] > 0
        # its cells cover its mesh: robust moments, FP64 sums (tiny pieces
        # dropped and cells short of the security radius stay within 2e-5)
        vt = _mesh_volume(v, t)
        assert abs(mom[a:b, 0].sum() - vt.sum()) <= 2e-5 * vt.sum(), f"instance {i}"
        # a block of tets, pair by pair
        rows, pst, pv = blocks[i]
        sel = slice(blk0[i], blk0[i] + BLK)
        otk = O.tet_knn(v, t[sel], s, TK)
        np.testing.assert_array_equal(rows - a, otk, err_msg=f"instance {i}")
        r = O.clip_tets(s, osk, v, t[sel], otk, geom="f32", pair_acc="f64", site_mode=O.SITE_FIXED,
                        moments="robust")
        np.testing.assert_array_equal(pst.astype(np.int32), r["pair_status"], err_msg=f"instance {i}")
        np.testing.assert_array_equal(pv, r["pair_vals"], err_msg=f"instance {i}")
        # the pieces of each sampled tet cover it (tets with a cell short of its
        # security radius excepted: such a cell may be too large)
        cov = pv[:, 0].reshape(BLK, TK).sum(1)
        ok = ~np.any(pst.reshape(BLK, TK) == 3, axis=1)
        np.testing.assert_allclose(cov[ok], vt[sel][ok], rtol=1e-4)
        if i in whole:
            print(f"instance {i}: {len(t)} tets clipped whole by the oracle", flush=True)
            rr = O.clip_tets(s, osk, v, t, O.tet_knn(v, t, s, TK), geom="f32", pair_acc="f64",
                             site_mode=O.SITE_FIXED, moments="robust", want_pairs=False)
            np.testing.assert_array_equal(mom[a:b], rr["site_vals"], err_msg=f"instance {i}")
            np.testing.assert_array_equal(new[a:b], O.lloyd_update(rr["site_vals"], s, fp32=False))


def _footprint(v, t, st, K=90, TK=64):
    from test_gpu_serve import _footprint as fp

    return fp(v, t, st, K=K, TK=TK)


@pytest.mark.timeout(600)
def test_served_workload_mix(vdn, O):
    from voronoi_amd import meshgen

    jobs = meshgen.serving_jobs(24, 8.0)
    kinds = [j[3] for j in jobs]
    assert kinds.count("torus") >= 2
    foot = [_footprint(*j[1]) for j in jobs]
    budget = int(1.25 * max(foot))
    assert budget < sum(foot) / 3  # the budget forces queueing
    cfg = dict(tet_k=64, **CFG)
    print(f"served: {len(jobs)} jobs ({kinds.count('torus')} tori), budget {budget / 2**30:.2f} GiB", flush=True)
    with vdn.Server(budget_bytes=budget, **cfg) as srv:
        ids = []
        nxt, step, waited = 0, 0, False
        # arrivals in step time: one step stands for 10 ms of the stream
        while True:
            while nxt < len(jobs) and jobs[nxt][0] <= 0.01 * step:
                ids.append(srv.submit(*jobs[nxt][1], jobs[nxt][2]))
                nxt += 1
            r, p = srv.step()
            step += 1
            assert srv.bytes_in_use() <= budget
            waited |= r > 0 and p > 0
            if r == 0 and p == 0 and nxt == len(jobs):
                break
            assert step < 20000
        assert waited
        results = []
        for jid, (_, prob, it, _) in zip(ids, jobs):
            info = srv.job(jid)
            assert info["state"] == vdn.JOB_DONE and info["iters"] == it
            results.append(srv.result(jid))
    print(f"served: {step} steps; comparing each job with the job alone", flush=True)
    for i, (_, (v, t, s), it, kind) in enumerate(jobs):
        with vdn.Engine(**cfg) as e:
            e.set_mesh(v, t)
            e.set_sites(s)
            e.lloyd(it)
            np.testing.assert_array_equal(results[i], e.sites(), err_msg=f"job {i} ({kind})")
    # the smallest torus job and the smallest cube job against the oracle chain
    for kind in ("torus", "cube"):
        i = min((j for j in range(len(jobs)) if kinds[j] == kind), key=lambda j: len(jobs[j][1][1]) * jobs[j][2])
        _, (v, t, s), it, _ = jobs[i]
        print(f"job {i} ({kind}, {len(t)} tets, {it} iterations) against the oracle chain", flush=True)
        for _ in range(it):
            s, _ = O.lloyd_step(v, t, s, k=90, tk=64, geom="f32", pair_acc="f64", site_mode=O.SITE_FIXED,
                                fp32_update=False, moments="robust")
        np.testing.assert_array_equal(results[i], s, err_msg=f"job {i} ({kind})")
