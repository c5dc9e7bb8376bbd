"""Batches of independent instances in one context: every instance's kNN,
moments and Lloyd result must be the ones it gets alone (bit for bit in the
FP64 accumulation mode, whose site sums do not depend on the order of the
atomics)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instances():
    from voronoi_amd import meshgen

    out = []
    for i, (n, ns) in enumerate([(6, 150), (9, 400), (7, 260), (11, 700), (5, 120)]):
        out.append(meshgen.make_problem("cube", n, ns, seed=100 + i))
    out.append(meshgen.make_problem("torus", 14, 300, seed=9))
    return out


@pytest.mark.parametrize("form", ["site", "tet"])
def test_batch_equals_instances_alone(vdn, form):
    probs = _instances()
    cfg = dict(accum="f64", tet_k=32, form=form, record_pairs=0)
    with vdn.Engine(**cfg) as e:
        e.set_batch(probs)
        k, tk, ns, nt = e.dims()
        assert (k, tk) == (min(90, min(len(p[2]) for p in probs) - 1), 32)
        assert ns == sum(len(p[2]) for p in probs) and nt == sum(len(p[1]) for p in probs)
        e.knn_step()
        bknn = e.site_knn()
        btk = e.tet_knn()
        e.clip_step()
        bmom = e.moments()
        e.update_step()
        e.lloyd(1)
        bsites = e.sites()
        off = e.batch_sites
    t0 = 0
    for i, (v, t, s) in enumerate(probs):
        a, b = off[i], off[i + 1]
        with vdn.Engine(**cfg) as e:
            e.set_mesh(v, t)
            e.set_sites(s)
            e.knn_step()
            assert e.dims()[0] == k
            np.testing.assert_array_equal(bknn[a:b] - a, e.site_knn())
            np.testing.assert_array_equal(btk[t0:t0 + len(t)] - a, e.tet_knn())
            e.clip_step()
            np.testing.assert_array_equal(bmom[a:b], e.moments())
            e.update_step()
            e.lloyd(1)
            np.testing.assert_array_equal(bsites[a:b], e.sites())
        t0 += len(t)


def test_batch_site_knn_matches_oracle(vdn, O):
    probs = _instances()
    with vdn.Engine(tet_k=32) as e:
        e.set_batch(probs)
        k = e.dims()[0]
        e.knn_step()
        bknn = e.site_knn()
        off = e.batch_sites
    for i, (_, _, s) in enumerate(probs):
        got = bknn[off[i]:off[i + 1]] - off[i]
        np.testing.assert_array_equal(got, O.site_knn(np.asarray(s, np.float32), k))


def test_batch_needs_one_candidate_count(vdn):
    from voronoi_amd import meshgen

    # reference heuristic (voronoi.cu:936): 90 candidates for 600 sites on 384 tets, 20 for 100 on 6000
    probs = [meshgen.make_problem("cube", 4, 600, seed=1), meshgen.make_problem("cube", 10, 100, seed=2)]
    with vdn.Engine(tet_k=0) as e:
        e.set_batch(probs)
        assert e.dims()[1] == -1
        with pytest.raises(RuntimeError, match="candidate counts"):
            e.knn_step()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replays_equal_eager_iterations(vdn, batched):
    """Each replay of the captured iteration, with eager work and host
    synchronisation between replays, gives the eager iteration's sites."""
    from voronoi_amd import meshgen

    probs = _instances()[:3] if batched else [meshgen.make_problem("cube", 10, 600, seed=3)]
    cfg = dict(accum="f64", tet_k=32)
    ref = []
    with vdn.Engine(**cfg) as e:
        e.set_batch(probs) if batched else (e.set_mesh(*probs[0][:2]), e.set_sites(probs[0][2]))
        for _ in range(5):
            e.lloyd(1)
            ref.append(e.sites())
    with vdn.Engine(**cfg) as g:
        g.set_batch(probs) if batched else (g.set_mesh(*probs[0][:2]), g.set_sites(probs[0][2]))
        g.lloyd_graph(2)  # eager sizing iteration, then the captured one
        g.sync()
        np.testing.assert_array_equal(g.sites(), ref[1])
        for i in range(2, 5):
            g.lloyd_graph(1)
            g.sync()
            np.testing.assert_array_equal(g.sites(), ref[i])
            g.moments()  # eager work on the same stream between replays
        t = g.timing()
        assert t["total_ms"] > 0


def test_autotune_keeps_results(vdn):
    """The tuned launch parameters change no result (FP64 accumulation)."""
    from voronoi_amd import meshgen

    v, t, s = meshgen.make_problem("cube", 12, 1500, seed=5)
    with vdn.Engine(accum="f64", tet_k=32) as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        e.lloyd(2)
        want = e.sites()
    with vdn.Engine(accum="f64", tet_k=32) as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        chosen = e.autotune(iters=1)
        assert chosen["site_tcap"] in (0, 12, 16, 20, 24) and chosen["cull_depth"] in (8, 16, 32, 64)
        assert chosen["clip_blocks"] in (0, 1536, 3072, 4608) and chosen["knn_waves"] in (0, 2, 4, 8)
        np.testing.assert_array_equal(e.sites(), s.astype(np.float32))  # restored
        e.lloyd(2)
        np.testing.assert_array_equal(e.sites(), want)


@pytest.mark.parametrize("knn_waves,clip_blocks", [(2, 1536), (8, 4608)])
def test_block_sizes_keep_results(vdn, knn_waves, clip_blocks):
    """The kNN workgroup size and the site-major launch's workgroup count
    change no result (FP64 accumulation, two iterations)."""
    from voronoi_amd import meshgen

    v, t, s = meshgen.make_problem("cube", 12, 1500, seed=5)
    out = []
    for cfg in (dict(), dict(knn_waves=knn_waves, clip_blocks=clip_blocks)):
        with vdn.Engine(accum="f64", tet_k=32, **cfg) as e:
            e.set_mesh(v, t)
            e.set_sites(s)
            e.lloyd(2)
            out.append((e.sites(), e.site_knn(), e.tet_knn()))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("moments", ["robust", "reference"])
def test_fresh_site_depth_keeps_results(vdn, moments):
    """The first site-major clipping after set_sites uses 64-bisector tables
    (automatic depth), later ones 40: the sites of every iteration, and of a
    chain restarted by set_sites, equal those at a fixed depth of 40 and of
    site_k (no pair recomputed deeper) bit for bit."""
    from voronoi_amd import meshgen

    v, t, s = meshgen.make_problem("cube", 16, 4000, seed=11)
    runs = []
    for depth in (0, 40, 90):
        with vdn.Engine(accum="f64", moments=moments, tet_k=64, site_k=90, site_depth=depth) as e:
            e.set_mesh(v, t)
            got = []
            for _ in range(2):  # the second chain restarts from the same sites
                e.set_sites(s)
                for _ in range(3):
                    e.lloyd(1)
                    got.append(e.sites())
            runs.append(got)
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("tol", [0.0, 1e-5])
def test_autotune_k(vdn, tol):
    """vd_autotune_k keeps only smaller neighbour/candidate counts whose sites
    after `iters` iterations stay within tol * |site| of the configured
    counts' (tol 0: the same bits); the reported error is checked here by
    running both configurations."""
    from voronoi_amd import meshgen

    v, t, s = meshgen.make_problem("cube", 12, 1500, seed=5)
    base = dict(accum="f64", tet_k=64, site_k=90)
    with vdn.Engine(**base) as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        e.lloyd(2)
        want = e.sites()
    with vdn.Engine(**base) as e:
        e.set_mesh(v, t)
        e.set_sites(s)
        chosen, err = e.autotune_k(iters=2, tol=tol)
        assert chosen["tet_k"] <= 64 and chosen["site_k"] <= 90 and 0.0 <= err <= tol
        np.testing.assert_array_equal(e.sites(), s.astype(np.float32))  # restored
        e.lloyd(2)
        got = e.sites()
    d = np.linalg.norm(got.astype(np.float64) - want, axis=1) / np.linalg.norm(want.astype(np.float64), axis=1)
    assert d.max() <= tol
    assert abs(d.max() - err) <= 1e-12
    print(f"tol {tol}: tet_k {chosen['tet_k']} site_k {chosen['site_k']} err {err:.3g}")
