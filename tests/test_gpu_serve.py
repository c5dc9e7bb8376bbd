"""The job server: jobs admitted under a memory budget and run together in
one batch must each give the result they get alone (FP64 accumulation, bit
for bit), whatever else runs beside them and whenever they are admitted."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(accum="f64", tet_k=32)


def _jobs():
    from voronoi_amd import meshgen

    specs = [("cube", 6, 200, 3), ("cube", 9, 500, 5), ("cube", 7, 300, 2), ("torus", 14, 300, 4),
             ("cube", 10, 800, 3), ("cube", 5, 150, 6)]
    return [(meshgen.make_problem(k, n, ns, seed=40 + i), it) for i, (k, n, ns, it) in enumerate(specs)]


def _alone(vdn, prob, n_iter):
    with vdn.Engine(**CFG) as e:
        e.set_mesh(*prob[:2])
        e.set_sites(prob[2])
        e.lloyd(n_iter)
        return e.sites()


def test_jobs_in_the_mix_equal_jobs_alone(vdn):
    jobs = _jobs()
    with vdn.Server(**CFG) as s:
        ids = [s.submit(*p, it) for p, it in jobs[:3]]
        s.step()
        ids += [s.submit(*p, it) for p, it in jobs[3:5]]  # arrive while others run
        s.step()
        s.step()
        ids.append(s.submit(*jobs[5][0], jobs[5][1]))
        for _ in range(20):
            r, p = s.step()
            if r == 0 and p == 0:
                break
        for jid, (prob, it) in zip(ids, jobs):
            info = s.job(jid)
            assert info["state"] == vdn.JOB_DONE and info["iters"] == it
            np.testing.assert_array_equal(s.result(jid), _alone(vdn, prob, it))
        assert s.bytes_in_use() == 0


def _footprint(v, t, st, K=90, TK=32):
    """The server's estimate of a running job's device bytes (site-major
    form, volume-pass records of 16 + 4 * 20 bytes, at the automatic vertex
    cap of 17, for one pair in ten)."""
    nv, nt, ns = len(v), len(t), len(st)
    own = 16 * nv + 64 * nt + 16 * ns + 1024
    work = 16 * nv + (16 + 48 + 12 * TK + (16 + 4 * 20) * TK // 10 + 8) * nt + (16 + 4 * K + 64) * ns
    return own + work


def test_budget_admission(vdn):
    jobs = _jobs()
    foot = [_footprint(*p) for p, _ in jobs]
    budget = int(1.5 * max(foot))  # every job fits alone, not every pair does
    with vdn.Server(budget_bytes=budget, **CFG) as s:
        ids = [s.submit(*p, it) for p, it in jobs]
        waited = False
        for _ in range(60):
            r, pend = s.step()
            assert s.bytes_in_use() <= budget
            waited |= r > 0 and pend > 0
            if r == 0 and pend == 0:
                break
        assert waited
        for jid, (p, it) in zip(ids, jobs):
            info = s.job(jid)
            assert info["state"] == vdn.JOB_DONE and info["iters"] == it
            np.testing.assert_array_equal(s.result(jid), _alone(vdn, p, it))


def test_submit_errors(vdn):
    from voronoi_amd import meshgen

    v, t, st = meshgen.make_problem("cube", 4, 60, seed=1)
    with vdn.Server(**CFG) as s:
        with pytest.raises(RuntimeError, match="more sites"):
            s.submit(v, t, st, 2)  # 60 sites <= site_k 90
    v, t, st = meshgen.make_problem("cube", 8, 400, seed=1)
    with vdn.Server(budget_bytes=1 << 16, **CFG) as s:
        with pytest.raises(RuntimeError, match="exceeds the memory budget"):
            s.submit(v, t, st, 2)
