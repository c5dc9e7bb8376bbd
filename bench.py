"""Lloyd-iteration throughput of the clipped-Voronoi engine on one MI355X.

Workload (the single-solve configuration): unit cube split into 55^3 cubes x 6
Kuhn tets (998,250 tets, vertices jittered by U(-0.1,0.1)*h), 100,000 uniform
FP32 sites (seeded), candidate sites per tet TK = 64, neighbours per site
K = 90, 10 Lloyd iterations from the seeded sites.  A step is one whole Lloyd
iteration on the device: grid kNN of the sites, candidate kNN of the tets,
clipping of all n_tets*TK (tet, site) pairs with volume/moment accumulation,
site update.  The timed steps run the 10-iteration chain from the seeded sites
(a longer run restarts the chain from them every 10 steps); inputs are resident
in HBM before the timed region.

Reported:
  value           Lloyd iterations/s of the fast form: site-major launch, FP32
                  clipping, FP64 accumulation, robust moment decomposition --
                  the form whose volumes and centroids are within 1e-5 of the
                  FP64 oracle (see accuracy)
  parity_form     the same launch with the reference's moment decomposition
                  (bit for bit the reference algorithm; its FP32 volumes miss
                  1e-5 against the same algorithm in FP64)
  plain_*         the plain form: tet-major launch (one lane per pair, the
                  reference's order), FP32 accumulation with FP32 atomics
  accuracy        max / mean relative error of cell volume, total volume and
                  centroid against the FP64 oracle on iteration 1, for the
                  single-solve workload and the 48,000-tet one
  roofline        clipping kernels: the reference algorithm's FLOPs, counted
                  by the oracle on the timed iterations' own sites, over the
                  kernels' event time, against the FP32 vector peak
  roofline_knn    kNN kernels: algorithmic bytes over their event time vs HBM
  cpu_baseline    the oracle's FP32 restatement of the reference's iteration
                  (site kNN, candidate kNN, kernel A) on the host cores
  batched         256 independent jittered-cube instances (6n^3 tets,
                  n = round(logU[12,45]), tets/10 sites, TK = 64; 3.85e7 tets,
                  2.46e9 pairs) in one context, each iteration a replay of one
                  HIP graph
  served          a Poisson stream of jobs of that size mix plus 5% torus jobs,
                  5..50 iterations each, through the job server (200 GB
                  admission budget); p50/p99 job latency
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-voronoi-gpu_amd"))

METRIC = "Lloyd iters/s at 1M tets,100k sites,k=64; batched/served p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
N_CUBE, N_SITES, TK, K, CHAIN = 55, 100000, 64, 90, 10
FAST = dict(form="site", accum="f64", moments="robust")

# FLOP weights per event of the reference's kernel A (add/mul = 1, fma = 2):
# pair setup (4 face planes), bisector, conflict test (det4x4), security-radius
# term per vertex (vertex position + distance), per security check, moment
# integration per vertex.
W_PAIR, W_CLIP, W_CONFLICT, W_SECURE_VTX, W_SECURE_CHECK, W_ACC_VTX = 80, 14, 45, 67, 9, 545

TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "clip_site_998k_tets_100k_sites_hbm_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per iteration of the clipping kernels at this workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O


def run_chain(native, verts, tets, sites, steps, warmup, chain=CHAIN, **cfg):
    """`warmup` untimed iterations, then `steps` timed ones: the chain of
    `chain` iterations from the seeded sites, restarted from them every
    `chain` steps.  Returns wall seconds, per-iteration phase times and the
    status counts of the last iteration."""
    e = native.Engine(0, site_k=K, tet_k=TK, **cfg)
    e.set_mesh(verts, tets)
    e.set_sites(sites)
    for _ in range(warmup):
        e.lloyd(1)
    e.set_sites(sites)
    e.sync()
    e.timing_reset()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chain, steps - done)
        e.lloyd(n)
        done += n
        if done < steps:
            e.set_sites(sites)  # next chain from the seeded sites
    e.sync()
    wall = time.perf_counter() - t0
    # the engine keeps the phase times of the last 256 iterations: the list
    # starts at a chain's first iteration
    n = steps if steps <= 256 else steps - -(-(steps - 256) // chain) * chain
    tl = e.timing_list(n)
    hist = e.status_hist()
    e.close()
    return wall, tl, hist


def chain_sites(native, verts, tets, sites, n, **cfg):
    """The sites entering each of the chain's first n iterations (untimed)."""
    out = [np.asarray(sites, np.float32)]
    with native.Engine(0, site_k=K, tet_k=TK, **cfg) as e:
        e.set_mesh(verts, tets)
        e.set_sites(sites)
        for _ in range(n - 1):
            e.lloyd(1)
            out.append(e.sites())
    return out


def mean_phases(tl, idx=None):
    rows = [tl[i] for i in (idx if idx is not None else range(len(tl)))]
    return {k: float(np.mean([r[k] for r in rows])) for k in rows[0]}


def gpu_moments(native, verts, tets, sites, **cfg):
    """Per-site (vol, mx, my, mz) of iteration 1."""
    with native.Engine(0, site_k=K, tet_k=TK, **cfg) as e:
        e.set_mesh(verts, tets)
        e.set_sites(sites)
        e.knn_step()
        e.clip_step()
        return e.moments()


def errors(a, b):
    """Relative errors of GPU site sums `a` against the FP64 oracle's `b`."""
    ok = b[:, 0] > 0
    a, b = a[ok], b[ok]
    vol = np.abs(a[:, 0] - b[:, 0]) / b[:, 0]
    ca, cb = a[:, 1:] / a[:, :1], b[:, 1:] / b[:, :1]
    d = np.linalg.norm(ca - cb, axis=1)
    cen = d / np.linalg.norm(cb, axis=1)
    cen_cell = d / np.cbrt(b[:, 0])
    tot = abs(a[:, 0].sum() - b[:, 0].sum()) / b[:, 0].sum()
    f = lambda x: float("%.3e" % x)  # noqa: E731
    return {"cell_volume": {"max": f(vol.max()), "mean": f(vol.mean())}, "total_volume": f(tot),
            "centroid": {"max": f(cen.max()), "mean": f(cen.mean())},
            "centroid_per_cell_size": {"max": f(cen_cell.max()), "mean": f(cen_cell.mean())},
            "sites": int(ok.sum())}


def accuracy(native, forms):
    """Iteration-1 site sums of each form on the single-solve workload and on
    the 48,000-tet one against the all-FP64 oracle of the same moment
    decomposition."""
    from voronoi_amd import meshgen

    O = oracle()
    out = {}
    for name, (n, ns) in (("998k_tets_100k_sites", (N_CUBE, N_SITES)), ("48k_tets_2k_sites", (20, 2000))):
        v, t, s = meshgen.make_problem("cube", n, ns, seed=7)
        sk = O.site_knn(s, K)
        tk = O.tet_knn(v, t, s, TK)
        ref = {}
        out[name] = {}
        for form, cfg in forms.items():
            m = cfg["moments"]
            if m not in ref:
                t0 = time.perf_counter()
                ref[m] = O.clip_tets(s, sk, v, t, tk, geom="f64", pair_acc="f64", site_mode=O.SITE_F64, moments=m,
                                     want_pairs=False)["site_vals"]
                log(f"accuracy: FP64 oracle, {name}, {m} moments ({time.perf_counter() - t0:.1f}s)")
            out[name][form] = errors(gpu_moments(native, v, t, s, **cfg), ref[m])
            log(f"accuracy: {name} {form}: {out[name][form]}")
    return out


def flops_per_pair(ev):
    f = (W_PAIR * ev["pairs"] + W_CLIP * ev["clips"] + W_CONFLICT * ev["conflict_tests"]
         + W_SECURE_VTX * ev["secure_evals"] + W_SECURE_CHECK * ev["secure_checks"] + W_ACC_VTX * ev["acc_tris"])
    return f / max(ev["pairs"], 1)


def count_flops(verts, tets, site_sets):
    """The reference algorithm's clipping FLOPs of each iteration (oracle
    event counts of its FP32 kernel A on that iteration's sites), and the
    first iteration's oracle times, which are the CPU baseline."""
    O = oracle()
    flops, base = [], None
    for i, s in enumerate(site_sets):
        t0 = time.perf_counter()
        sk = O.site_knn(s, K)
        t1 = time.perf_counter()
        tk = O.tet_knn(verts, tets, s, TK)
        t2 = time.perf_counter()
        r = O.clip_tets(s, sk, verts, tets, tk, geom="f32", pair_acc="f32", site_mode=O.SITE_F32, want_pairs=False)
        t3 = time.perf_counter()
        flops.append(flops_per_pair(r["events"]) * len(tets) * TK)
        if i == 0:
            base = dict(site_knn=t1 - t0, tet_knn=t2 - t1, clip=t3 - t2, threads=O.nthreads())
        log(f"flops: iteration {i + 1}: {flops[-1]:.4e} ({t3 - t0:.1f}s)")
    return flops, base


def run_batch(native, steps):
    """The batched workload: 256 independent jittered-cube instances of
    6n^3 tets, n = round(logU[12, 45]), tets/10 uniform sites each, TK = 64,
    solved together in one context, one Lloyd iteration per graph replay."""
    from voronoi_amd import meshgen

    t0 = time.perf_counter()
    probs = meshgen.make_batch(256)
    n_t = sum(len(p[1]) for p in probs)
    n_s = sum(len(p[2]) for p in probs)
    log(f"batch: {len(probs)} instances, {n_t} tets, {n_s} sites ({time.perf_counter() - t0:.1f}s to generate)")
    e = native.Engine(0, site_k=K, tet_k=TK, **{k: v for k, v in FAST.items() if k != "form"})
    e.set_batch(probs)
    del probs
    e.lloyd_graph(2)  # eager sizing iteration + capture + one replay
    e.sync()
    e.timing_reset()
    t0 = time.perf_counter()
    e.lloyd_graph(steps)
    e.sync()
    wall = time.perf_counter() - t0
    tl = e.timing_list(min(steps, 256))
    e.close()
    dev_ms = float(np.mean([r["total_ms"] for r in tl]))
    log(f"batch: {1e3 * wall / steps:.1f} ms/iter wall, {dev_ms:.1f} ms device per replayed iteration")
    return {"instances": 256, "n_tets": n_t, "n_sites": n_s, "pairs_per_iter": n_t * TK, "steps": steps,
            "ms_per_iter": round(1e3 * wall / steps, 3), "device_ms_per_iter": round(dev_ms, 3),
            "iters_per_s": round(steps / wall, 4), "pairs_per_s": round(n_t * TK * steps / wall, 1),
            "launch": "one context, hipGraph replay of a whole iteration",
            "accumulation": "fp64", "moments": FAST["moments"]}


def run_serving(native, n_jobs, rate, budget_gb=200):
    """Served workload: a Poisson stream of jobs (the batched workload's size
    mix plus 5% torus jobs, 5..50 iterations each) through the job server with
    a device-memory admission budget, in real time.  Job latency = submission
    to completion."""
    from voronoi_amd import meshgen

    t0 = time.perf_counter()
    jobs = meshgen.serving_jobs(n_jobs, rate)
    log(f"serving: {n_jobs} jobs at {rate}/s generated ({time.perf_counter() - t0:.1f}s)")
    pair_iters = sum(len(p[1]) * TK * it for _, p, it, _ in jobs)
    s = native.Server(0, budget_bytes=budget_gb << 30, site_k=K, tet_k=TK,
                      **{k: v for k, v in FAST.items() if k != "form"})
    ids = []
    nxt = 0
    steps = 0
    start = time.perf_counter()
    while True:
        now = time.perf_counter() - start
        while nxt < n_jobs and jobs[nxt][0] <= now:
            ids.append(s.submit(*jobs[nxt][1], jobs[nxt][2]))
            nxt += 1
        running, pending = s.step()
        steps += 1
        if running == 0 and pending == 0:
            if nxt == n_jobs:
                break
            time.sleep(max(0.0, jobs[nxt][0] - (time.perf_counter() - start)))
    span = time.perf_counter() - start
    lat = np.array([s.job(j)["wait_ms"] + s.job(j)["run_ms"] for j in ids])
    wait = np.array([s.job(j)["wait_ms"] for j in ids])
    s.close()
    offered = pair_iters / span
    res = {"jobs": n_jobs, "arrival_rate_per_s": rate, "budget_gb": budget_gb, "steps": steps,
           "span_s": round(span, 3), "jobs_per_s": round(n_jobs / span, 3),
           "pairs_per_s": round(offered, 1),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                          "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)},
           "wait_ms": {"p50": round(float(np.percentile(wait, 50)), 2), "p99": round(float(np.percentile(wait, 99)), 2)}}
    log(f"serving: {res}")
    return res


def device_copy(mb):
    """Device-to-device copy of `mb` MiB through the HIP runtime the engine links:
    measured HBM bandwidth, and the calibration dispatch for PMC byte counters."""
    import ctypes

    L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert L.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    L.hipMemset(a, 1, ctypes.c_size_t(n))
    L.hipDeviceSynchronize()
    best = 1e9
    for _ in range(5):
        t0 = time.perf_counter()
        assert L.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # device to device
        L.hipDeviceSynchronize()
        best = min(best, time.perf_counter() - t0)
    L.hipFree(a)
    L.hipFree(b)
    return 2 * n / best / 1e9


def knn_bytes(n_tets, n_verts, n_sites):
    """Algorithmic HBM bytes of one iteration's kNN: site kNN
    n_s*(12 read + 12 sorted + 4 permutation + 4K lists) + 8 per grid cell,
    candidate kNN n_t*(16 tet ids + 4TK lists) + 12 n_verts + 12 n_s."""
    cells = n_sites / 3.1
    return n_sites * (12 + 12 + 4 + 4 * K) + 8 * cells + n_tets * (16 + 4 * TK) + 12 * n_verts + 12 * n_sites


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3, help="timed iterations of the plain and cell forms (0 skips)")
    ap.add_argument("--parity-steps", type=int, default=10, help="timed iterations of the parity form (0 skips)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle legs (FLOP count, CPU baseline)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0, help="also time a device copy of this size (MiB)")
    ap.add_argument("--batch-steps", type=int, default=10, help="graph replays of the 256-instance batch (0 skips)")
    ap.add_argument("--serve-jobs", type=int, default=120, help="jobs of the served workload (0 skips)")
    ap.add_argument("--serve-rate", type=float, nargs="+", default=[8.0, 16.0], help="job arrivals per second")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is not built)"}))
        sys.exit(2)

    from voronoi_amd import meshgen, native

    t0 = time.perf_counter()
    verts, tets, sites = meshgen.make_problem("cube", N_CUBE, N_SITES, seed=7)
    log(f"workload: {len(tets)} tets, {len(verts)} verts, {len(sites)} sites ({time.perf_counter() - t0:.1f}s)")
    n_pairs = len(tets) * TK

    wall, tl, hist = run_chain(native, verts, tets, sites, args.steps, args.warmup, **FAST)
    ms_step = 1e3 * wall / args.steps
    value = args.steps / wall
    it_idx = [[j for j in range(len(tl)) if j % CHAIN == i] for i in range(min(CHAIN, len(tl)))]
    ph = mean_phases(tl)
    ph1 = mean_phases(tl, it_idx[0])
    log(f"fast: {ms_step:.3f} ms/iter wall; per-iter device: knn {ph['knn_ms']:.3f} pairs {ph['pairs_ms']:.3f} "
        f"clip {ph['clip_ms']:.3f} (overflow {ph['overflow_ms']:.3f}) update {ph['update_ms']:.3f}; iteration 1 "
        f"{ph1['total_ms']:.3f} ms; status {hist}")

    parity = None
    if args.parity_steps > 0:
        cfg = dict(FAST, moments="reference")
        pw, ptl, _ = run_chain(native, verts, tets, sites, args.parity_steps, 1, **cfg)
        parity = {"iters_per_s": round(args.parity_steps / pw, 4), "ms_per_iter": round(1e3 * pw / args.parity_steps, 3),
                  "moments": "reference", "iteration_ms": [round(r["total_ms"], 3) for r in ptl[:CHAIN]]}
        log(f"parity form: {parity}")

    plain = cells = None
    if args.plain_steps > 0:
        cw, ctl, _ = run_chain(native, verts, tets, sites, args.plain_steps, 1, form="cells", accum="f64",
                               moments="robust")
        cells = {"iters_per_s": round(args.plain_steps / cw, 4),
                 "clip_ms": round(mean_phases(ctl)["clip_ms"], 3)}
        log(f"cells form: {cells}")
        pw, ptl, _ = run_chain(native, verts, tets, sites, args.plain_steps, 1, form="tet", accum="f32",
                               moments="robust")
        same = [j for j in range(len(tl)) if j % CHAIN < args.plain_steps]
        fast_same = float(np.mean([tl[j]["total_ms"] for j in same]))
        plain_ms = 1e3 * pw / args.plain_steps
        plain = {"iters_per_s": round(args.plain_steps / pw, 4), "ms_per_iter": round(plain_ms, 3),
                 "clip_ms": round(mean_phases(ptl)["clip_ms"], 3),
                 "iterations": f"1..{args.plain_steps} of the chain",
                 "speedup_same_iterations": round(plain_ms / fast_same, 2)}
        log(f"plain: {plain}")

    acc = None
    if not args.no_accuracy:
        acc = accuracy(native, {"fast": FAST, "parity_form": dict(FAST, moments="reference"),
                                "plain": dict(form="tet", accum="f32", moments="robust")})

    batch = run_batch(native, args.batch_steps) if args.batch_steps > 0 else None
    served = [run_serving(native, args.serve_jobs, r) for r in args.serve_rate] if args.serve_jobs > 0 else None
    copy_gbs = device_copy(args.copy_mb) if args.copy_mb > 0 else None
    if copy_gbs:
        log(f"device copy of {args.copy_mb} MiB: {copy_gbs:.0f} GB/s (read + write)")

    roof = cpu = None
    knn_b = knn_bytes(len(tets), len(verts), len(sites))
    knn_s = ph["knn_ms"] / 1e3
    roof_knn = {"bound": "hbm", "kernel": "grid build + k_knn_query (sites, tet centroids)",
                "achieved": round(knn_b / knn_s / 1e9, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(knn_b / knn_s / 1e9 / PEAK_HBM_GBS, 5), "bytes_per_iteration": knn_b,
                "kernel_ms": round(ph["knn_ms"], 4)}
    if not args.no_cpu:
        sets = chain_sites(native, verts, tets, sites, min(CHAIN, args.steps), **FAST)
        flops, base = count_flops(verts, tets, sets)
        # the three launches of the clipping (emptiness pass, clipping, volume
        # pass) on the engine's stream: the clip phase minus the exposed part
        # of the overflow re-clip beside the volume pass
        kern_ms = [tl[j]["clip_ms"] - tl[j]["overflow_ms"] for j in range(len(tl))]
        f_timed = [flops[j % CHAIN] for j in range(len(tl))]
        achieved = sum(f_timed) / (sum(kern_ms) / 1e3) / 1e12
        roof = {"bound": "valu", "kernel": "k_cull_site + k_clip_site + k_site_moments (emptiness test; clipping; "
                                           "volumes, moments, accumulation)",
                "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": profiled_traffic(),
                "flop_per_launch": round(float(np.mean(f_timed)), 1), "kernel_ms": round(float(np.mean(kern_ms)), 4),
                "flop_per_iteration": [round(f, 1) for f in flops],
                "kernel_ms_per_iteration": [round(float(np.mean([kern_ms[j] for j in ix])), 4) for ix in it_idx]}
        t_iter = base["site_knn"] + base["tet_knn"] + base["clip"]
        cpu = {"value": round(1.0 / t_iter, 5), "unit": "iters/s", "cores": base["threads"], "kind": "port",
               "sample": f"iteration 1 of the chain, whole mesh ({n_pairs} pairs): site kNN {base['site_knn']:.2f}s + "
                         f"candidate kNN {base['tet_knn']:.2f}s + FP32 kernel A {base['clip']:.2f}s"}
        log(f"roofline: {roof}\ncpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "jittered cube 55^3x6 tets (998250), 100000 uniform sites, TK=64, K=90, "
                               "10 Lloyd iterations from the seeded sites (restarted every 10 steps)",
                   "n_tets": int(len(tets)), "n_sites": int(N_SITES), "tet_k": TK, "site_k": K,
                   "clip_precision": "fp32", "accumulation": "fp64 (exact fixed-point site sums)",
                   "moments": FAST["moments"], "launch": "site-major"},
        "phase_ms": {k: round(v, 4) for k, v in ph.items()},
        "iteration_ms": [round(float(np.mean([tl[j]["total_ms"] for j in ix])), 4) for ix in it_idx],
        "iteration1_ms": round(ph1["total_ms"], 4),
        "parity_form": parity,
        "plain_fp32": plain,
        "plain_fp32_iters_per_s": None if plain is None else plain["iters_per_s"],
        "speedup_vs_plain": None if plain is None else plain["speedup_same_iterations"],
        "cells_form": cells,
        "accuracy": acc,
        "device_copy_gbs": None if copy_gbs is None else round(copy_gbs, 1),
        "batched": batch,
        "served": served,
        "roofline": roof,
        "roofline_knn": roof_knn,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
