"""Lloyd-iteration throughput of the clipped-Voronoi engine on one MI355X.

Workload (the single-solve configuration): unit cube split into 55^3 cubes x 6
Kuhn tets (998,250 tets, vertices jittered by U(-0.1,0.1)*h), 100,000 uniform
FP32 sites, candidate sites per tet TK = 64, neighbours per site K = 90.  A
step is one whole Lloyd iteration on the device: grid kNN of the sites,
candidate kNN of the tets, clipping of all n_tets*TK (tet, site) pairs with
volume/moment accumulation, site update.  Inputs are resident in HBM before
the timed region.

Reported:
  value           Lloyd iterations/s of the engine's default fast form
                  (site-major launch, FP32 clipping, FP64 accumulation)
  plain_*         the plain form: tet-major launch (one lane per pair, the
                  reference's order), FP32 accumulation with FP32 atomics
  roofline        clipping kernel: algorithmic FLOPs per launch / measured
                  kernel time vs the FP32 vector peak
  cpu_baseline    the oracle's FP32 restatement of the reference's kernel A
                  on the host cores, on a bounded sample of the same pairs
  batched         a second workload, not part of `value`: 256 independent
                  jittered-cube instances (6n^3 tets, n = round(logU[12,45]),
                  tets/10 sites, TK = 64; 3.85e7 tets, 2.46e9 pairs) solved in
                  one context, each iteration a replay of one HIP graph
  served          a third workload: a Poisson stream of jobs of that size
                  mix plus 5% torus jobs, 5..50 iterations each, through the
                  job server (200 GB admission budget); p50/p99 job latency
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
N_CUBE, N_SITES, TK, K = 55, 100000, 64, 90

# FLOP weights per event of the reference's kernel A (add/mul = 1, fma = 2):
# pair setup (4 face planes), bisector, conflict test (det4x4), security-radius
# term per vertex (vertex position + distance), per security check, moment
# integration per vertex.
W_PAIR, W_CLIP, W_CONFLICT, W_SECURE_VTX, W_SECURE_CHECK, W_ACC_VTX = 80, 14, 45, 67, 9, 545


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "clip_site_998k_tets_100k_sites_hbm_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per launch of the clipping kernel at this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_engine(native, verts, tets, sites, steps, warmup, **cfg):
    e = native.Engine(0, site_k=K, tet_k=TK, **cfg)
    e.set_mesh(verts, tets)
    e.set_sites(sites)
    for _ in range(warmup):
        e.lloyd(1)
    e.sync()
    e.timing_reset()
    t0 = time.perf_counter()
    e.lloyd(steps)
    e.sync()
    wall = time.perf_counter() - t0
    tsum = e.timing_sum()
    hist = e.status_hist()
    e.close()
    return wall, tsum, hist


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
    e = native.Engine(0, site_k=K, tet_k=TK, accum="f64")
    e.set_batch(probs)
    del probs
    e.lloyd_graph(2)  # eager sizing iteration + capture + one replay
    e.sync()
    e.timing_reset()
    t0 = time.perf_counter()
    e.lloyd_graph(steps)
    e.sync()
    wall = time.perf_counter() - t0
    ts = e.timing_sum()
    e.close()
    dev_ms = ts["total_ms"] / max(ts["iters"], 1)
    log(f"batch: {1e3 * wall / steps:.1f} ms/iter wall, {dev_ms:.1f} ms device per replayed iteration")
    return {"instances": 256, "n_tets": n_t, "n_sites": n_s, "pairs_per_iter": n_t * TK, "steps": steps,
            "ms_per_iter": round(1e3 * wall / steps, 3), "device_ms_per_iter": round(dev_ms, 3),
            "iters_per_s": round(steps / wall, 4), "pairs_per_s": round(n_t * TK * steps / wall, 1),
            "launch": "one context, hipGraph replay of a whole iteration", "accumulation": "fp64"}


def run_serving(native, n_jobs, rate, budget_gb=200):
    """Served workload: a Poisson stream of jobs (the batched workload's size
    mix plus 5% torus jobs, 5..50 iterations each) through the job server with
    a device-memory admission budget, in real time.  Job latency = submission
    to completion."""
    from voronoi_amd import meshgen

    t0 = time.perf_counter()
    jobs = meshgen.serving_jobs(n_jobs, rate)
    log(f"serving: {n_jobs} jobs generated ({time.perf_counter() - t0:.1f}s)")
    pair_iters = sum(len(p[1]) * TK * it for _, p, it in jobs)
    s = native.Server(0, budget_bytes=budget_gb << 30, site_k=K, tet_k=TK, accum="f64")
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
    res = {"jobs": n_jobs, "arrival_rate_per_s": rate, "budget_gb": budget_gb, "steps": steps,
           "span_s": round(span, 3), "jobs_per_s": round(n_jobs / span, 3),
           "pairs_per_s": round(pair_iters / span, 1),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                          "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)},
           "wait_ms": {"p50": round(float(np.percentile(wait, 50)), 2), "p99": round(float(np.percentile(wait, 99)), 2)}}
    log(f"serving: {res}")
    return res


def cpu_baseline(verts, tets, sites, budget_s):
    """Oracle FP32 kernel A on a bounded, evenly spread sample of tets."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = O.nthreads()
    t0 = time.perf_counter()
    sk = O.site_knn(sites, K)
    t_site_knn = time.perf_counter() - t0
    # candidate kNN and clipping on a sample; first a small probe to size it
    sample = []
    for frac in (1 / 512.0, None):
        if frac is None:
            per_tet = probe_s / len(probe)
            n = int(min(len(tets), max(len(probe), 0.5 * budget_s / per_tet)))
        else:
            n = max(64, int(len(tets) * frac))
        sel = np.linspace(0, len(tets) - 1, n).astype(np.int64)
        t0 = time.perf_counter()
        cent = O.tet_centroids(verts, tets[sel])
        tk = O.knn(sites, cent, TK, False)
        t_knn = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = O.clip_tets(sites, sk, verts, tets[sel], tk, geom="f32", pair_acc="f32", site_mode=O.SITE_F32,
                        want_pairs=False)
        t_clip = time.perf_counter() - t0
        if frac is not None:
            probe, probe_s = sel, t_clip + t_knn
        else:
            sample = (sel, t_knn, t_clip, r["events"])
    sel, t_knn, t_clip, ev = sample
    scale = len(tets) / len(sel)
    t_iter = t_site_knn + (t_knn + t_clip) * scale
    return dict(
        iters_per_s=1.0 / t_iter,
        cores=threads,
        sample=f"{len(sel)} of {len(tets)} tets (evenly spaced) x TK={TK} pairs; site kNN of all {len(sites)} sites; "
               f"clip {t_clip:.2f}s + candidate kNN {t_knn:.2f}s + site kNN {t_site_knn:.2f}s, scaled to the full mesh",
        events=ev,
        sample_pairs=len(sel) * TK,
        clip_s_per_pair=t_clip / (len(sel) * TK),
    )


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


def flops_per_pair(ev):
    f = (W_PAIR * ev["pairs"] + W_CLIP * ev["clips"] + W_CONFLICT * ev["conflict_tests"]
         + W_SECURE_VTX * ev["secure_evals"] + W_SECURE_CHECK * ev["secure_checks"] + W_ACC_VTX * ev["acc_tris"])
    return f / max(ev["pairs"], 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU work for the baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0, help="also time a device copy of this size (MiB)")
    ap.add_argument("--batch-steps", type=int, default=3, help="graph replays of the 256-instance batch (0 skips)")
    ap.add_argument("--serve-jobs", type=int, default=120, help="jobs of the served workload (0 skips)")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="job arrivals per second")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(sharding over several GPUs is not built)"}))
        sys.exit(2)

    from voronoi_amd import meshgen, native

    t0 = time.perf_counter()
    verts, tets, sites = meshgen.make_problem("cube", N_CUBE, N_SITES, seed=7)
    log(f"workload: {len(tets)} tets, {len(verts)} verts, {len(sites)} sites ({time.perf_counter() - t0:.1f}s)")

    wall, ts, hist = run_engine(native, verts, tets, sites, args.steps, args.warmup, form="site", accum="f64")
    n = ts["iters"]
    ms_step = 1e3 * wall / args.steps
    log(f"fast: {ms_step:.2f} ms/iter wall; per-iter device: knn {ts['knn_ms']/n:.2f} pairs {ts['pairs_ms']/n:.2f} "
        f"clip {ts['clip_ms']/n:.2f} (overflow {ts['overflow_ms']/n:.2f}) update {ts['update_ms']/n:.3f} ms; "
        f"status {hist}")

    # the cell -> tet formulation (kernels B + C), the reference's source of the new sites
    cwall, cts, _ = run_engine(native, verts, tets, sites, max(1, args.plain_steps), 1, form="cells", accum="f64")
    cells = {"iters_per_s": round(max(1, args.plain_steps) / cwall, 4),
             "clip_ms": round(cts["clip_ms"] / max(cts["iters"], 1), 3)}
    log(f"cells form: {1e3 * cwall / max(1, args.plain_steps):.2f} ms/iter wall; clip {cells['clip_ms']:.2f} ms")

    plain = None
    if args.plain_steps > 0:
        pwall, pts, _ = run_engine(native, verts, tets, sites, args.plain_steps, 1, form="tet", accum="f32")
        pn = pts["iters"]
        plain = (args.plain_steps / pwall, pts["clip_ms"] / pn)
        log(f"plain: {1e3 * pwall / args.plain_steps:.2f} ms/iter wall; clip {pts['clip_ms']/pn:.2f} ms")
    batch = run_batch(native, args.batch_steps) if args.batch_steps > 0 else None
    served = run_serving(native, args.serve_jobs, args.serve_rate) if args.serve_jobs > 0 else None
    copy_gbs = device_copy(args.copy_mb) if args.copy_mb > 0 else None
    if copy_gbs:
        log(f"device copy of {args.copy_mb} MiB: {copy_gbs:.0f} GB/s (read + write)")

    n_pairs = len(tets) * TK
    cpu = None
    fpp = None
    if not args.no_cpu:
        cpu = cpu_baseline(verts, tets, sites, args.cpu_budget)
        fpp = flops_per_pair(cpu["events"])
        log(f"cpu baseline: {cpu['iters_per_s']:.4f} iters/s on {cpu['cores']} threads; {fpp:.0f} FLOP/pair")

    # the site-major clipping with its emptiness pass and volume pass: the
    # clip phase minus the re-clipping of pairs that overflowed its caps (HIP
    # events on the engine's stream, around the three launches)
    kern_s = (ts["clip_ms"] - ts["overflow_ms"]) / n / 1e3
    roof = None
    if fpp is not None:
        achieved = fpp * n_pairs / kern_s / 1e12
        roof = {"bound": "valu", "kernel": "k_cull_site + k_clip_site + k_site_moments (emptiness test; clipping; volumes, "
                                             "moments, accumulation)", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                "traffic": profiled_traffic(), "flop_per_launch": fpp * n_pairs, "kernel_ms": round(kern_s * 1e3, 3)}
    value = args.steps / wall
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
                               "1 Lloyd iteration per step",
                   "n_tets": int(len(tets)), "n_sites": int(N_SITES), "tet_k": TK, "site_k": K,
                   "clip_precision": "fp32", "accumulation": "fp64 (exact fixed-point site sums)",
                   "launch": "site-major"},
        "phase_ms": {k: round(ts[k] / n, 3) for k in ("knn_ms", "pairs_ms", "clip_ms", "overflow_ms", "update_ms",
                                                      "total_ms")},
        "plain_fp32_iters_per_s": None if plain is None else round(plain[0], 4),
        "plain_clip_ms": None if plain is None else round(plain[1], 3),
        "speedup_vs_plain": None if plain is None else round(value / plain[0], 3),
        "device_copy_gbs": None if copy_gbs is None else round(copy_gbs, 1),
        "cells_form": cells,
        "batched": batch,
        "served": served,
        "roofline": roof,
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["iters_per_s"], 5), "unit": "iters/s", "cores": cpu["cores"], "kind": "port",
            "sample": cpu["sample"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
