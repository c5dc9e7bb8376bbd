"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it checks the HIP path and is never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libvd_oracle.so")

STATUS_NAMES = [
    "triangle_overflow",
    "vertex_overflow",
    "inconsistent_boundary",
    "security_radius_not_reached",
    "success",
    "needs_exact_predicates",
    "no_intersection",
]
EVENT_NAMES = [
    "pairs", "clips", "conflict_tests", "secure_evals", "secure_checks",
    "new_tris", "acc_tris", "acc_pairs", "max_nv", "max_nt",
]

SITE_F32, SITE_FIXED, SITE_F64 = 0, 1, 2
MOMENTS = {"reference": 0, "robust": 1}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int, ctypes.c_int64
        L.vdo_knn.argtypes = [P, i32, P, i32, i32, i32, i32, P, i32]
        L.vdo_tet_centroids.argtypes = [P, P, i32, P]
        L.vdo_tet_k.argtypes = [i32, i32, i32]
        L.vdo_clip_tets.argtypes = [i32, i32, P, i32, P, i32, P, P, i32, P, i32, i32, i32, P, P, P, i32, P, P,
                                    i32, i64, i64, i32, i32]
        L.vdo_clip_cells.argtypes = [i32, i32, P, i32, P, i32, P, P, i32, P, i32, i32, i32, P, P, P, i32, P, i32,
                                     i32, i32]
        L.vdo_lloyd_update.argtypes = [P, i32, i32, P]
        L.vdo_det4.argtypes = [P, P, P, P, i32]
        L.vdo_det4.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def nthreads() -> int:
    """VD_ORACLE_THREADS, else OMP_NUM_THREADS, else the CPUs this process may run on."""
    for var in ("VD_ORACLE_THREADS", "OMP_NUM_THREADS"):
        if os.environ.get(var, "").strip().isdigit() and int(os.environ[var]) > 0:
            return int(os.environ[var])
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def knn(ref, qry, k, self_exclude, grid=True, threads=None):
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    qry = np.ascontiguousarray(qry, dtype=np.float32)
    out = np.empty((len(qry), k), dtype=np.int32)
    rc = lib().vdo_knn(_p(ref), len(ref), _p(qry), len(qry), k, int(self_exclude), int(grid), _p(out),
                       threads or nthreads())
    assert rc == 0
    return out


def site_knn(sites, k, grid=True):
    return knn(sites, sites, k, True, grid)


def tet_centroids(verts, tets):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    out = np.empty((len(tets), 3), dtype=np.float32)
    lib().vdo_tet_centroids(_p(verts), _p(tets), len(tets), _p(out))
    return out


def tet_knn(verts, tets, sites, tk, grid=True):
    return knn(sites, tet_centroids(verts, tets), tk, False, grid)


def tet_k(n_site, n_tet, preferred=0):
    return lib().vdo_tet_k(n_site, n_tet, preferred)


def clip_tets(sites, sknn, verts, tets, tknn, geom="f32", pair_acc="f32", site_mode=SITE_F32,
              pcap=64, tcap=96, pair_range=None, want_pairs=True, threads=None, moments="reference", filter=0):
    """Kernel A over all (tet, candidate) pairs (filter: arithmetic filter mode
    of the conflict test, 0 off / 1 the reference's / 2 FP64 re-check). Returns dict with
    site_vals (n_s,4) [vol, mx, my, mz], pair_status, pair_vals, events, peak_hist."""
    sites = np.ascontiguousarray(sites, dtype=np.float32)
    sknn = np.ascontiguousarray(sknn, dtype=np.int32)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    tknn = np.ascontiguousarray(tknn, dtype=np.int32)
    k = sknn.shape[1]
    tk = tknn.shape[1]
    npairs = len(tets) * tk
    b, e = pair_range if pair_range else (0, npairs)
    status = np.full(npairs, -1, dtype=np.int32) if want_pairs else None
    vals = np.zeros((npairs, 4), dtype=np.float64) if want_pairs else None
    site_vals = np.zeros((len(sites), 4), dtype=np.float64)
    ev = np.zeros(10, dtype=np.int64)
    hist = np.zeros(1024, dtype=np.int64)
    rc = lib().vdo_clip_tets(0 if geom == "f32" else 1, 0 if pair_acc == "f32" else 1, _p(sites), len(sites),
                             _p(sknn), k, _p(verts), _p(tets), len(tets), _p(tknn), tk, pcap, tcap, _p(status),
                             _p(vals), _p(site_vals), site_mode, _p(ev), _p(hist), threads or nthreads(), b, e,
                             MOMENTS[moments], filter)
    assert rc == 0
    return dict(site_vals=site_vals, pair_status=status, pair_vals=vals,
                events=dict(zip(EVENT_NAMES, ev.tolist())), peak_hist=hist)


def clip_cells(sites, sknn, verts, tets, tknn, geom="f32", pair_acc="f32", site_mode=SITE_F32,
               pcap=64, tcap=96, threads=None, moments="reference", filter=0):
    """Kernels B then C. Returns dict like clip_tets plus cell_status."""
    sites = np.ascontiguousarray(sites, dtype=np.float32)
    sknn = np.ascontiguousarray(sknn, dtype=np.int32)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    tknn = np.ascontiguousarray(tknn, dtype=np.int32)
    npairs = len(tets) * tknn.shape[1]
    status = np.full(npairs, -1, dtype=np.int32)
    vals = np.zeros((npairs, 4), dtype=np.float64)
    site_vals = np.zeros((len(sites), 4), dtype=np.float64)
    cstat = np.zeros(len(sites), dtype=np.int32)
    rc = lib().vdo_clip_cells(0 if geom == "f32" else 1, 0 if pair_acc == "f32" else 1, _p(sites), len(sites),
                              _p(sknn), sknn.shape[1], _p(verts), _p(tets), len(tets), _p(tknn), tknn.shape[1],
                              pcap, tcap, _p(status), _p(vals), _p(site_vals), site_mode, _p(cstat),
                              threads or nthreads(), MOMENTS[moments], filter)
    assert rc == 0
    return dict(site_vals=site_vals, pair_status=status, pair_vals=vals, cell_status=cstat)


def lloyd_update(site_vals, sites, fp32=True):
    out = np.array(sites, dtype=np.float32, copy=True)
    sv = np.ascontiguousarray(site_vals, dtype=np.float64)
    lib().vdo_lloyd_update(_p(sv), len(out), int(fp32), _p(out))
    return out


def det4(p1, p2, p3, e, fp32=False):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (p1, p2, p3, e)]
    return lib().vdo_det4(*[_p(x) for x in a], int(fp32))


def lloyd_step(verts, tets, sites, k=90, tk=64, geom="f32", pair_acc="f32", site_mode=SITE_F32, fp32_update=True,
               moments="reference"):
    """One full Lloyd iteration of the reference path on the CPU."""
    kk = min(k, len(sites) - 1)
    sknn = site_knn(sites, kk)
    tknn = tet_knn(verts, tets, sites, min(tk, len(sites)))
    r = clip_tets(sites, sknn, verts, tets, tknn, geom=geom, pair_acc=pair_acc, site_mode=site_mode,
                  want_pairs=False, moments=moments)
    return lloyd_update(r["site_vals"], sites, fp32=fp32_update), r
