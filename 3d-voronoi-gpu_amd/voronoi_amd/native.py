"""ctypes binding of libvoronoi_hip.so (the C ABI in include/vd_api.h).

The library must have been built (`make -C 3d-voronoi-gpu_amd`); loading fails
loudly otherwise -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "_build", "libvoronoi_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "vd_api.h")

VD_OK = 0
ACCUM = {"f32": 0, "f64": 1}
MOMENTS = {"reference": 0, "robust": 1}
FORM = {"tet": 0, "site": 1, "cells": 2}
STATUS_NAMES = [
    "triangle_overflow", "vertex_overflow", "inconsistent_boundary",
    "security_radius_not_reached", "success", "needs_exact_predicates", "no_intersection",
]


class Config(ctypes.Structure):
    _fields_ = [
        ("site_k", ctypes.c_int),
        ("tet_k", ctypes.c_int),
        ("pcap", ctypes.c_int),
        ("tcap", ctypes.c_int),
        ("accum", ctypes.c_int),
        ("moments", ctypes.c_int),
        ("form", ctypes.c_int),
        ("grid_density", ctypes.c_float),
        ("record_pairs", ctypes.c_int),
        ("clip_blocks", ctypes.c_int),
        ("site_tcap", ctypes.c_int),
        ("cull_depth", ctypes.c_int),
        ("srnr_k", ctypes.c_int),
        ("filter", ctypes.c_int),
        ("site_records", ctypes.c_int),
        ("site_depth", ctypes.c_int),
        ("knn_waves", ctypes.c_int),
        ("cells_pcap", ctypes.c_int),
        ("cells_tcap", ctypes.c_int),
    ]


class Timing(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("knn_ms", "pairs_ms", "clip_ms", "update_ms", "total_ms", "overflow_ms")]


class KnInfo(ctypes.Structure):
    _fields_ = [
        ("allocated_points", ctypes.c_int), ("dimx", ctypes.c_int), ("dimy", ctypes.c_int), ("dimz", ctypes.c_int),
        ("num_cell_offsets", ctypes.c_int), ("d_stored_points", ctypes.c_void_p), ("d_knearests", ctypes.c_void_p),
        ("d_permutation", ctypes.c_void_p), ("d_counters", ctypes.c_void_p), ("d_ptrs", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}`")
        L = ctypes.CDLL(LIB_PATH)
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        L.vd_config_default.argtypes = [P]
        L.vd_ctx_create.argtypes = [P, I, P]
        L.vd_ctx_destroy.argtypes = [P]
        L.vd_ctx_error.argtypes = [P]
        L.vd_ctx_error.restype = ctypes.c_char_p
        L.vd_ctx_set_config.argtypes = [P, P]
        L.vd_set_mesh.argtypes = [P, P, I, P, I]
        L.vd_set_sites.argtypes = [P, P, I]
        L.vd_set_batch.argtypes = [P, I, P, P, P, P, P, P]
        L.vd_get_batch.argtypes = [P, P]
        L.vd_get_sites.argtypes = [P, P]
        for f in ("vd_knn_step", "vd_clip_step", "vd_update_step", "vd_sync"):
            getattr(L, f).argtypes = [P]
        L.vd_lloyd.argtypes = [P, I]
        L.vd_lloyd_graph.argtypes = [P, I]
        L.vd_autotune.argtypes = [P, I, P]
        L.vd_autotune_k.argtypes = [P, I, ctypes.c_double, P, P]
        L.vd_get_timing.argtypes = [P, P]
        L.vd_timing_reset.argtypes = [P]
        L.vd_get_timing_sum.argtypes = [P, P, P]
        L.vd_get_timing_list.argtypes = [P, P, I]
        L.vd_get_dims.argtypes = [P, P, P, P, P]
        L.vd_get_site_knn.argtypes = [P, P]
        L.vd_get_tet_knn.argtypes = [P, P]
        L.vd_get_moments.argtypes = [P, P]
        L.vd_get_pairs.argtypes = [P, P, P]
        L.vd_get_pairs_range.argtypes = [P, ctypes.c_longlong, ctypes.c_longlong, P, P]
        L.vd_get_tet_knn_range.argtypes = [P, I, I, P]
        L.vd_get_status_hist.argtypes = [P, P]
        L.vd_knn_dev.argtypes = [P, I, ctypes.c_size_t, P, I, ctypes.c_size_t, I, I, P, ctypes.c_size_t, P]
        L.vd_knn_dev_noalloc.argtypes = L.vd_knn_dev.argtypes
        L.vd_kn_prepare.argtypes = [P, P, I, I]
        L.vd_kn_solve.argtypes = [P]
        L.vd_kn_free.argtypes = [P]
        L.vd_kn_info_get.argtypes = [P, P]
        L.vd_kn_get_points.argtypes = [P, P]
        L.vd_kn_get_knearests.argtypes = [P, P]
        L.vd_kn_get_permutation.argtypes = [P, P]
        L.vd_compute_clipped_voro_diagram.argtypes = [P, I, P, I, P, I, I, I, I, P, ctypes.c_char_p]
        L.vd_server_create.argtypes = [P, I, P, ctypes.c_size_t]
        L.vd_server_destroy.argtypes = [P]
        L.vd_server_error.argtypes = [P]
        L.vd_server_error.restype = ctypes.c_char_p
        L.vd_server_submit.argtypes = [P, P, I, P, I, P, I, I, P]
        L.vd_server_step.argtypes = [P, P, P]
        L.vd_server_job.argtypes = [P, I, P, P, P, P]
        L.vd_server_result.argtypes = [P, I, P]
        L.vd_server_bytes_in_use.argtypes = [P]
        L.vd_server_bytes_in_use.restype = ctypes.c_size_t
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_config(**kw) -> Config:
    c = Config()
    lib().vd_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if k == "accum":
            v = ACCUM.get(v, v)
        elif k == "moments":
            v = MOMENTS.get(v, v)
        elif k == "form":
            v = FORM.get(v, v)
        setattr(c, k, v)
    return c


class Engine:
    """One device context: a mesh and a site set resident on the GPU."""

    def __init__(self, device: int = 0, **cfg):
        self.cfg = default_config(**cfg)
        h = ctypes.c_void_p()
        rc = lib().vd_ctx_create(ctypes.byref(h), device, ctypes.byref(self.cfg))
        if rc != VD_OK:
            raise RuntimeError(f"vd_ctx_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vd_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc, what):
        if rc != VD_OK:
            msg = lib().vd_ctx_error(self.h).decode()
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    def configure(self, **cfg):
        c = default_config()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self.cfg), ctypes.sizeof(Config))
        for k, v in cfg.items():
            v = {"accum": ACCUM, "moments": MOMENTS, "form": FORM}.get(k, {}).get(v, v)
            setattr(c, k, v)
        self._ck(lib().vd_ctx_set_config(self.h, ctypes.byref(c)), "vd_ctx_set_config")
        self.cfg = c

    def set_mesh(self, verts, tets):
        v = np.ascontiguousarray(verts, dtype=np.float32)
        t = np.ascontiguousarray(tets, dtype=np.int32)
        self._ck(lib().vd_set_mesh(self.h, _p(v), len(v), _p(t), len(t)), "vd_set_mesh")

    def set_sites(self, sites):
        s = np.ascontiguousarray(sites, dtype=np.float32)
        self._ck(lib().vd_set_sites(self.h, _p(s), len(s)), "vd_set_sites")

    def set_batch(self, problems):
        """problems: sequence of (verts, tets, sites), tets with local vertex ids."""
        nv = np.array([len(p[0]) for p in problems], dtype=np.int32)
        nt = np.array([len(p[1]) for p in problems], dtype=np.int32)
        ns = np.array([len(p[2]) for p in problems], dtype=np.int32)
        v = np.ascontiguousarray(np.concatenate([np.asarray(p[0], np.float32).reshape(-1, 3) for p in problems]))
        t = np.ascontiguousarray(np.concatenate([np.asarray(p[1], np.int32).reshape(-1, 4) for p in problems]))
        st = np.ascontiguousarray(np.concatenate([np.asarray(p[2], np.float32).reshape(-1, 3) for p in problems]))
        self._ck(lib().vd_set_batch(self.h, len(problems), _p(nv), _p(nt), _p(ns), _p(v), _p(t), _p(st)),
                 "vd_set_batch")
        self.batch_sites = np.concatenate([[0], np.cumsum(ns)])

    def dims(self):
        k, tk, ns, nt = (ctypes.c_int() for _ in range(4))
        self._ck(lib().vd_get_dims(self.h, ctypes.byref(k), ctypes.byref(tk), ctypes.byref(ns), ctypes.byref(nt)),
                 "vd_get_dims")
        return k.value, tk.value, ns.value, nt.value

    def sites(self):
        _, _, ns, _ = self.dims()
        out = np.empty((ns, 3), dtype=np.float32)
        self._ck(lib().vd_get_sites(self.h, _p(out)), "vd_get_sites")
        return out

    def knn_step(self):
        self._ck(lib().vd_knn_step(self.h), "vd_knn_step")

    def clip_step(self):
        self._ck(lib().vd_clip_step(self.h), "vd_clip_step")

    def update_step(self):
        self._ck(lib().vd_update_step(self.h), "vd_update_step")

    def lloyd(self, n_iter=1):
        self._ck(lib().vd_lloyd(self.h, n_iter), "vd_lloyd")

    def autotune(self, iters=2):
        """Times candidate launch parameters and keeps the fastest; returns them."""
        c = Config()
        self._ck(lib().vd_autotune(self.h, iters, ctypes.byref(c)), "vd_autotune")
        self.cfg = c
        return {n: getattr(c, n) for n, _ in Config._fields_}

    def autotune_k(self, iters=2, tol=1e-5):
        """Smaller site_k / tet_k where the sites after `iters` iterations move
        by at most tol * |site|; returns (chosen config, error)."""
        c, err = Config(), ctypes.c_double()
        self._ck(lib().vd_autotune_k(self.h, iters, tol, ctypes.byref(c), ctypes.byref(err)), "vd_autotune_k")
        self.cfg = c
        return {n: getattr(c, n) for n, _ in Config._fields_}, err.value

    def lloyd_graph(self, n_iter=1):
        """n_iter iterations replayed from one captured HIP graph."""
        self._ck(lib().vd_lloyd_graph(self.h, n_iter), "vd_lloyd_graph")

    def sync(self):
        self._ck(lib().vd_sync(self.h), "vd_sync")

    def timing(self):
        t = Timing()
        self._ck(lib().vd_get_timing(self.h, ctypes.byref(t)), "vd_get_timing")
        return {n: getattr(t, n) for n, _ in Timing._fields_}

    def timing_reset(self):
        self._ck(lib().vd_timing_reset(self.h), "vd_timing_reset")

    def timing_sum(self):
        t, n = Timing(), ctypes.c_int()
        self._ck(lib().vd_get_timing_sum(self.h, ctypes.byref(t), ctypes.byref(n)), "vd_get_timing_sum")
        d = {k: getattr(t, k) for k, _ in Timing._fields_}
        d["iters"] = n.value
        return d

    def timing_list(self, n):
        """Phase times of each of the last n completed iterations, oldest first."""
        t = (Timing * max(n, 1))()
        self._ck(lib().vd_get_timing_list(self.h, t, n), "vd_get_timing_list")
        return [{k: getattr(t[i], k) for k, _ in Timing._fields_} for i in range(n)]

    def site_knn(self):
        k, _, ns, _ = self.dims()
        out = np.empty((ns, k), dtype=np.int32)
        self._ck(lib().vd_get_site_knn(self.h, _p(out)), "vd_get_site_knn")
        return out

    def tet_knn(self, first=None, count=None):
        """Candidate rows of all tets, or of tets first .. first+count-1."""
        _, tk, _, nt = self.dims()
        if first is None:
            out = np.empty((nt, tk), dtype=np.int32)
            self._ck(lib().vd_get_tet_knn(self.h, _p(out)), "vd_get_tet_knn")
            return out
        out = np.empty((count, tk), dtype=np.int32)
        self._ck(lib().vd_get_tet_knn_range(self.h, first, count, _p(out)), "vd_get_tet_knn_range")
        return out

    def moments(self):
        _, _, ns, _ = self.dims()
        out = np.empty((ns, 4), dtype=np.float64)
        self._ck(lib().vd_get_moments(self.h, _p(out)), "vd_get_moments")
        return out

    def pairs(self, first=None, count=None, vals=True):
        """Per-pair status and (vol, mx, my, mz) of all pairs, or of pairs
        first .. first+count-1 (pair = tet * TK + rank); vals=False reads the
        statuses only."""
        _, tk, _, nt = self.dims()
        if first is None:
            first, count = 0, nt * tk
        st = np.empty(count, dtype=np.uint8)
        v = np.empty((count, 4), dtype=np.float64) if vals else None
        self._ck(lib().vd_get_pairs_range(self.h, first, count, _p(st), _p(v)), "vd_get_pairs_range")
        return (st, v) if vals else st

    def status_hist(self):
        h = np.zeros(7, dtype=np.uint64)
        self._ck(lib().vd_get_status_hist(self.h, _p(h)), "vd_get_status_hist")
        return dict(zip(STATUS_NAMES, h.tolist()))


class KnProblem:
    """The reference's grid-kNN problem (kn_prepare / kn_solve / getters)."""

    def __init__(self, points, k=90):
        pts = np.ascontiguousarray(points, dtype=np.float32)
        h = ctypes.c_void_p()
        rc = lib().vd_kn_prepare(ctypes.byref(h), _p(pts), len(pts), k)
        if rc != VD_OK:
            raise RuntimeError(f"vd_kn_prepare failed ({rc})")
        self.h, self.n, self.k = h, len(pts), k

    def solve(self):
        rc = lib().vd_kn_solve(self.h)
        if rc != VD_OK:
            raise RuntimeError(f"vd_kn_solve failed ({rc})")

    def info(self):
        i = KnInfo()
        lib().vd_kn_info_get(self.h, ctypes.byref(i))
        return i

    def points(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        assert lib().vd_kn_get_points(self.h, _p(out)) == VD_OK
        return out

    def knearests(self):
        out = np.empty((self.n, self.k), dtype=np.uint32)
        assert lib().vd_kn_get_knearests(self.h, _p(out)) == VD_OK
        return out

    def permutation(self):
        out = np.empty(self.n, dtype=np.uint32)
        assert lib().vd_kn_get_permutation(self.h, _p(out)) == VD_OK
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().vd_kn_free(self.h)
            self.h = None

    __del__ = close


def compute_clipped_voro_diagram(verts, tets, sites, nb_iter=1, tet_k=0, out_dir=None, transposed=False, **cfg):
    """Runs the whole-program driver; returns the new sites (same layout as given)."""
    v = np.ascontiguousarray(verts, dtype=np.float32)
    t = np.ascontiguousarray(tets, dtype=np.int32)
    s = np.array(sites, dtype=np.float32, copy=True, order="C")
    n = s.shape[1] if transposed else s.shape[0]
    c = default_config(**cfg)
    rc = lib().vd_compute_clipped_voro_diagram(_p(v), len(v), _p(t), len(t), _p(s), n, int(transposed), nb_iter,
                                               tet_k, ctypes.byref(c), (out_dir or "").encode())
    if rc != VD_OK:
        raise RuntimeError(f"vd_compute_clipped_voro_diagram failed ({rc})")
    return s


JOB_PENDING, JOB_RUNNING, JOB_DONE = 0, 1, 2


class Server:
    """Job server: continuous batching of independent Lloyd jobs on one device."""

    def __init__(self, device: int = 0, budget_bytes: int = 200 << 30, **cfg):
        self.cfg = default_config(**cfg)
        h = ctypes.c_void_p()
        rc = lib().vd_server_create(ctypes.byref(h), device, ctypes.byref(self.cfg), budget_bytes)
        if rc != VD_OK:
            raise RuntimeError(f"vd_server_create failed ({rc})")
        self.h = h
        self.n_sites = {}

    def close(self):
        if getattr(self, "h", None):
            lib().vd_server_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc, what):
        if rc != VD_OK:
            raise RuntimeError(f"{what} failed ({rc}): {lib().vd_server_error(self.h).decode()}")

    def submit(self, verts, tets, sites, n_iter):
        v = np.ascontiguousarray(verts, dtype=np.float32)
        t = np.ascontiguousarray(tets, dtype=np.int32)
        st = np.ascontiguousarray(sites, dtype=np.float32)
        jid = ctypes.c_int()
        self._ck(lib().vd_server_submit(self.h, _p(v), len(v), _p(t), len(t), _p(st), len(st), n_iter,
                                        ctypes.byref(jid)), "vd_server_submit")
        self.n_sites[jid.value] = len(st)
        return jid.value

    def step(self):
        r, p = ctypes.c_int(), ctypes.c_int()
        self._ck(lib().vd_server_step(self.h, ctypes.byref(r), ctypes.byref(p)), "vd_server_step")
        return r.value, p.value

    def job(self, jid):
        st, it = ctypes.c_int(), ctypes.c_int()
        w, r = ctypes.c_double(), ctypes.c_double()
        self._ck(lib().vd_server_job(self.h, jid, ctypes.byref(st), ctypes.byref(it), ctypes.byref(w), ctypes.byref(r)),
                 "vd_server_job")
        return {"state": st.value, "iters": it.value, "wait_ms": w.value, "run_ms": r.value}

    def result(self, jid):
        out = np.empty((self.n_sites[jid], 3), dtype=np.float32)
        self._ck(lib().vd_server_result(self.h, jid, _p(out)), "vd_server_result")
        return out

    def bytes_in_use(self):
        return lib().vd_server_bytes_in_use(self.h)
