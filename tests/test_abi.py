"""The C ABI library: it loads on a machine without a GPU, exports every
function that include/vd_api.h declares, and reports a missing device as an
error code instead of exiting."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "vd_api.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function(vdn):
    L = vdn.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_config_layout_and_defaults(vdn):
    c = vdn.default_config()
    assert (c.site_k, c.tet_k, c.pcap, c.tcap) == (90, 0, 64, 96)
    assert abs(c.grid_density - 3.1) < 1e-6 and c.cull_depth == 16
    assert (c.site_tcap, c.srnr_k, c.filter, c.site_records, c.site_depth) == (0, 0, 0, 0, 0)
    src = open(os.path.join(ROOT, "include", "vd_api.h")).read()
    body = src[src.index("typedef struct vd_config {"):src.index("} vd_config;")]
    fields = re.findall(r"^\s*(?:int|float)\s+(\w+);", body, flags=re.M)
    assert fields == [f for f, _ in vdn.Config._fields_]


def test_errors_are_returned_not_raised(vdn):
    L = vdn.lib()
    h = ctypes.c_void_p()
    bad = vdn.default_config(site_k=0)
    assert L.vd_ctx_create(ctypes.byref(h), 0, ctypes.byref(bad)) == -1  # VD_ERR_ARG
    assert L.vd_knn_dev(None, 0, 0, None, 0, 0, 3, 5, None, 0, None) == -1
    assert L.vd_compute_clipped_voro_diagram(None, 0, None, 0, None, 0, 0, 1, 0, None, b"") == -1
    cfg = vdn.default_config()
    rc = L.vd_ctx_create(ctypes.byref(h), 0, ctypes.byref(cfg))
    if rc == 0:  # a GPU is present: the context works and is released
        L.vd_ctx_destroy(h)
    else:
        assert rc == -2  # VD_ERR_HIP: no device


def test_product_does_not_import_the_oracle():
    """Only tests, smoke() and bench.py's CPU-baseline leg may touch oracle/."""
    pkg = os.path.join(ROOT, "3d-voronoi-gpu_amd")
    for d, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(d, f), errors="ignore").read()
                assert "vd_oracle" not in txt and "import oracle" not in txt, f
