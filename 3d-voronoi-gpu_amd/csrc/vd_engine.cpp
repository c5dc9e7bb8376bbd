// Host engine of libvoronoi_hip.so: contexts, device buffers, the Lloyd
// iteration and the C ABI declared in include/vd_api.h.
//
// One context = one mesh + one site set resident in HBM.  An iteration is
//   grid build -> site kNN -> candidate kNN -> (pairs by site) -> clip -> update
// all enqueued on the context's stream with no host round trip; the host only
// reads results when asked.  Layout in HBM:
//   verts float4[n_verts], tets int4[n_tets], sites float4[n_sites]
//   site_knn int32[n_sites][K], tet_knn int32[n_tets][TK]
//   grid: sorted float4[n_sites] + ids + cell starts
//   accumulators: float[n_sites][4] or uint64[n_sites][4]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "vd_api.h"
#include "vd_clip.hpp"
#include "vd_knn.hpp"
#include "vd_internal.hpp"

namespace {

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e == hipSuccess) cap = std::max<size_t>(n, 1);
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Bisectors per site table of the site-major launch for K neighbours: by
// default 40, which keeps the kernel's LDS at 9 waves per CU (the kernel is
// latency bound and scales with them); the few pairs that need more are
// recomputed at the full depth.
constexpr int kAutoSiteDepth = 40;
int site_depth_of(const vd_config& c, int K) { return std::min(c.site_depth > 0 ? c.site_depth : kAutoSiteDepth, K); }
// The automatic depth of the first clipping after the caller sets sites.  Such
// sites have not been moved by a Lloyd update yet (random samples, typically):
// their cells are irregular and many more pairs need over 40 bisectors (336k
// of 63.9M for 100k uniform sites against none after a few updates), whose
// re-clip from scratch cost more than the deeper tables' lost occupancy (one
// wave per CU fewer at the same vertex cap).
constexpr int kFreshSiteDepth = 64;
// kNN steps after the caller sets sites that do without the previous rows'
// bounds (see vd_knn_step)
constexpr int kHintAfter = 2;

int tet_k_heuristic(int n_site, int n_tet, int preferred) {
  // voronoi.cu:936
  if (preferred) return preferred;
  if (n_tet <= 0) return 20;
  const int h = 10 * (int)std::ceil(8.0 * std::log10(7.0 * n_site / n_tet + 1.0));
  return std::max(h, 20);
}

}  // namespace

namespace vdi {

// By default the largest vertex cap (at most 32 and the configured cap) whose
// LDS at the configured table depth, in the 512-byte granules a CU allocates,
// still lets 9 waves share a CU's 160 KB: 17 at 40 bisectors.  A batch whose
// instances have fewer neighbours has smaller tables and the same cap.
constexpr int kSiteWavesPerCu = 9;
double fixed_point_load(const float* verts_xyz, int v0, int nv, const int32_t* tets, long long t0, int nt) {
  double m = 0.0, vol = 0.0;
  for (long long i = 3LL * v0; i < 3LL * (v0 + (long long)nv); ++i) m = std::max(m, std::fabs((double)verts_xyz[i]));
  for (long long t = t0; t < t0 + nt; ++t) {
    const float* p[4];
    for (int k = 0; k < 4; ++k) p[k] = verts_xyz + 3 * (size_t)(v0 + tets[4 * t + k]);
    double a[3], b[3], c[3];
    for (int x = 0; x < 3; ++x) {
      a[x] = (double)p[1][x] - p[0][x];
      b[x] = (double)p[2][x] - p[0][x];
      c[x] = (double)p[3][x] - p[0][x];
    }
    vol += std::fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                     a[2] * (b[0] * c[1] - b[1] * c[0])) /
           6.0;
  }
  return vol * m;
}

int site_tcap_of(const vd_config& c) {
  if (c.site_tcap > 0) return std::min(c.site_tcap, c.tcap);
  const int depth = std::max(site_depth_of(c, c.site_k), 1);
  int t = std::min(c.tcap, 32);
  while (t > 8 && ((vd::clip_site_lds(depth, t) + 511) & ~(size_t)511) * kSiteWavesPerCu > 160 * 1024) --t;
  return t;
}

}  // namespace vdi

struct vd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of the kNN step (the site queries beside the tet queries),
  // forked from and joined back into `stream` through the two events
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  vd_config cfg{};
  std::string err;

  int n_verts = 0, n_tets = 0, n_sites = 0;
  int K = 0, TK = 0;
  // batch of independent instances (vd_set_batch); empty: one instance made of
  // the whole mesh and site set
  struct Inst {
    int v0, nv, t0, nt, s0, ns;
  };
  std::vector<Inst> inst;
  bool batch() const { return inst.size() > 1; }
  DevBuf<int> b_site_begin, b_cell_base, b_site_inst, b_tet_inst, b_voff;
  DevBuf<float> b_box;
  bool have_knn = false, have_clip = false;
  float mesh_lo[3] = {0, 0, 0}, mesh_hi[3] = {0, 0, 0};
  // largest vdi::fixed_point_load of the mesh (or of the batch's instances)
  double fixed_load = 0.0;

  DevBuf<float4> verts, sites;
  DevBuf<float4> tet_xyz;  // per tet its 12 corner coordinates (3 float4)
  DevBuf<float4> tet_prep; // per tet its face planes and start vertices (vd::tet_prep)
  DevBuf<int4> tets;
  DevBuf<int4> offsets;
  int n_offsets = 0;
  // grid
  DevBuf<vd::GridShape> shape;
  DevBuf<int> g_count, g_start, g_cursor, g_cell_of, g_sorted_ids, g_scan_tmp;
  DevBuf<unsigned> p_scan_tmp;
  DevBuf<float4> g_sorted_pts;
  int max_cells = 0;
  // kNN results
  DevBuf<int32_t> site_knn, tet_knn;
  // site_knn / tet_knn hold a complete result (rows of rows_K / rows_TK
  // distinct ids) of the current mesh, batch layout and site count: the next
  // kNN step bounds each query by its previous row
  bool rows_valid = false;
  int rows_K = 0, rows_TK = 0;
  // sites set by the caller and not clipped since (the next site-major
  // clipping at the automatic depth uses kFreshSiteDepth)
  bool fresh_sites = false;
  // kNN steps since the caller set sites: the first two do without bounds
  // from the previous rows (kHintAfter)
  int knn_since_set = 0;
  // site-major pairs
  DevBuf<int> p_count, p_cursor;
  DevBuf<unsigned> p_start, p_ids;
  DevBuf<unsigned> p_surv;  // survivors of the emptiness test per site
  // accumulators
  DevBuf<float> acc_f32;
  DevBuf<unsigned long long> acc_fix;
  DevBuf<double> moments;
  DevBuf<uint8_t> pair_status;
  DevBuf<double> pair_vals;
  DevBuf<unsigned long long> hist;
  DevBuf<unsigned> retry_list;
  DevBuf<unsigned> retry_count;
  DevBuf<unsigned> retry_list2, retry_count2;  // overflows of the first re-clip
  DevBuf<unsigned> site_counter;
  // site-major form: records of finished polytopes for the volume pass
  DevBuf<uint4> rec_meta;
  DevBuf<uint32_t> rec_tris;
  DevBuf<unsigned> rec_count;
  // pairs retried with deeper neighbour lists (srnr_k)
  DevBuf<unsigned> srnr_list, srnr_count;
  DevBuf<int> srnr_row_of, srnr_row_site;
  DevBuf<int32_t> knn2;
  // cells of the cell -> tet form
  DevBuf<int> c_status, c_nv, c_nt;
  DevBuf<float> c_r2;
  DevBuf<float4> c_planes;
  DevBuf<uint32_t> c_tris;
  DevBuf<unsigned> c_redo;  // kernel B's list of cells built again at the configured caps

  // HIP events of the phases of the last kRing iterations: slot = iteration % kRing;
  // [0] start, [1] after kNN, [2] after pair grouping, [3] after clip, [4] after update,
  // [5] after the main clipping launch (before the overflow re-clipping)
  static constexpr int kRing = 256;
  hipEvent_t ev[6][kRing] = {};
  int it_slot = 0;        // slot of the iteration being enqueued
  long long it_done = 0;  // iterations completed since the last timing reset
  bool ev_knn = false, ev_pairs = false, ev_clip = false, ev_update = false;
  bool graph_slot[kRing] = {};  // iteration replayed from the graph: whole-iteration time only
  hipEvent_t& E(int phase) { return ev[phase][it_slot]; }

  // one captured Lloyd iteration (vd_lloyd_graph)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool capturing = false;
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? VD_ERR_NOMEM : VD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
};

#define VD_HIP(call, what)                          \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) return ctx->hip_fail(e_, what); \
  } while (0)

namespace {

// Phase event of the iteration being enqueued; none while a graph is being
// captured (a replay is timed as a whole).
hipError_t record_phase(vd_ctx* ctx, int phase) {
  if (ctx->capturing) return hipSuccess;
  return hipEventRecord(ctx->E(phase), ctx->stream);
}


int check_config(vd_ctx* ctx, const vd_config& c) {
  if (c.site_k < 1 || c.site_k > 256) return ctx->fail(VD_ERR_ARG, "site_k must be in [1,256], got %d", c.site_k);
  if (c.tet_k < 0 || c.tet_k > 256) return ctx->fail(VD_ERR_ARG, "tet_k must be in [0,256], got %d", c.tet_k);
  if (c.pcap < 8 || c.pcap > 254) return ctx->fail(VD_ERR_ARG, "pcap must be in [8,254], got %d", c.pcap);
  if (c.tcap < 8 || c.tcap > 1024) return ctx->fail(VD_ERR_ARG, "tcap must be in [8,1024], got %d", c.tcap);
  if (c.accum != VD_ACCUM_F32 && c.accum != VD_ACCUM_F64) return ctx->fail(VD_ERR_ARG, "bad accum %d", c.accum);
  if (c.moments != VD_MOMENTS_REFERENCE && c.moments != VD_MOMENTS_ROBUST)
    return ctx->fail(VD_ERR_ARG, "bad moments %d", c.moments);
  if (c.form != VD_FORM_TET_MAJOR && c.form != VD_FORM_SITE_MAJOR && c.form != VD_FORM_CELLS)
    return ctx->fail(VD_ERR_ARG, "bad form %d", c.form);
  if (c.form == VD_FORM_CELLS &&
      (vd::cells_lds(c.pcap, c.tcap) > 160 * 1024 || vd::cells_clip_lds(c.pcap, c.tcap) > 160 * 1024 || c.pcap > 250))
    return ctx->fail(VD_ERR_ARG, "pcap/tcap too large for the cell kernels' LDS");
  if (!(c.grid_density > 0.05f && c.grid_density < 1000.f))
    return ctx->fail(VD_ERR_ARG, "grid_density out of range");
  // the tet-major kernel also serves the site-major form's overflow retries
  if (vd::clip_plain_lds(c.pcap, c.tcap) > 160 * 1024)
    return ctx->fail(VD_ERR_ARG, "pcap/tcap too large for the tet-major kernel's LDS");
  const int fast_tcap = vdi::site_tcap_of(c);
  if (c.form == VD_FORM_SITE_MAJOR && vd::clip_site_lds(site_depth_of(c, c.site_k), fast_tcap) > 160 * 1024)
    return ctx->fail(VD_ERR_ARG, "site_k/site_tcap too large for the site-major form's LDS");
  if (c.cull_depth < 0 || c.cull_depth > 256)
    return ctx->fail(VD_ERR_ARG, "cull_depth must be in [0,256], got %d", c.cull_depth);
  if (c.site_tcap != 0 && (c.site_tcap < 8 || c.site_tcap > 32))
    return ctx->fail(VD_ERR_ARG, "site_tcap must be 0 or in [8,32], got %d", c.site_tcap);
  if (c.cells_pcap != 0 && (c.cells_pcap < 8 || c.cells_pcap > c.pcap))
    return ctx->fail(VD_ERR_ARG, "cells_pcap must be 0 or in [8,pcap], got %d", c.cells_pcap);
  if (c.cells_tcap != 0 && (c.cells_tcap < 8 || c.cells_tcap > c.tcap))
    return ctx->fail(VD_ERR_ARG, "cells_tcap must be 0 or in [8,tcap], got %d", c.cells_tcap);
  if (c.filter < 0 || c.filter > 2) return ctx->fail(VD_ERR_ARG, "filter must be 0, 1 or 2, got %d", c.filter);
  if (c.site_depth < 0 || c.site_depth > 256)
    return ctx->fail(VD_ERR_ARG, "site_depth must be in [0,256], got %d", c.site_depth);
  if (c.site_records < -1)
    return ctx->fail(VD_ERR_ARG, "site_records must be -1, 0 or a record count, got %d", c.site_records);
  if (c.srnr_k != 0 && (c.srnr_k <= c.site_k || c.srnr_k > 256))
    return ctx->fail(VD_ERR_ARG, "srnr_k must be 0 or in (site_k, 256], got %d", c.srnr_k);
  if (c.form == VD_FORM_SITE_MAJOR && c.site_k + 4 > 254)
    return ctx->fail(VD_ERR_ARG, "site-major form needs site_k <= 250");
  if (c.knn_waves != 0 && c.knn_waves != 2 && c.knn_waves != 4 && c.knn_waves != 8)
    return ctx->fail(VD_ERR_ARG, "knn_waves must be 0, 2, 4 or 8, got %d", c.knn_waves);
  if (c.clip_blocks < 0) return ctx->fail(VD_ERR_ARG, "clip_blocks must be >= 0, got %d", c.clip_blocks);
  return VD_OK;
}

int ensure_grid(vd_ctx* ctx, int n) {
  int mc = vd::grid_max_cells(n, ctx->cfg.grid_density);
  if (ctx->batch()) {
    // cells of instance i: cell_base[i] .. cell_base[i+1]
    std::vector<int> base(ctx->inst.size() + 1, 0);
    long long tot = 0;
    for (size_t i = 0; i < ctx->inst.size(); ++i) {
      base[i] = (int)tot;
      tot += vd::grid_max_cells(std::max(ctx->inst[i].ns, 1), ctx->cfg.grid_density);
    }
    if (tot > 0x7fffffffLL) return ctx->fail(VD_ERR_ARG, "batch grid too large");
    base.back() = (int)tot;
    mc = (int)tot;
    VD_HIP(ctx->b_cell_base.ensure(base.size()), "alloc grid");
    VD_HIP(hipMemcpy(ctx->b_cell_base.p, base.data(), sizeof(int) * base.size(), hipMemcpyHostToDevice),
           "upload grid");
  }
  VD_HIP(ctx->shape.ensure(std::max<size_t>(1, ctx->inst.size())), "alloc grid shape");
  VD_HIP(ctx->g_count.ensure(mc), "alloc grid");
  VD_HIP(ctx->g_start.ensure((size_t)mc + 1), "alloc grid");
  VD_HIP(ctx->g_cursor.ensure(mc), "alloc grid");
  VD_HIP(ctx->g_cell_of.ensure(n), "alloc grid");
  VD_HIP(ctx->g_sorted_ids.ensure(n), "alloc grid");
  VD_HIP(ctx->g_sorted_pts.ensure(n), "alloc grid");
  VD_HIP(ctx->g_scan_tmp.ensure(vd::scan_tmp_ints(mc)), "alloc grid");
  ctx->max_cells = mc;
  return VD_OK;
}

vd::GridBufs grid_bufs(vd_ctx* ctx) {
  vd::GridBufs gb;
  gb.shape = ctx->shape.p;
  if (ctx->batch()) {
    gb.n_inst = (int)ctx->inst.size();
    gb.cell_base = ctx->b_cell_base.p;
    gb.pt_begin = ctx->b_site_begin.p;
    gb.pt_inst = ctx->b_site_inst.p;
    gb.box = ctx->b_box.p;
  }
  gb.max_cells = ctx->max_cells;
  gb.count = ctx->g_count.p;
  gb.start = ctx->g_start.p;
  gb.cursor = ctx->g_cursor.p;
  gb.cell_of = ctx->g_cell_of.p;
  gb.sorted_ids = ctx->g_sorted_ids.p;
  gb.sorted_pts = ctx->g_sorted_pts.p;
  gb.scan_tmp = ctx->g_scan_tmp.p;
  return gb;
}

int upload_xyz(vd_ctx* ctx, DevBuf<float4>& dst, const float* xyz, int n) {
  std::vector<float4> h((size_t)std::max(n, 1));
  for (int i = 0; i < n; ++i) h[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
  VD_HIP(dst.ensure(std::max(n, 1)), "alloc points");
  if (n > 0)
    VD_HIP(hipMemcpyAsync(dst.p, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream), "upload points");
  VD_HIP(hipStreamSynchronize(ctx->stream), "upload points");
  return VD_OK;
}

int ensure_offsets(vd_ctx* ctx) {
  if (ctx->n_offsets) return VD_OK;
  const std::vector<int4> t = vd::knn_offset_table();
  VD_HIP(ctx->offsets.ensure(t.size()), "alloc offsets");
  VD_HIP(hipMemcpy(ctx->offsets.p, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice), "upload offsets");
  ctx->n_offsets = (int)t.size();
  return VD_OK;
}

}  // namespace

extern "C" {

void vd_config_default(vd_config* c) {
  if (!c) return;
  c->site_k = 90;
  c->tet_k = 0;
  c->pcap = 64;
  c->tcap = 96;
  c->accum = VD_ACCUM_F32;
  c->moments = VD_MOMENTS_REFERENCE;
  c->form = VD_FORM_SITE_MAJOR;
  c->grid_density = 3.1f;
  c->record_pairs = 0;
  c->clip_blocks = 0;
  c->site_tcap = 0;
  c->cells_pcap = 0;
  c->cells_tcap = 0;
  c->cull_depth = 16;
  c->srnr_k = 0;
  c->filter = 0;
  c->site_records = 0;
  c->site_depth = 0;
  c->knn_waves = 0;
}

int vd_ctx_create(vd_ctx** out, int device, const vd_config* cfg) {
  if (!out) return VD_ERR_ARG;
  *out = nullptr;
  vd_ctx* ctx = new vd_ctx();
  ctx->device = device;
  if (cfg)
    ctx->cfg = *cfg;
  else
    vd_config_default(&ctx->cfg);
  int rc = check_config(ctx, ctx->cfg);
  if (rc != VD_OK) {
    fprintf(stderr, "vd_ctx_create: %s\n", ctx->err.c_str());
    delete ctx;
    return rc;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    fprintf(stderr, "vd_ctx_create: no HIP device %d (found %d)\n", device, ndev);
    delete ctx;
    return VD_ERR_HIP;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    fprintf(stderr, "vd_ctx_create: cannot create a stream on device %d\n", device);
    delete ctx;
    return VD_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) {
    fprintf(stderr, "vd_ctx_create: cannot create the kNN step's second stream on device %d\n", device);
    vd_ctx_destroy(ctx);
    return VD_ERR_HIP;
  }
  for (auto& row : ctx->ev)
    for (auto& e : row) (void)hipEventCreate(&e);
  if (ensure_offsets(ctx) != VD_OK || ctx->hist.ensure(8) != hipSuccess) {
    fprintf(stderr, "vd_ctx_create: %s\n", ctx->err.c_str());
    vd_ctx_destroy(ctx);
    return VD_ERR_HIP;
  }
  *out = ctx;
  return VD_OK;
}

void vd_ctx_destroy(vd_ctx* ctx) {
  if (!ctx) return;
  ctx->drop_graph();
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  ctx->verts.release(); ctx->sites.release(); ctx->tet_xyz.release(); ctx->tet_prep.release();
  ctx->b_site_begin.release(); ctx->b_voff.release(); ctx->b_cell_base.release(); ctx->b_site_inst.release(); ctx->b_tet_inst.release();
  ctx->b_box.release(); ctx->tets.release(); ctx->offsets.release();
  ctx->shape.release(); ctx->g_count.release(); ctx->g_start.release(); ctx->g_cursor.release();
  ctx->g_cell_of.release(); ctx->g_sorted_ids.release(); ctx->g_sorted_pts.release();
  ctx->g_scan_tmp.release(); ctx->p_scan_tmp.release();
  ctx->site_knn.release(); ctx->tet_knn.release();
  ctx->p_count.release(); ctx->p_start.release(); ctx->p_cursor.release(); ctx->p_ids.release(); ctx->p_surv.release();
  ctx->acc_f32.release(); ctx->acc_fix.release(); ctx->moments.release();
  ctx->pair_status.release(); ctx->pair_vals.release(); ctx->hist.release();
  ctx->retry_list.release(); ctx->retry_count.release(); ctx->retry_list2.release(); ctx->retry_count2.release(); ctx->site_counter.release();
  ctx->srnr_list.release(); ctx->srnr_count.release(); ctx->srnr_row_of.release(); ctx->srnr_row_site.release();
  ctx->knn2.release();
  ctx->rec_meta.release(); ctx->rec_tris.release(); ctx->rec_count.release();
  ctx->c_status.release(); ctx->c_nv.release(); ctx->c_nt.release(); ctx->c_r2.release();
  ctx->c_planes.release(); ctx->c_tris.release(); ctx->c_redo.release();
  for (auto& row : ctx->ev)
    for (auto& e : row)
      if (e) (void)hipEventDestroy(e);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* vd_ctx_error(const vd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vd_ctx_set_config(vd_ctx* ctx, const vd_config* cfg) {
  if (!ctx || !cfg) return VD_ERR_ARG;
  ctx->drop_graph();
  int rc = check_config(ctx, *cfg);
  if (rc != VD_OK) return rc;
  const bool grid_changed = cfg->grid_density != ctx->cfg.grid_density;
  ctx->cfg = *cfg;
  if (grid_changed && ctx->n_sites > 0) {
    rc = ensure_grid(ctx, ctx->n_sites);
    if (rc != VD_OK) return rc;
  }
  ctx->have_knn = false;
  ctx->have_clip = false;
  return VD_OK;
}

namespace {
// Uploads vertices, tets (global vertex ids, validated by the caller) and the
// per-tet corner records; sets the mesh bounding box.
int load_mesh(vd_ctx* ctx, const float* verts_xyz, int n_verts, const int32_t* tets, int n_tets) {
  for (int a = 0; a < 3; ++a) {
    ctx->mesh_lo[a] = n_verts ? verts_xyz[a] : 0.f;
    ctx->mesh_hi[a] = n_verts ? verts_xyz[a] : 0.f;
  }
  for (int i = 0; i < n_verts; ++i)
    for (int a = 0; a < 3; ++a) {
      ctx->mesh_lo[a] = std::min(ctx->mesh_lo[a], verts_xyz[3 * i + a]);
      ctx->mesh_hi[a] = std::max(ctx->mesh_hi[a], verts_xyz[3 * i + a]);
    }
  int rc = upload_xyz(ctx, ctx->verts, verts_xyz, n_verts);
  if (rc != VD_OK) return rc;
  VD_HIP(ctx->tets.ensure(std::max(n_tets, 1)), "alloc tets");
  if (n_tets > 0)
    VD_HIP(hipMemcpy(ctx->tets.p, tets, sizeof(int4) * (size_t)n_tets, hipMemcpyHostToDevice), "upload tets");
  {
    std::vector<float> rec((size_t)12 * std::max(n_tets, 1));
    for (size_t t = 0; t < (size_t)n_tets; ++t)
      for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) rec[12 * t + 3 * k + a] = verts_xyz[3 * (size_t)tets[4 * t + k] + a];
    VD_HIP(ctx->tet_xyz.ensure((size_t)3 * std::max(n_tets, 1)), "alloc tet records");
    VD_HIP(hipMemcpy(ctx->tet_xyz.p, rec.data(), sizeof(float) * rec.size(), hipMemcpyHostToDevice),
           "upload tet records");
    VD_HIP(ctx->tet_prep.ensure((size_t)vd::kTetPrepWords * std::max(n_tets, 1)), "alloc tet records");
    VD_HIP(vd::tet_prep(ctx->tet_xyz.p, n_tets, ctx->tet_prep.p, ctx->stream), "tet records");
  }
  ctx->n_verts = n_verts;
  ctx->n_tets = n_tets;
  ctx->have_knn = ctx->have_clip = false;
  ctx->rows_valid = false;
  return VD_OK;
}
}  // namespace

int vd_set_mesh(vd_ctx* ctx, const float* verts_xyz, int n_verts, const int32_t* tets, int n_tets) {
  if (!ctx || n_verts < 0 || n_tets < 0 || (n_verts && !verts_xyz) || (n_tets && !tets))
    return ctx ? ctx->fail(VD_ERR_ARG, "vd_set_mesh: bad arguments") : VD_ERR_ARG;
  VD_HIP(hipSetDevice(ctx->device), "set device");
  for (long long i = 0; i < 4LL * n_tets; ++i)
    if (tets[i] < 0 || tets[i] >= n_verts)
      return ctx->fail(VD_ERR_PRECONDITION, "tet %lld refers to vertex %d of %d", i / 4, tets[i], n_verts);
  ctx->drop_graph();
  ctx->inst.clear();
  const int rc = load_mesh(ctx, verts_xyz, n_verts, tets, n_tets);
  if (rc == VD_OK) ctx->fixed_load = vdi::fixed_point_load(verts_xyz, 0, n_verts, tets, 0, n_tets);
  return rc;
}

int vd_set_batch(vd_ctx* ctx, int n_inst, const int32_t* n_verts, const int32_t* n_tets, const int32_t* n_sites,
                 const float* verts_xyz, const int32_t* tets, const float* sites_xyz) {
  if (!ctx || n_inst < 1 || !n_verts || !n_tets || !n_sites)
    return ctx ? ctx->fail(VD_ERR_ARG, "vd_set_batch: bad arguments") : VD_ERR_ARG;
  VD_HIP(hipSetDevice(ctx->device), "set device");
  std::vector<vd_ctx::Inst> in((size_t)n_inst);
  long long V = 0, T = 0, S = 0;
  for (int i = 0; i < n_inst; ++i) {
    if (n_verts[i] < 0 || n_tets[i] < 0 || n_sites[i] < 1)
      return ctx->fail(VD_ERR_ARG, "vd_set_batch: instance %d has bad sizes", i);
    in[i] = {(int)V, n_verts[i], (int)T, n_tets[i], (int)S, n_sites[i]};
    V += n_verts[i];
    T += n_tets[i];
    S += n_sites[i];
    if (V > 0x7fffffffLL || T > 0x7fffffffLL || S > 0x7fffffffLL) return ctx->fail(VD_ERR_ARG, "vd_set_batch: too large");
  }
  if ((V && !verts_xyz) || (T && !tets) || !sites_xyz) return ctx->fail(VD_ERR_ARG, "vd_set_batch: bad arguments");
  ctx->drop_graph();
  // tets to global vertex ids; per-instance boxes; instance of every site and tet
  std::vector<int32_t> gt((size_t)4 * T);
  std::vector<float> box((size_t)6 * n_inst);
  std::vector<int> site_begin((size_t)n_inst + 1), site_inst((size_t)S), tet_inst((size_t)std::max(T, 1LL));
  for (int i = 0; i < n_inst; ++i) {
    const vd_ctx::Inst& q = in[i];
    for (long long j = 4LL * q.t0; j < 4LL * (q.t0 + q.nt); ++j) {
      if (tets[j] < 0 || tets[j] >= q.nv)
        return ctx->fail(VD_ERR_PRECONDITION, "instance %d: tet %lld refers to vertex %d of %d", i, j / 4 - q.t0,
                         tets[j], q.nv);
      gt[j] = tets[j] + q.v0;
    }
    for (int a = 0; a < 3; ++a) {
      float lo = q.nv ? verts_xyz[3 * (size_t)q.v0 + a] : 0.f, hi = lo;
      for (int v = q.v0; v < q.v0 + q.nv; ++v) {
        lo = std::min(lo, verts_xyz[3 * (size_t)v + a]);
        hi = std::max(hi, verts_xyz[3 * (size_t)v + a]);
      }
      box[6 * i + a] = lo;
      box[6 * i + 3 + a] = hi;
    }
    site_begin[i] = q.s0;
    for (int k = q.s0; k < q.s0 + q.ns; ++k) site_inst[k] = i;
    for (int k = q.t0; k < q.t0 + q.nt; ++k) tet_inst[k] = i;
  }
  site_begin[n_inst] = (int)S;
  double fixed_load = 0.0;
  for (int i = 0; i < n_inst; ++i)
    fixed_load = std::max(fixed_load, vdi::fixed_point_load(verts_xyz, in[i].v0, in[i].nv, tets, in[i].t0, in[i].nt));
  int rc = load_mesh(ctx, verts_xyz, (int)V, gt.data(), (int)T);
  ctx->fixed_load = fixed_load;
  if (rc != VD_OK) return rc;
  rc = upload_xyz(ctx, ctx->sites, sites_xyz, (int)S);
  if (rc != VD_OK) return rc;
  ctx->n_sites = (int)S;
  ctx->inst = in;
  if (n_inst > 1) {
    VD_HIP(ctx->b_site_begin.ensure(site_begin.size()), "alloc batch");
    VD_HIP(ctx->b_site_inst.ensure(site_inst.size()), "alloc batch");
    VD_HIP(ctx->b_tet_inst.ensure(tet_inst.size()), "alloc batch");
    VD_HIP(ctx->b_box.ensure(box.size()), "alloc batch");
    VD_HIP(hipMemcpy(ctx->b_site_begin.p, site_begin.data(), sizeof(int) * site_begin.size(), hipMemcpyHostToDevice),
           "upload batch");
    VD_HIP(hipMemcpy(ctx->b_site_inst.p, site_inst.data(), sizeof(int) * site_inst.size(), hipMemcpyHostToDevice),
           "upload batch");
    VD_HIP(hipMemcpy(ctx->b_tet_inst.p, tet_inst.data(), sizeof(int) * tet_inst.size(), hipMemcpyHostToDevice),
           "upload batch");
    VD_HIP(hipMemcpy(ctx->b_box.p, box.data(), sizeof(float) * box.size(), hipMemcpyHostToDevice), "upload batch");
  }
  rc = ensure_grid(ctx, std::max(ctx->n_sites, 1));
  if (rc != VD_OK) return rc;
  ctx->have_knn = ctx->have_clip = false;
  ctx->fresh_sites = true;
  ctx->knn_since_set = 0;
  return VD_OK;
}

int vd_get_batch(vd_ctx* ctx, int* n_inst) {
  if (!ctx || !n_inst) return VD_ERR_ARG;
  *n_inst = std::max<int>(1, (int)ctx->inst.size());
  return VD_OK;
}

int vd_set_sites(vd_ctx* ctx, const float* sites_xyz, int n_sites) {
  if (!ctx || n_sites < 0 || (n_sites && !sites_xyz)) return ctx ? ctx->fail(VD_ERR_ARG, "vd_set_sites: bad arguments") : VD_ERR_ARG;
  VD_HIP(hipSetDevice(ctx->device), "set device");
  if (ctx->batch() && n_sites != ctx->n_sites)
    return ctx->fail(VD_ERR_STATE, "vd_set_sites: a batch holds %d sites, got %d", ctx->n_sites, n_sites);
  if (n_sites != ctx->n_sites) {
    ctx->drop_graph();
    ctx->rows_valid = false;
  }
  if (!ctx->batch()) ctx->inst.clear();
  int rc = upload_xyz(ctx, ctx->sites, sites_xyz, n_sites);
  if (rc != VD_OK) return rc;
  rc = ensure_grid(ctx, std::max(n_sites, 1));
  if (rc != VD_OK) return rc;
  ctx->n_sites = n_sites;
  ctx->have_knn = ctx->have_clip = false;
  ctx->fresh_sites = true;
  ctx->knn_since_set = 0;
  return VD_OK;
}

int vd_get_sites(vd_ctx* ctx, float* out) {
  if (!ctx || (!out && ctx->n_sites)) return VD_ERR_ARG;
  std::vector<float4> h((size_t)std::max(ctx->n_sites, 1));
  if (ctx->n_sites)
    VD_HIP(hipMemcpyAsync(h.data(), ctx->sites.p, sizeof(float4) * ctx->n_sites, hipMemcpyDeviceToHost, ctx->stream),
           "read sites");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read sites");
  for (int i = 0; i < ctx->n_sites; ++i) {
    out[3 * i] = h[i].x;
    out[3 * i + 1] = h[i].y;
    out[3 * i + 2] = h[i].z;
  }
  return VD_OK;
}

int vd_get_dims(vd_ctx* ctx, int* k, int* tk, int* n_sites, int* n_tets) {
  if (!ctx) return VD_ERR_ARG;
  int K = std::max(0, std::min(ctx->cfg.site_k, ctx->n_sites - 1));
  int TK = std::max(0, std::min(tet_k_heuristic(ctx->n_sites, ctx->n_tets, ctx->cfg.tet_k), ctx->n_sites));
  if (ctx->batch()) {
    // one K and TK for all instances: K from the smallest site set, TK the
    // instances' common value (-1 when they differ)
    K = ctx->cfg.site_k;
    TK = -2;
    for (const auto& q : ctx->inst) {
      K = std::min(K, q.ns - 1);
      const int t = std::max(0, std::min(tet_k_heuristic(q.ns, q.nt, ctx->cfg.tet_k), q.ns));
      TK = TK == -2 || TK == t ? t : -1;
    }
    K = std::max(K, 0);
  }
  if (k) *k = K;
  if (tk) *tk = TK;
  if (n_sites) *n_sites = ctx->n_sites;
  if (n_tets) *n_tets = ctx->n_tets;
  return VD_OK;
}

int vd_knn_step(vd_ctx* ctx) {
  if (!ctx) return VD_ERR_ARG;
  if (ctx->n_sites <= 0) return ctx->fail(VD_ERR_STATE, "vd_knn_step: no sites");
  VD_HIP(hipSetDevice(ctx->device), "set device");
  vd_get_dims(ctx, &ctx->K, &ctx->TK, nullptr, nullptr);
  if (ctx->TK < 0)
    return ctx->fail(VD_ERR_ARG, "the instances of the batch have different candidate counts; set tet_k");
  const int n = ctx->n_sites;
  VD_HIP(ctx->site_knn.ensure((size_t)n * std::max(ctx->K, 1)), "alloc site knn");
  VD_HIP(ctx->tet_knn.ensure((size_t)std::max(ctx->n_tets, 1) * std::max(ctx->TK, 1)), "alloc tet knn");
  VD_HIP(record_phase(ctx, (0)), "event");
  const vd::GridBufs gb = grid_bufs(ctx);
  VD_HIP(vd::grid_build(gb, ctx->sites.p, n, ctx->n_verts ? ctx->mesh_lo : nullptr,
                        ctx->n_verts ? ctx->mesh_hi : nullptr, ctx->cfg.grid_density, ctx->stream),
         "grid build");
  // The site and the tet queries read the same grid and write tables of their
  // own: with both, the site queries run on the second stream beside the tet
  // queries, so that neither launch's last partial wave of workgroups leaves
  // the GPU idle, and the step ends when both have.
  const bool tets = ctx->n_tets > 0 && ctx->TK > 0;
  const bool fork = ctx->K > 0 && tets;
  // the rows of the last kNN step bound this one's queries (same points, moved)
  // The previous rows bound each query (any k of its instance's points do)
  // but cost a gather per row; when the sites have moved far since those rows
  // were made the bound is loose and costs more than it saves.  Sites set by
  // the caller (random samples, typically) and their first Lloyd update move
  // far: kNN iteration 1 2.31-2.50 ms with the bounds against 2.17-2.21
  // without, iteration 2 2.26-2.30 against 2.15-2.20, iteration 3 2.08-2.11
  // either way (1M tets, 100k uniform sites).
  const float4* hint = ctx->rows_valid && ctx->rows_K == ctx->K && ctx->rows_TK == ctx->TK &&
                               ctx->knn_since_set >= kHintAfter
                           ? ctx->sites.p
                           : nullptr;
  ++ctx->knn_since_set;
  if (fork) {
    VD_HIP(hipEventRecord(ctx->fork_ev, ctx->stream), "event");
    VD_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0), "stream wait");
  }
  if (ctx->K > 0)
    VD_HIP(vd::knn_query(gb, n, ctx->offsets.p, ctx->n_offsets, ctx->K, 0, n, nullptr, nullptr, nullptr, nullptr,
                         ctx->site_knn.p, 0, fork ? ctx->side : ctx->stream, hint, ctx->cfg.knn_waves),
           "site knn");
  if (fork) VD_HIP(hipEventRecord(ctx->join_ev, ctx->side), "event");
  if (tets)
    VD_HIP(vd::knn_query(gb, n, ctx->offsets.p, ctx->n_offsets, ctx->TK, 1, ctx->n_tets, ctx->verts.p, ctx->tets.p,
                         nullptr, ctx->batch() ? ctx->b_tet_inst.p : nullptr, ctx->tet_knn.p, 0, ctx->stream, hint,
                         ctx->cfg.knn_waves),
           "tet knn");
  if (fork) VD_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0), "stream wait");
  VD_HIP(record_phase(ctx, (1)), "event");
  ctx->rows_valid = true;
  ctx->rows_K = ctx->K;
  ctx->rows_TK = ctx->TK;
  ctx->ev_knn = true;
  ctx->ev_pairs = ctx->ev_clip = ctx->ev_update = false;
  ctx->have_knn = true;
  ctx->have_clip = false;
  return VD_OK;
}

int vd_clip_step(vd_ctx* ctx) {
  if (!ctx) return VD_ERR_ARG;
  if (!ctx->have_knn) return ctx->fail(VD_ERR_STATE, "vd_clip_step: run vd_knn_step first");
  VD_HIP(hipSetDevice(ctx->device), "set device");
  const vd_config& c = ctx->cfg;
  const int n = ctx->n_sites;
  const long long n_pairs = (long long)ctx->n_tets * ctx->TK;
  if (n_pairs > 0xffffffffLL) return ctx->fail(VD_ERR_ARG, "more than 2^32 - 1 pairs in one context");
  if (c.accum == VD_ACCUM_F64 && ctx->fixed_load > vdi::kFixedPointMaxLoad)
    return ctx->fail(VD_ERR_PRECONDITION,
                     "mesh volume x largest |coordinate| = %.3g exceeds %.3g, the range of the exact FP64 "
                     "fixed-point sums: map the mesh into the [0,1000]^3 frame or use VD_ACCUM_F32",
                     ctx->fixed_load, vdi::kFixedPointMaxLoad);
  if (c.accum == VD_ACCUM_F32) {
    VD_HIP(ctx->acc_f32.ensure(4 * (size_t)n), "alloc acc");
    VD_HIP(vd::fill_u32(ctx->acc_f32.p, 4LL * n, 0u, ctx->stream), "clear acc");
  } else {
    VD_HIP(ctx->acc_fix.ensure(4 * (size_t)n), "alloc acc");
    VD_HIP(vd::fill_u32(ctx->acc_fix.p, 8LL * n, 0u, ctx->stream), "clear acc");
  }
  VD_HIP(vd::fill_u32(ctx->hist.p, 16, 0u, ctx->stream), "clear hist");
  if (c.record_pairs) {
    VD_HIP(ctx->pair_status.ensure(std::max(n_pairs, 1LL)), "alloc pair status");
    VD_HIP(ctx->pair_vals.ensure(4 * std::max(n_pairs, 1LL)), "alloc pair vals");
  }
  vd::ClipArgs a{};
  a.sites = ctx->sites.p;
  a.site_knn = ctx->site_knn.p;
  a.K = ctx->K;
  a.verts = ctx->verts.p;
  a.tets = ctx->tets.p;
  a.tet_xyz = ctx->tet_xyz.p;
  a.tet_prep = ctx->tet_prep.p;
  a.tet_knn = ctx->tet_knn.p;
  a.TK = ctx->TK;
  a.n_pairs = n_pairs;
  a.n_sites = n;
  a.pcap = c.pcap;
  a.tcap = c.tcap;
  a.acc_f32 = ctx->acc_f32.p;
  a.acc_fix = ctx->acc_fix.p;
  a.pair_status = c.record_pairs ? ctx->pair_status.p : nullptr;
  a.pair_vals = c.record_pairs ? ctx->pair_vals.p : nullptr;
  a.status_hist = ctx->hist.p;
  a.filter = c.filter;
  int blocks = 0;
  if (c.form == VD_FORM_SITE_MAJOR || c.form == VD_FORM_CELLS) {
    VD_HIP(ctx->p_count.ensure(n), "alloc pairs");
    VD_HIP(ctx->p_start.ensure((size_t)n + 1), "alloc pairs");
    VD_HIP(ctx->p_cursor.ensure(n), "alloc pairs");
    VD_HIP(ctx->p_ids.ensure(std::max(n_pairs, 1LL)), "alloc pairs");
    VD_HIP(ctx->p_scan_tmp.ensure(vd::scan_tmp_ints(n)), "alloc pairs");
    VD_HIP(vd::fill_u32(ctx->p_count.p, n, 0u, ctx->stream), "clear pairs");
    VD_HIP(vd::pairs_by_site(ctx->tet_knn.p, n_pairs, n, ctx->p_count.p, ctx->p_start.p, ctx->p_cursor.p, ctx->p_ids.p,
                             ctx->p_scan_tmp.p, ctx->stream),
           "pairs by site");
    VD_HIP(record_phase(ctx, (2)), "event");
    ctx->ev_pairs = true;
    a.pair_start = ctx->p_start.p;
    a.pair_ids = ctx->p_ids.p;
    // sites in the kNN grid's cell order: nearby sites share most of their tets
    a.site_order = ctx->g_sorted_ids.p;
    if (c.form == VD_FORM_CELLS) {
      // kernel B (every site's cell), then kernel C (cell x tet per pair)
      VD_HIP(ctx->c_status.ensure(n), "alloc cells");
      VD_HIP(ctx->c_nv.ensure(n), "alloc cells");
      VD_HIP(ctx->c_nt.ensure(n), "alloc cells");
      VD_HIP(ctx->c_r2.ensure(n), "alloc cells");
      VD_HIP(ctx->c_planes.ensure((size_t)n * c.pcap), "alloc cells");
      VD_HIP(ctx->c_tris.ensure((size_t)n * c.tcap), "alloc cells");
      VD_HIP(ctx->c_redo.ensure((size_t)n + 1), "alloc cells");
      vd::CellStore cs{ctx->c_status.p, ctx->c_nv.p, ctx->c_nt.p, ctx->c_r2.p, ctx->c_planes.p, ctx->c_tris.p,
                       ctx->c_redo.p};
      a.cull_depth = c.cull_depth;
      if (c.cull_depth > 0) VD_HIP(ctx->p_surv.ensure(n), "alloc pairs");
      VD_HIP(vd::cells_launch(a, cs, c.moments, c.accum, ctx->stream, ctx->p_ids.p, c.cull_depth > 0 ? ctx->p_surv.p : nullptr,
                              c.cells_pcap, c.cells_tcap),
             "cells");
      VD_HIP(record_phase(ctx, (5)), "event");
      VD_HIP(record_phase(ctx, (3)), "event");
      ctx->ev_clip = true;
      ctx->have_clip = true;
      return VD_OK;
    }
    blocks = c.clip_blocks;
    VD_HIP(ctx->site_counter.ensure(1), "alloc site counter");
    VD_HIP(vd::fill_u32(ctx->site_counter.p, 1, 0u, ctx->stream), "clear site counter");
    a.site_counter = ctx->site_counter.p;
    const int fast_tcap = vdi::site_tcap_of(c);
    a.site_depth = site_depth_of(c, ctx->K);
    if (ctx->fresh_sites && c.site_depth == 0) a.site_depth = std::min(kFreshSiteDepth, ctx->K);
    ctx->fresh_sites = false;
    // (the robust moments also need it: the site-major kernel retries the
    // polytopes it cannot hand to the volume pass; and it is allocated alike
    // at either depth, so that an iteration captured in a graph after a first
    // eager one allocates nothing)
    if (fast_tcap < c.tcap || std::min(a.site_depth, site_depth_of(c, ctx->K)) < ctx->K ||
        c.moments == VD_MOMENTS_ROBUST) {
      VD_HIP(ctx->retry_list.ensure(std::max(n_pairs, 1LL)), "alloc retry list");
      VD_HIP(ctx->retry_count.ensure(1), "alloc retry list");
      VD_HIP(vd::fill_u32(ctx->retry_count.p, 1, 0u, ctx->stream), "clear retry count");
      a.retry_list = ctx->retry_list.p;
      a.retry_count = ctx->retry_count.p;
    }
    a.tcap = fast_tcap;
    a.tcap_full = c.tcap;
    // (a depth at which every skipped pair is what the configured caps give)
    a.cull_depth = vd::exact_cull_depth(c.cull_depth, c.pcap, c.tcap);
    if (std::min(a.cull_depth, a.site_depth) > 0) {
      // the emptiness test as a pass of its own; the clipping takes the survivors
      VD_HIP(ctx->p_surv.ensure(n), "alloc pairs");
      VD_HIP(vd::cull_sites(a, ctx->p_ids.p, ctx->p_surv.p, ctx->stream), "cull");
      a.pair_count = ctx->p_surv.p;
      a.cull_depth = 0;
    }
    if (c.site_records >= 0) {
      // automatic: one pair in ten, plus the slots the waves reserve ahead
      const long long want = c.site_records > 0 ? c.site_records : n_pairs / 10 + 262144;
      const unsigned cap = (unsigned)std::max(1LL, std::min({want, n_pairs, (long long)UINT32_MAX}));
      VD_HIP(ctx->rec_meta.ensure(cap), "alloc records");
      VD_HIP(ctx->rec_tris.ensure((size_t)cap * vd::rec_stride(fast_tcap)), "alloc records");
      VD_HIP(ctx->rec_count.ensure(1), "alloc records");
      VD_HIP(vd::fill_u32(ctx->rec_count.p, 1, 0u, ctx->stream), "clear record count");
      a.rec_meta = ctx->rec_meta.p;
      a.rec_tris = ctx->rec_tris.p;
      a.rec_count = ctx->rec_count.p;
      a.rec_cap = cap;
    }
  } else {
    VD_HIP(record_phase(ctx, (2)), "event");
    ctx->ev_pairs = true;
  }
  // pairs whose cell misses the security radius: deferred, then clipped again
  // with deeper neighbour lists (srnr_k)
  int K2 = 0;
  if (c.srnr_k > 0) {
    K2 = std::min(c.srnr_k, n - 1);
    for (const auto& q : ctx->inst) K2 = std::min(K2, q.ns - 1);
    if (K2 <= ctx->K) K2 = 0;
  }
  if (K2 > 0) {
    VD_HIP(ctx->srnr_list.ensure(std::max(n_pairs, 1LL)), "alloc srnr list");
    VD_HIP(ctx->srnr_count.ensure(2), "alloc srnr list");
    VD_HIP(vd::fill_u32(ctx->srnr_count.p, 2, 0u, ctx->stream), "clear srnr list");
    a.srnr_list = ctx->srnr_list.p;
    a.srnr_count = ctx->srnr_count.p;
  }
  // The site-major form's re-clip of its overflow pairs depends on the
  // clipping launch only: it runs on the second stream beside the volume pass
  // over the clipping's records (both only add to the site sums), and the
  // step waits for both.
  const bool vol_pass = c.form == VD_FORM_SITE_MAJOR && a.rec_meta && a.rec_cap > 0;
  const bool beside = vol_pass && a.retry_list;
  VD_HIP(vd::clip_launch(a, c.form, c.moments, c.accum, blocks, ctx->stream, !beside), "clip");
  hipStream_t rs = ctx->stream;
  if (beside) {
    VD_HIP(hipEventRecord(ctx->fork_ev, ctx->stream), "event");
    VD_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0), "stream wait");
    rs = ctx->side;
    VD_HIP(vd::site_volume_pass(a, c.moments, c.accum, ctx->stream), "volume pass");
  }
  VD_HIP(record_phase(ctx, (5)), "event");
  if (a.retry_list) {
    // pairs that overflowed the site-major launch's vertex cap or needed more
    // bisectors than its tables hold: re-clipped by the tet-major kernel with
    // cached vertex distances at caps of 32 planes and 32 vertices (3 waves
    // per CU instead of 1 at the configured 64 planes), and those that
    // overflow that too by the plain one at the configured caps
    vd::ClipArgs r = a;
    r.pair_list = a.retry_list;
    r.pair_list_count = a.retry_count;
    r.retry_list = nullptr;
    r.retry_count = nullptr;
    const int mid = std::min(c.tcap, 32), pmid = std::min(c.pcap, 32);
    if (mid >= a.tcap && vd::clip_plain_cached_lds(pmid, mid) <= 160 * 1024) {
      r.tcap = mid;
      r.tcap_full = c.tcap;
      r.pcap = pmid;
      r.pcap_full = c.pcap;
      if (mid < c.tcap || pmid < c.pcap) {
        VD_HIP(ctx->retry_list2.ensure(std::max(n_pairs, 1LL)), "alloc retry list");
        VD_HIP(ctx->retry_count2.ensure(1), "alloc retry list");
        VD_HIP(vd::fill_u32(ctx->retry_count2.p, 1, 0u, rs), "clear retry count");
        r.retry_list = ctx->retry_list2.p;
        r.retry_count = ctx->retry_count2.p;
      }
      VD_HIP(vd::clip_launch(r, vd::kFormTetCached, c.moments, c.accum, 768, rs), "clip retry");
      if (mid < c.tcap || pmid < c.pcap) {
        vd::ClipArgs r3 = a;
        r3.tcap = c.tcap;
        r3.pair_list = ctx->retry_list2.p;
        r3.pair_list_count = ctx->retry_count2.p;
        r3.retry_list = nullptr;
        r3.retry_count = nullptr;
        VD_HIP(vd::clip_launch(r3, VD_FORM_TET_MAJOR, c.moments, c.accum, 256, rs), "clip retry");
      }
    } else {
      r.tcap = c.tcap;
      VD_HIP(vd::clip_launch(r, VD_FORM_TET_MAJOR, c.moments, c.accum, 256, rs), "clip retry");
    }
  }
  if (beside) {
    VD_HIP(hipEventRecord(ctx->join_ev, ctx->side), "event");
    VD_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0), "stream wait");
  }
  if (K2 > 0) {
    VD_HIP(ctx->srnr_row_of.ensure(n), "alloc srnr rows");
    VD_HIP(ctx->srnr_row_site.ensure(n), "alloc srnr rows");
    VD_HIP(ctx->knn2.ensure((size_t)n * K2), "alloc srnr rows");
    VD_HIP(vd::fill_u32(ctx->srnr_row_of.p, n, 0xffffffffu, ctx->stream), "clear srnr rows");
    unsigned* n_rows = ctx->srnr_count.p + 1;
    VD_HIP(vd::srnr_rows(ctx->srnr_list.p, ctx->srnr_count.p, n_pairs, ctx->tet_knn.p, n, ctx->srnr_row_of.p,
                         ctx->srnr_row_site.p, n_rows, ctx->stream),
           "srnr rows");
    VD_HIP(vd::knn_query_ids(grid_bufs(ctx), ctx->offsets.p, ctx->n_offsets, K2, ctx->sites.p, ctx->srnr_row_site.p,
                             n_rows, n, ctx->knn2.p, ctx->stream),
           "srnr knn");
    vd::ClipArgs r = a;
    r.tcap = c.tcap;
    r.pair_list = ctx->srnr_list.p;
    r.pair_list_count = ctx->srnr_count.p;
    r.retry_list = nullptr;
    r.retry_count = nullptr;
    r.srnr_list = nullptr;
    r.srnr_count = nullptr;
    r.knn2 = ctx->knn2.p;
    r.K2 = K2;
    r.knn2_row = ctx->srnr_row_of.p;
    VD_HIP(vd::clip_launch(r, VD_FORM_TET_MAJOR, c.moments, c.accum, 1024, ctx->stream), "clip srnr");
  }
  VD_HIP(record_phase(ctx, (3)), "event");
  ctx->ev_clip = true;
  ctx->have_clip = true;
  return VD_OK;
}

int vd_update_step(vd_ctx* ctx) {
  if (!ctx) return VD_ERR_ARG;
  if (!ctx->have_clip) return ctx->fail(VD_ERR_STATE, "vd_update_step: run vd_clip_step first");
  VD_HIP(vd::lloyd_update(ctx->sites.p, ctx->n_sites, ctx->cfg.accum, ctx->acc_f32.p, ctx->acc_fix.p, ctx->stream),
         "update");
  VD_HIP(record_phase(ctx, (4)), "event");
  ctx->ev_update = true;
  if (!ctx->capturing) {
    ctx->graph_slot[ctx->it_slot] = false;
    ctx->it_slot = (ctx->it_slot + 1) % vd_ctx::kRing;
    ctx->it_done++;
  }
  ctx->ev_knn = ctx->ev_pairs = ctx->ev_clip = false;
  ctx->have_knn = false;  // sites moved
  return VD_OK;
}

int vd_lloyd(vd_ctx* ctx, int n_iter) {
  if (!ctx || n_iter < 0) return VD_ERR_ARG;
  for (int i = 0; i < n_iter; ++i) {
    int rc = vd_knn_step(ctx);
    if (rc == VD_OK) rc = vd_clip_step(ctx);
    if (rc == VD_OK) rc = vd_update_step(ctx);
    if (rc != VD_OK) return rc;
  }
  return VD_OK;
}

int vd_lloyd_graph(vd_ctx* ctx, int n_iter) {
  if (!ctx || n_iter < 0) return VD_ERR_ARG;
  if (n_iter == 0) return VD_OK;
  VD_HIP(hipSetDevice(ctx->device), "set device");
  if (!ctx->graph_exec) {
    // an eager iteration sizes every buffer, so that the captured one
    // allocates nothing
    int rc = vd_lloyd(ctx, 1);
    if (rc != VD_OK) return rc;
    if (--n_iter == 0) return VD_OK;
    VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
    // the captured iteration stands for every later one: it takes the
    // previous rows' bounds whenever the rows are valid
    ctx->knn_since_set = std::max(ctx->knn_since_set, kHintAfter);
    VD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    ctx->capturing = true;
    rc = vd_knn_step(ctx);
    if (rc == VD_OK) rc = vd_clip_step(ctx);
    if (rc == VD_OK) rc = vd_update_step(ctx);
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != VD_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc != VD_OK ? rc : ctx->hip_fail(e, "end capture");
    }
    ctx->graph = g;
    const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      ctx->drop_graph();
      return ctx->hip_fail(ei, "instantiate graph");
    }
  }
  for (int i = 0; i < n_iter; ++i) {
    VD_HIP(hipEventRecord(ctx->E(0), ctx->stream), "event");
    VD_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream), "graph launch");
    VD_HIP(hipEventRecord(ctx->E(4), ctx->stream), "event");
    ctx->graph_slot[ctx->it_slot] = true;
    ctx->it_slot = (ctx->it_slot + 1) % vd_ctx::kRing;
    ctx->it_done++;
  }
  ctx->have_knn = false;
  ctx->have_clip = true;
  ctx->ev_knn = ctx->ev_pairs = ctx->ev_clip = false;
  return VD_OK;
}

int vd_autotune(vd_ctx* ctx, int iters, vd_config* chosen) {
  if (!ctx || iters < 1) return VD_ERR_ARG;
  if (ctx->n_sites <= 0) return ctx->fail(VD_ERR_STATE, "vd_autotune: no sites");
  VD_HIP(hipSetDevice(ctx->device), "set device");
  DevBuf<float4> saved;
  VD_HIP(saved.ensure(ctx->n_sites), "alloc");
  VD_HIP(hipMemcpyAsync(saved.p, ctx->sites.p, sizeof(float4) * ctx->n_sites, hipMemcpyDeviceToDevice, ctx->stream),
         "save sites");
  auto restore = [&]() -> int {
    VD_HIP(hipMemcpyAsync(ctx->sites.p, saved.p, sizeof(float4) * ctx->n_sites, hipMemcpyDeviceToDevice, ctx->stream),
           "restore sites");
    VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
    ctx->have_knn = ctx->have_clip = false;
    return VD_OK;
  };
  // device time of `iters` iterations (after one untimed) under cfg, or a
  // negative value when cfg is rejected
  auto trial = [&](const vd_config& cfg, double* ms) -> int {
    *ms = -1.0;
    if (check_config(ctx, cfg) != VD_OK) return VD_OK;
    int rc = vd_ctx_set_config(ctx, &cfg);
    if (rc == VD_OK) rc = restore();
    if (rc == VD_OK) rc = vd_lloyd(ctx, 1);
    if (rc == VD_OK) rc = vd_timing_reset(ctx);
    if (rc == VD_OK) rc = vd_lloyd(ctx, iters);
    vd_timing t{};
    int n = 0;
    if (rc == VD_OK) rc = vd_get_timing_sum(ctx, &t, &n);
    if (rc != VD_OK) return rc;
    *ms = n > 0 ? t.total_ms / n : -1.0;
    return VD_OK;
  };
  vd_config best = ctx->cfg;
  double best_ms = 0.0;
  int rc = trial(best, &best_ms);
  if (rc != VD_OK) return rc;
  // coordinate search: one parameter at a time, the others at their best
  const int tcaps[] = {12, 16, 20, 24};
  const int depths[] = {8, 16, 32, 64};
  const float densities[] = {2.0f, 3.1f, 5.0f};
  const int clip_blocks[] = {1536, 3072, 4608};  // 0: as many as stay resident
  const int knn_waves[] = {2, 4, 8};
  for (int dim = 0; dim < 5; ++dim) {
    const int nv = dim == 2 || dim >= 3 ? 3 : 4;
    for (int j = 0; j < nv; ++j) {
      vd_config c = best;
      if (dim == 0) {
        if (c.form != VD_FORM_SITE_MAJOR) break;
        c.site_tcap = tcaps[j];
      } else if (dim == 1) {
        if (c.form != VD_FORM_SITE_MAJOR) break;
        c.cull_depth = depths[j];
      } else if (dim == 2) {
        c.grid_density = densities[j];
      } else if (dim == 3) {
        if (c.form != VD_FORM_SITE_MAJOR) break;
        c.clip_blocks = clip_blocks[j];
      } else {
        c.knn_waves = knn_waves[j];
      }
      if (std::memcmp(&c, &best, sizeof c) == 0) continue;
      double ms = 0.0;
      if ((rc = trial(c, &ms)) != VD_OK) return rc;
      if (ms > 0.0 && ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  if ((rc = vd_ctx_set_config(ctx, &best)) != VD_OK) return rc;
  if ((rc = restore()) != VD_OK) return rc;
  VD_HIP(vd_timing_reset(ctx) == VD_OK ? hipSuccess : hipErrorUnknown, "timing reset");
  if (chosen) *chosen = best;
  return VD_OK;
}

int vd_autotune_k(vd_ctx* ctx, int iters, double tol, vd_config* chosen, double* err_out) {
  if (!ctx || iters < 1 || !(tol >= 0.0)) return VD_ERR_ARG;
  if (ctx->n_sites <= 0) return ctx->fail(VD_ERR_STATE, "vd_autotune_k: no sites");
  VD_HIP(hipSetDevice(ctx->device), "set device");
  const int n = ctx->n_sites;
  DevBuf<float4> saved;
  VD_HIP(saved.ensure(n), "alloc");
  VD_HIP(hipMemcpyAsync(saved.p, ctx->sites.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, ctx->stream),
         "save sites");
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  std::vector<float> start(3 * (size_t)n), ref(3 * (size_t)n), got(3 * (size_t)n);
  int rc = vd_get_sites(ctx, start.data());
  if (rc != VD_OK) return rc;
  auto restore = [&]() -> int {
    VD_HIP(hipMemcpyAsync(ctx->sites.p, saved.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, ctx->stream),
           "restore sites");
    VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
    ctx->have_knn = ctx->have_clip = false;
    return VD_OK;
  };
  // iters iterations under cfg from the saved sites: device ms per iteration
  // (after the first, or of the first when iters == 1) and the final sites
  auto trial = [&](const vd_config& cfg, double* ms, std::vector<float>& out) -> int {
    *ms = -1.0;
    int r = vd_ctx_set_config(ctx, &cfg);
    if (r == VD_OK) r = restore();
    if (r == VD_OK) r = vd_timing_reset(ctx);
    if (r == VD_OK) r = vd_lloyd(ctx, iters);
    if (r != VD_OK) return r;
    std::vector<vd_timing> t((size_t)iters);
    if ((r = vd_get_timing_list(ctx, t.data(), iters)) != VD_OK) return r;
    double sum = 0.0;
    for (int i = iters > 1 ? 1 : 0; i < iters; ++i) sum += t[i].total_ms;
    *ms = sum / (iters > 1 ? iters - 1 : 1);
    return vd_get_sites(ctx, out.data());
  };
  // largest displacement relative to |site|, over all sites
  auto error = [&](const std::vector<float>& a, const std::vector<float>& b) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
      double d2 = 0.0, r2 = 0.0;
      for (int x = 0; x < 3; ++x) {
        const double d = (double)a[3 * i + x] - b[3 * i + x];
        d2 += d * d;
        r2 += (double)b[3 * i + x] * b[3 * i + x];
      }
      if (d2 > 0.0) e = std::max(e, std::sqrt(d2 / std::max(r2, 1e-300)));
    }
    return e;
  };
  const vd_config base = ctx->cfg;
  int K0 = 0, TK0 = 0;
  vd_get_dims(ctx, &K0, &TK0, nullptr, nullptr);
  if (TK0 < 0) return ctx->fail(VD_ERR_ARG, "vd_autotune_k: the batch's instances need one tet_k");
  vd_config best = base;
  best.site_k = std::max(K0, 1);
  best.tet_k = std::max(TK0, 1);
  double best_ms = 0.0, best_err = 0.0;
  if ((rc = trial(best, &best_ms, ref)) != VD_OK) return rc;
  const double fr[] = {0.875, 0.75, 0.625, 0.5, 0.4, 0.3};
  for (int dim = 0; dim < 2; ++dim) {
    for (double f : fr) {
      vd_config c = best;
      if (dim == 0) {
        c.tet_k = std::max(8, (int)std::lround(f * TK0));
        if (c.tet_k >= best.tet_k) continue;
      } else {
        c.site_k = std::max(16, (int)std::lround(f * K0));
        if (c.site_k >= best.site_k) continue;
        if (c.srnr_k != 0 && c.srnr_k <= c.site_k) c.srnr_k = 0;
      }
      if (check_config(ctx, c) != VD_OK) break;
      double ms = 0.0;
      if ((rc = trial(c, &ms, got)) != VD_OK) return rc;
      const double e = error(got, ref);
      if (e > tol) break;  // smaller counts only move the sites further
      if (ms > 0.0 && ms < best_ms) {
        best_ms = ms;
        best = c;
        best_err = e;
      }
    }
  }
  if ((rc = vd_ctx_set_config(ctx, &best)) != VD_OK) return rc;
  if ((rc = restore()) != VD_OK) return rc;
  if ((rc = vd_timing_reset(ctx)) != VD_OK) return rc;
  if (chosen) *chosen = best;
  if (err_out) *err_out = best_err;
  return VD_OK;
}

int vd_sync(vd_ctx* ctx) {
  if (!ctx) return VD_ERR_ARG;
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  return VD_OK;
}

namespace {
void slot_timing(vd_ctx* ctx, int slot, bool knn, bool pairs, bool clip, bool update, vd_timing* t) {
  auto el = [&](int a, int b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[a][slot], ctx->ev[b][slot]);
    return ms;
  };
  if (ctx->graph_slot[slot]) {
    if (update) t->total_ms += el(0, 4);
    return;
  }
  if (knn) t->knn_ms += el(0, 1);
  if (knn && pairs) t->pairs_ms += el(1, 2);
  if (pairs && clip) t->clip_ms += el(2, 3);
  if (pairs && clip) t->overflow_ms += el(5, 3);
  if (clip && update) t->update_ms += el(3, 4);
  if (knn && update) t->total_ms += el(0, 4);
}
}  // namespace

int vd_get_timing(vd_ctx* ctx, vd_timing* t) {
  if (!ctx || !t) return VD_ERR_ARG;
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  std::memset(t, 0, sizeof *t);
  // the iteration in flight (no update yet) or else the last completed one
  int slot = ctx->it_slot;
  bool upd = false;
  if (!ctx->ev_knn && ctx->it_done > 0) {
    slot = (ctx->it_slot + vd_ctx::kRing - 1) % vd_ctx::kRing;
    upd = true;
    slot_timing(ctx, slot, true, true, true, true, t);
    return VD_OK;
  }
  (void)upd;
  slot_timing(ctx, slot, ctx->ev_knn, ctx->ev_pairs, ctx->ev_clip, false, t);
  return VD_OK;
}

int vd_timing_reset(vd_ctx* ctx) {
  if (!ctx) return VD_ERR_ARG;
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  ctx->it_done = 0;
  return VD_OK;
}

int vd_get_timing_sum(vd_ctx* ctx, vd_timing* t, int* n_iters) {
  if (!ctx || !t) return VD_ERR_ARG;
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  std::memset(t, 0, sizeof *t);
  const int n = (int)std::min<long long>(ctx->it_done, vd_ctx::kRing);
  for (int i = 1; i <= n; ++i) slot_timing(ctx, (ctx->it_slot + vd_ctx::kRing - i) % vd_ctx::kRing, true, true, true, true, t);
  if (n_iters) *n_iters = n;
  return VD_OK;
}

int vd_get_timing_list(vd_ctx* ctx, vd_timing* t, int n) {
  if (!ctx || n < 0 || (n && !t)) return VD_ERR_ARG;
  if (n > std::min<long long>(ctx->it_done, vd_ctx::kRing))
    return ctx->fail(VD_ERR_ARG, "vd_get_timing_list: %d iterations asked, %lld kept", n,
                     std::min<long long>(ctx->it_done, vd_ctx::kRing));
  VD_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  for (int i = 0; i < n; ++i) {
    std::memset(&t[i], 0, sizeof t[i]);
    slot_timing(ctx, (ctx->it_slot + vd_ctx::kRing - n + i) % vd_ctx::kRing, true, true, true, true, &t[i]);
  }
  return VD_OK;
}

int vd_get_site_knn(vd_ctx* ctx, int32_t* out) {
  if (!ctx || !out) return VD_ERR_ARG;
  if (!ctx->have_knn && !ctx->have_clip) return ctx->fail(VD_ERR_STATE, "no kNN result");
  if ((size_t)ctx->n_sites * ctx->K)
    VD_HIP(hipMemcpyAsync(out, ctx->site_knn.p, sizeof(int32_t) * (size_t)ctx->n_sites * ctx->K,
                          hipMemcpyDeviceToHost, ctx->stream), "read site knn");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read site knn");
  return VD_OK;
}

int vd_get_tet_knn(vd_ctx* ctx, int32_t* out) {
  if (!ctx || !out) return VD_ERR_ARG;
  if (!ctx->have_knn && !ctx->have_clip) return ctx->fail(VD_ERR_STATE, "no kNN result");
  if ((size_t)ctx->n_tets * ctx->TK)
    VD_HIP(hipMemcpyAsync(out, ctx->tet_knn.p, sizeof(int32_t) * (size_t)ctx->n_tets * ctx->TK,
                          hipMemcpyDeviceToHost, ctx->stream), "read tet knn");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read tet knn");
  return VD_OK;
}

int vd_get_moments(vd_ctx* ctx, double* out) {
  if (!ctx || !out) return VD_ERR_ARG;
  if (!ctx->have_clip && !ctx->ev_update) return ctx->fail(VD_ERR_STATE, "no clip result");
  const int n = ctx->n_sites;
  VD_HIP(ctx->moments.ensure(4 * (size_t)std::max(n, 1)), "alloc moments");
  VD_HIP(vd::moments_out(n, ctx->cfg.accum, ctx->acc_f32.p, ctx->acc_fix.p, ctx->moments.p, ctx->stream), "moments");
  VD_HIP(hipMemcpyAsync(out, ctx->moments.p, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream),
         "read moments");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read moments");
  return VD_OK;
}

int vd_get_pairs(vd_ctx* ctx, uint8_t* status, double* vals) {
  if (!ctx) return VD_ERR_ARG;
  if (!ctx->cfg.record_pairs) return ctx->fail(VD_ERR_STATE, "record_pairs is off");
  if (!ctx->have_clip && !ctx->ev_update) return ctx->fail(VD_ERR_STATE, "no clip result");
  const size_t np = (size_t)ctx->n_tets * ctx->TK;
  if (status && np)
    VD_HIP(hipMemcpyAsync(status, ctx->pair_status.p, np, hipMemcpyDeviceToHost, ctx->stream), "read pairs");
  if (vals && np)
    VD_HIP(hipMemcpyAsync(vals, ctx->pair_vals.p, 4 * sizeof(double) * np, hipMemcpyDeviceToHost, ctx->stream),
           "read pairs");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read pairs");
  return VD_OK;
}

int vd_get_pairs_range(vd_ctx* ctx, long long first, long long count, uint8_t* status, double* vals) {
  if (!ctx) return VD_ERR_ARG;
  if (!ctx->cfg.record_pairs) return ctx->fail(VD_ERR_STATE, "record_pairs is off");
  if (!ctx->have_clip && !ctx->ev_update) return ctx->fail(VD_ERR_STATE, "no clip result");
  const long long np = (long long)ctx->n_tets * ctx->TK;
  if (first < 0 || count < 0 || first > np || count > np - first)
    return ctx->fail(VD_ERR_ARG, "pair range [%lld, %lld) outside [0, %lld)", first, first + count, np);
  if (status && count)
    VD_HIP(hipMemcpyAsync(status, ctx->pair_status.p + first, (size_t)count, hipMemcpyDeviceToHost, ctx->stream),
           "read pairs");
  if (vals && count)
    VD_HIP(hipMemcpyAsync(vals, ctx->pair_vals.p + 4 * first, 4 * sizeof(double) * (size_t)count,
                          hipMemcpyDeviceToHost, ctx->stream),
           "read pairs");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read pairs");
  return VD_OK;
}

int vd_get_tet_knn_range(vd_ctx* ctx, int first_tet, int n_tets, int32_t* out) {
  if (!ctx || (!out && n_tets)) return VD_ERR_ARG;
  if (!ctx->have_knn && !ctx->have_clip) return ctx->fail(VD_ERR_STATE, "no kNN result");
  if (first_tet < 0 || n_tets < 0 || first_tet > ctx->n_tets || n_tets > ctx->n_tets - first_tet)
    return ctx->fail(VD_ERR_ARG, "tet range [%d, %d) outside [0, %d)", first_tet, first_tet + n_tets, ctx->n_tets);
  if ((size_t)n_tets * ctx->TK)
    VD_HIP(hipMemcpyAsync(out, ctx->tet_knn.p + (size_t)first_tet * ctx->TK, sizeof(int32_t) * (size_t)n_tets * ctx->TK,
                          hipMemcpyDeviceToHost, ctx->stream),
           "read tet knn");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read tet knn");
  return VD_OK;
}

int vd_get_status_hist(vd_ctx* ctx, uint64_t* hist7) {
  if (!ctx || !hist7) return VD_ERR_ARG;
  unsigned long long h[8];
  VD_HIP(hipMemcpyAsync(h, ctx->hist.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "read hist");
  VD_HIP(hipStreamSynchronize(ctx->stream), "read hist");
  for (int i = 0; i < 7; ++i) hist7[i] = h[i];
  return VD_OK;
}

}  // extern "C"

// ------------------------------------------------- batches from device jobs --

namespace {
// Tet vertex ids from instance-local to batch-global.
__global__ void k_offset_tets(int4* tets, long long n, const int* tet_inst, const int* voff) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int o = voff[tet_inst[t]];
    int4 v = tets[t];
    v.x += o;
    v.y += o;
    v.z += o;
    v.w += o;
    tets[t] = v;
  }
}
}  // namespace

namespace vdi {

hipStream_t stream(vd_ctx* ctx) { return ctx->stream; }
void set_error(vd_ctx* ctx, const char* msg) { ctx->err = msg; }

int assemble(vd_ctx* ctx, const JobDev* jobs, int n) {
  if (!ctx || !jobs || n < 1) return VD_ERR_ARG;
  VD_HIP(hipSetDevice(ctx->device), "set device");
  ctx->drop_graph();
  std::vector<vd_ctx::Inst> in((size_t)n);
  long long V = 0, T = 0, S = 0;
  for (int i = 0; i < n; ++i) {
    in[i] = {(int)V, jobs[i].nv, (int)T, jobs[i].nt, (int)S, jobs[i].ns};
    V += jobs[i].nv;
    T += jobs[i].nt;
    S += jobs[i].ns;
  }
  if (V > 0x7fffffffLL || T > 0x7fffffffLL || S > 0x7fffffffLL) return ctx->fail(VD_ERR_ARG, "batch too large");
  hipStream_t st = ctx->stream;
  VD_HIP(ctx->verts.ensure(std::max(V, 1LL)), "alloc batch");
  VD_HIP(ctx->tets.ensure(std::max(T, 1LL)), "alloc batch");
  VD_HIP(ctx->tet_xyz.ensure(3 * std::max(T, 1LL)), "alloc batch");
  VD_HIP(ctx->tet_prep.ensure((size_t)vd::kTetPrepWords * std::max(T, 1LL)), "alloc batch");
  VD_HIP(ctx->sites.ensure(std::max(S, 1LL)), "alloc batch");
  for (int i = 0; i < n; ++i) {
    const JobDev& j = jobs[i];
    const vd_ctx::Inst& q = in[i];
    if (j.nv)
      VD_HIP(hipMemcpyAsync(ctx->verts.p + q.v0, j.verts, sizeof(float4) * j.nv, hipMemcpyDeviceToDevice, st), "copy");
    if (j.nt) {
      VD_HIP(hipMemcpyAsync(ctx->tets.p + q.t0, j.tets, sizeof(int4) * j.nt, hipMemcpyDeviceToDevice, st), "copy");
      VD_HIP(hipMemcpyAsync(ctx->tet_xyz.p + 3LL * q.t0, j.tet_xyz, 3 * sizeof(float4) * j.nt,
                            hipMemcpyDeviceToDevice, st), "copy");
    }
    VD_HIP(hipMemcpyAsync(ctx->sites.p + q.s0, j.sites, sizeof(float4) * j.ns, hipMemcpyDeviceToDevice, st), "copy");
  }
  for (int a = 0; a < 3; ++a) {
    ctx->mesh_lo[a] = jobs[0].box[a];
    ctx->mesh_hi[a] = jobs[0].box[3 + a];
    for (int i = 1; i < n; ++i) {
      ctx->mesh_lo[a] = std::min(ctx->mesh_lo[a], jobs[i].box[a]);
      ctx->mesh_hi[a] = std::max(ctx->mesh_hi[a], jobs[i].box[3 + a]);
    }
  }
  // the tets' start records, from the copied corner records
  VD_HIP(vd::tet_prep(ctx->tet_xyz.p, T, ctx->tet_prep.p, st), "tet records");
  ctx->n_verts = (int)V;
  ctx->n_tets = (int)T;
  ctx->n_sites = (int)S;
  ctx->inst = in;
  ctx->rows_valid = false;
  ctx->fixed_load = 0.0;
  for (int i = 0; i < n; ++i) ctx->fixed_load = std::max(ctx->fixed_load, jobs[i].fixed_load);
  if (n > 1) {
    std::vector<int> site_begin((size_t)n + 1), voff((size_t)n);
    std::vector<float> box((size_t)6 * n);
    for (int i = 0; i < n; ++i) {
      site_begin[i] = in[i].s0;
      voff[i] = in[i].v0;
      for (int a = 0; a < 6; ++a) box[6 * i + a] = jobs[i].box[a];
    }
    site_begin[n] = (int)S;
    VD_HIP(ctx->b_site_begin.ensure(site_begin.size()), "alloc batch");
    VD_HIP(ctx->b_site_inst.ensure(std::max(S, 1LL)), "alloc batch");
    VD_HIP(ctx->b_tet_inst.ensure(std::max(T, 1LL)), "alloc batch");
    VD_HIP(ctx->b_box.ensure(box.size()), "alloc batch");
    VD_HIP(ctx->b_voff.ensure(voff.size()), "alloc batch");
    VD_HIP(hipMemcpyAsync(ctx->b_site_begin.p, site_begin.data(), sizeof(int) * site_begin.size(),
                          hipMemcpyHostToDevice, st), "upload batch");
    VD_HIP(hipMemcpyAsync(ctx->b_box.p, box.data(), sizeof(float) * box.size(), hipMemcpyHostToDevice, st),
           "upload batch");
    VD_HIP(hipMemcpyAsync(ctx->b_voff.p, voff.data(), sizeof(int) * voff.size(), hipMemcpyHostToDevice, st),
           "upload batch");
    for (int i = 0; i < n; ++i) {
      VD_HIP(vd::fill_u32(ctx->b_site_inst.p + in[i].s0, in[i].ns, (uint32_t)i, st), "batch tables");
      VD_HIP(vd::fill_u32(ctx->b_tet_inst.p + in[i].t0, in[i].nt, (uint32_t)i, st), "batch tables");
    }
    if (T > 0) {
      hipLaunchKernelGGL(k_offset_tets, dim3((unsigned)std::min<long long>((T + 255) / 256, 8192)), dim3(256), 0, st,
                         ctx->tets.p, T, ctx->b_tet_inst.p, ctx->b_voff.p);
      VD_HIP(hipGetLastError(), "offset tets");
    }
    // the host vectors above must outlive the asynchronous uploads
    VD_HIP(hipStreamSynchronize(st), "synchronize");
  }
  int rc = ensure_grid(ctx, std::max(ctx->n_sites, 1));
  if (rc != VD_OK) return rc;
  ctx->have_knn = ctx->have_clip = false;
  return VD_OK;
}

int scatter_sites(vd_ctx* ctx, const JobDev* jobs, const int* idx, int n_idx) {
  if (!ctx || (n_idx && (!jobs || !idx))) return VD_ERR_ARG;
  for (int k = 0; k < n_idx; ++k) {
    const int i = idx[k];
    if (i < 0 || i >= (int)std::max<size_t>(ctx->inst.size(), 1)) return ctx->fail(VD_ERR_ARG, "bad instance");
    const int s0 = ctx->inst.empty() ? 0 : ctx->inst[i].s0;
    VD_HIP(hipMemcpyAsync(jobs[i].sites, ctx->sites.p + s0, sizeof(float4) * jobs[i].ns, hipMemcpyDeviceToDevice,
                          ctx->stream), "copy sites");
  }
  return VD_OK;
}

}  // namespace vdi
