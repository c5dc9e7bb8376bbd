// Job server: continuous batching of independent clipped-Voronoi Lloyd jobs
// on one device.
//
// A job (mesh, sites, iteration count) waits in a FIFO until its estimated
// device footprint fits the memory budget, then its data moves to device
// buffers of its own and it joins the running set.  Every step runs one Lloyd
// iteration over all running jobs as one batch (the engine's batch of
// instances); when the running set changes the batch is re-assembled on the
// device from the jobs' buffers, the continuing jobs' sites first copied back
// to them.  A job leaves the set after its last iteration with its final
// sites read back to the host.  Because every instance of a batch gets the
// result it would get alone, a job's result does not depend on the other
// jobs or on when it was admitted.
//
// The reference has no counterpart: it runs one mesh per process
// (test_voronoi.cu:150-229).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "vd_api.h"
#include "vd_clip.hpp"
#include "vd_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
  int id = 0;
  int state = VD_JOB_PENDING;
  int n_iter = 0, done_iters = 0;
  int nv = 0, nt = 0, ns = 0;
  std::vector<float> hv, hs;  // host copies until admission
  std::vector<int32_t> ht;
  std::vector<float> result;  // final sites (xyz) once done
  float box[6] = {};
  size_t bytes = 0;
  void* block = nullptr;  // device buffers: verts, tets, tet records, sites (from the arena)
  size_t block_bytes = 0;
  vdi::JobDev dev{};
  Clock::time_point t_submit, t_admit, t_done;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device memory of the jobs' own buffers: chunks taken from the runtime as the
// running set first needs them and kept until the server ends, handed out by
// first fit with neighbouring free ranges merged on release, so that admitting
// and retiring jobs allocates and frees nothing on the device (a hipFree
// waits for the whole device).
class Arena {
 public:
  static constexpr size_t kChunk = (size_t)1 << 30;

  ~Arena() {
    for (auto& c : chunks_) (void)hipFree(c.base);
  }

  // A 256-byte-aligned range of `bytes`, or nullptr when the device has no
  // memory for a new chunk.
  void* take(size_t bytes) {
    bytes = align256(std::max<size_t>(bytes, 1));
    for (auto& c : chunks_) {
      for (auto it = c.free.begin(); it != c.free.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        c.free.erase(it);
        if (len > bytes) c.free.emplace(off + bytes, len - bytes);
        return c.base + off;
      }
    }
    Chunk c;
    c.size = std::max(kChunk, bytes);
    if (hipMalloc((void**)&c.base, c.size) != hipSuccess) return nullptr;
    if (c.size > bytes) c.free.emplace(bytes, c.size - bytes);
    chunks_.push_back(std::move(c));
    return chunks_.back().base;
  }

  void give(void* p, size_t bytes) {
    if (!p) return;
    bytes = align256(std::max<size_t>(bytes, 1));
    for (auto& c : chunks_) {
      char* q = (char*)p;
      if (q < c.base || q >= c.base + c.size) continue;
      size_t off = (size_t)(q - c.base), len = bytes;
      auto next = c.free.lower_bound(off);
      if (next != c.free.end() && next->first == off + len) {
        len += next->second;
        next = c.free.erase(next);
      }
      if (next != c.free.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
          off = prev->first;
          len += prev->second;
          c.free.erase(prev);
        }
      }
      c.free.emplace(off, len);
      return;
    }
  }

  size_t reserved() const {
    size_t n = 0;
    for (const auto& c : chunks_) n += c.size;
    return n;
  }

 private:
  struct Chunk {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length
  };
  std::vector<Chunk> chunks_;
};

}  // namespace

struct vd_server {
  vd_ctx* ctx = nullptr;
  Arena arena;
  int device = 0;
  vd_config cfg{};
  size_t budget = 0, used = 0;
  std::deque<int> pending;
  std::vector<int> running;    // admission order
  std::vector<int> assembled;  // job ids in the order of the current batch
  bool changed = false;
  std::unordered_map<int, Job> jobs;
  int next_id = 0;
  std::string err;

  int fail(int code, const char* msg) {
    err = msg;
    return code;
  }

  // Device bytes of a running job: its own buffers plus its share of the
  // batch's working memory (batch copies, the tets' start records, kNN
  // tables, grouped pairs, retry list, grid, the site-major form's
  // volume-pass records: one per ten pairs, and the cell form's cells).
  size_t footprint(int nv, int nt, int ns) const {
    const size_t K = (size_t)cfg.site_k, TK = (size_t)cfg.tet_k;
    const size_t own = 16 * (size_t)nv + (16 + 48) * (size_t)nt + 16 * (size_t)ns + 4 * 256;
    size_t rec = 0, cells = 0;
    if (cfg.form == VD_FORM_SITE_MAJOR && cfg.site_records >= 0) {
      const int ft = vdi::site_tcap_of(cfg);
      rec = (16 + 4 * (size_t)vd::rec_stride(ft)) * TK / 10;
    }
    // status, counts, radius, planes, triangles and the rebuild list per site
    if (cfg.form == VD_FORM_CELLS) cells = 20 + 16 * (size_t)cfg.pcap + 4 * (size_t)cfg.tcap;
    const size_t work =
        16 * (size_t)nv + (16 + 48 + 16 * vd::kTetPrepWords + 12 * TK + rec + 8) * (size_t)nt +
        (16 + 4 * K + 64 + cells) * (size_t)ns;
    return own + work;
  }

  int admit(Job& j) {
    const size_t bv = align256(16 * (size_t)std::max(j.nv, 1)), bt = align256(16 * (size_t)std::max(j.nt, 1)),
                 bx = align256(48 * (size_t)std::max(j.nt, 1)), bs = align256(16 * (size_t)j.ns);
    j.block_bytes = bv + bt + bx + bs;
    if (hipSetDevice(device) != hipSuccess || !(j.block = arena.take(j.block_bytes))) {
      j.block = nullptr;
      return fail(VD_ERR_NOMEM, "job buffers: device allocation failed");
    }
    char* b = (char*)j.block;
    std::vector<float4> v4((size_t)std::max(j.nv, 1)), s4((size_t)j.ns), xyz((size_t)3 * std::max(j.nt, 1));
    for (int i = 0; i < j.nv; ++i) v4[i] = make_float4(j.hv[3 * i], j.hv[3 * i + 1], j.hv[3 * i + 2], 0.f);
    for (int i = 0; i < j.ns; ++i) s4[i] = make_float4(j.hs[3 * i], j.hs[3 * i + 1], j.hs[3 * i + 2], 0.f);
    float* r = (float*)xyz.data();
    for (size_t t = 0; t < (size_t)j.nt; ++t)
      for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) r[12 * t + 3 * k + a] = j.hv[3 * (size_t)j.ht[4 * t + k] + a];
    bool ok = hipMemcpy(b, v4.data(), 16 * (size_t)j.nv, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b + bv, j.ht.data(), 16 * (size_t)j.nt, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b + bv + bt, xyz.data(), 48 * (size_t)j.nt, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b + bv + bt + bx, s4.data(), 16 * (size_t)j.ns, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
      arena.give(j.block, j.block_bytes);
      j.block = nullptr;
      return fail(VD_ERR_HIP, "job upload failed");
    }
    j.dev.verts = (const float4*)b;
    j.dev.nv = j.nv;
    j.dev.tets = (const int4*)(b + bv);
    j.dev.tet_xyz = (const float4*)(b + bv + bt);
    j.dev.nt = j.nt;
    j.dev.sites = (float4*)(b + bv + bt + bx);
    j.dev.ns = j.ns;
    std::memcpy(j.dev.box, j.box, sizeof j.box);
    j.hv.clear();
    j.hv.shrink_to_fit();
    j.ht.clear();
    j.ht.shrink_to_fit();
    j.hs.clear();
    j.hs.shrink_to_fit();
    j.state = VD_JOB_RUNNING;
    j.t_admit = Clock::now();
    used += j.bytes;
    return VD_OK;
  }

  std::vector<vdi::JobDev> devs(const std::vector<int>& ids) {
    std::vector<vdi::JobDev> d;
    d.reserve(ids.size());
    for (int id : ids) d.push_back(jobs[id].dev);
    return d;
  }
};

extern "C" {

int vd_server_create(vd_server** out, int device, const vd_config* cfg, size_t hbm_budget_bytes) {
  if (!out) return VD_ERR_ARG;
  *out = nullptr;
  vd_config c;
  if (cfg)
    c = *cfg;
  else
    vd_config_default(&c);
  if (c.tet_k <= 0) {
    fprintf(stderr, "vd_server_create: jobs share one batch, so tet_k must be set (> 0)\n");
    return VD_ERR_ARG;
  }
  vd_server* s = new vd_server();
  s->device = device;
  s->cfg = c;
  s->budget = hbm_budget_bytes;
  const int rc = vd_ctx_create(&s->ctx, device, &c);
  if (rc != VD_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return VD_OK;
}

void vd_server_destroy(vd_server* s) {
  if (!s) return;
  if (s->ctx) (void)vd_sync(s->ctx);
  vd_ctx_destroy(s->ctx);
  delete s;  // the arena returns its chunks
}

const char* vd_server_error(const vd_server* s) { return s ? s->err.c_str() : "null server"; }

int vd_server_submit(vd_server* s, const float* verts_xyz, int n_verts, const int32_t* tets, int n_tets,
                     const float* sites_xyz, int n_sites, int n_iter, int* job_id) {
  if (!s || n_verts < 0 || n_tets < 0 || n_sites < 1 || n_iter < 1 || (n_verts && !verts_xyz) ||
      (n_tets && !tets) || !sites_xyz || !job_id)
    return s ? s->fail(VD_ERR_ARG, "vd_server_submit: bad arguments") : VD_ERR_ARG;
  if (n_sites <= s->cfg.site_k || n_sites < s->cfg.tet_k)
    return s->fail(VD_ERR_PRECONDITION, "vd_server_submit: a job needs more sites than site_k and tet_k");
  for (long long i = 0; i < 4LL * n_tets; ++i)
    if (tets[i] < 0 || tets[i] >= n_verts) return s->fail(VD_ERR_PRECONDITION, "vd_server_submit: bad tet vertex id");
  Job j;
  j.id = s->next_id++;
  j.n_iter = n_iter;
  j.nv = n_verts;
  j.nt = n_tets;
  j.ns = n_sites;
  j.hv.assign(verts_xyz, verts_xyz + 3 * (size_t)n_verts);
  j.ht.assign(tets, tets + 4 * (size_t)n_tets);
  j.hs.assign(sites_xyz, sites_xyz + 3 * (size_t)n_sites);
  for (int a = 0; a < 3; ++a) {
    float lo = n_verts ? verts_xyz[a] : 0.f, hi = lo;
    for (int v = 0; v < n_verts; ++v) {
      lo = std::min(lo, verts_xyz[3 * (size_t)v + a]);
      hi = std::max(hi, verts_xyz[3 * (size_t)v + a]);
    }
    j.box[a] = lo;
    j.box[3 + a] = hi;
  }
  j.dev.fixed_load = vdi::fixed_point_load(verts_xyz, 0, n_verts, tets, 0, n_tets);
  if (s->cfg.accum == VD_ACCUM_F64 && j.dev.fixed_load > vdi::kFixedPointMaxLoad)
    return s->fail(VD_ERR_PRECONDITION,
                   "vd_server_submit: mesh volume x largest |coordinate| exceeds the range of the exact FP64 "
                   "fixed-point sums; map the mesh into the [0,1000]^3 frame");
  j.bytes = s->footprint(n_verts, n_tets, n_sites);
  if (j.bytes > s->budget) return s->fail(VD_ERR_NOMEM, "vd_server_submit: the job alone exceeds the memory budget");
  j.t_submit = Clock::now();
  *job_id = j.id;
  s->pending.push_back(j.id);
  s->jobs.emplace(j.id, std::move(j));
  return VD_OK;
}

int vd_server_step(vd_server* s, int* n_running, int* n_pending) {
  if (!s) return VD_ERR_ARG;
  // admission in arrival order while the budget allows
  while (!s->pending.empty()) {
    Job& j = s->jobs[s->pending.front()];
    if (s->used + j.bytes > s->budget) break;
    const int rc = s->admit(j);
    if (rc != VD_OK) return rc;
    s->running.push_back(j.id);
    s->pending.pop_front();
    s->changed = true;
  }
  if (!s->running.empty()) {
    if (s->changed) {
      // the continuing jobs' current sites live in the batch: keep them
      std::vector<int> keep_idx;
      for (size_t i = 0; i < s->assembled.size(); ++i)
        if (s->jobs[s->assembled[i]].state == VD_JOB_RUNNING) keep_idx.push_back((int)i);
      if (!keep_idx.empty()) {
        const std::vector<vdi::JobDev> prev = s->devs(s->assembled);
        int rc = vdi::scatter_sites(s->ctx, prev.data(), keep_idx.data(), (int)keep_idx.size());
        if (rc != VD_OK) return s->fail(rc, vd_ctx_error(s->ctx));
      }
      const std::vector<vdi::JobDev> cur = s->devs(s->running);
      int rc = vdi::assemble(s->ctx, cur.data(), (int)cur.size());
      if (rc != VD_OK) return s->fail(rc, vd_ctx_error(s->ctx));
      s->assembled = s->running;
      s->changed = false;
    }
    int rc = vd_lloyd(s->ctx, 1);
    if (rc != VD_OK) return s->fail(rc, vd_ctx_error(s->ctx));
    std::vector<int> fin_idx;
    for (size_t i = 0; i < s->assembled.size(); ++i) {
      Job& j = s->jobs[s->assembled[i]];
      if (++j.done_iters == j.n_iter) fin_idx.push_back((int)i);
    }
    if (!fin_idx.empty()) {
      const std::vector<vdi::JobDev> cur = s->devs(s->assembled);
      rc = vdi::scatter_sites(s->ctx, cur.data(), fin_idx.data(), (int)fin_idx.size());
      if (rc != VD_OK) return s->fail(rc, vd_ctx_error(s->ctx));
      // the host waits for the device only when a job's result is due; other
      // steps return once their iteration is enqueued
      rc = vd_sync(s->ctx);
      if (rc != VD_OK) return s->fail(rc, vd_ctx_error(s->ctx));
    }
    const Clock::time_point now = Clock::now();
    for (int i : fin_idx) {
      Job& j = s->jobs[s->assembled[i]];
      std::vector<float4> h((size_t)j.ns);
      if (hipMemcpy(h.data(), j.dev.sites, 16 * (size_t)j.ns, hipMemcpyDeviceToHost) != hipSuccess)
        return s->fail(VD_ERR_HIP, "job readback failed");
      j.result.resize(3 * (size_t)j.ns);
      for (int k = 0; k < j.ns; ++k) {
        j.result[3 * k] = h[k].x;
        j.result[3 * k + 1] = h[k].y;
        j.result[3 * k + 2] = h[k].z;
      }
      s->arena.give(j.block, j.block_bytes);
      j.block = nullptr;
      s->used -= j.bytes;
      j.state = VD_JOB_DONE;
      j.t_done = now;
      s->running.erase(std::find(s->running.begin(), s->running.end(), j.id));
      s->changed = true;
    }
  }
  if (n_running) *n_running = (int)s->running.size();
  if (n_pending) *n_pending = (int)s->pending.size();
  return VD_OK;
}

int vd_server_job(vd_server* s, int job_id, int* state, int* iters_done, double* wait_ms, double* run_ms) {
  if (!s) return VD_ERR_ARG;
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return s->fail(VD_ERR_ARG, "unknown job");
  const Job& j = it->second;
  if (state) *state = j.state;
  if (iters_done) *iters_done = j.done_iters;
  const Clock::time_point now = Clock::now();
  if (wait_ms) *wait_ms = ms_between(j.t_submit, j.state == VD_JOB_PENDING ? now : j.t_admit);
  if (run_ms) *run_ms = j.state == VD_JOB_PENDING ? 0.0 : ms_between(j.t_admit, j.state == VD_JOB_DONE ? j.t_done : now);
  return VD_OK;
}

int vd_server_result(vd_server* s, int job_id, float* sites_xyz) {
  if (!s || !sites_xyz) return VD_ERR_ARG;
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return s->fail(VD_ERR_ARG, "unknown job");
  if (it->second.state != VD_JOB_DONE) return s->fail(VD_ERR_STATE, "job not done");
  std::memcpy(sites_xyz, it->second.result.data(), sizeof(float) * it->second.result.size());
  return VD_OK;
}

size_t vd_server_bytes_in_use(const vd_server* s) { return s ? s->used : 0; }

}  // extern "C"
