// C ABI entry points that stand in for the reference's own host functions:
//   vd_knn_dev                       <- knn_cuda_global_dev (knncuda.cu:271-349)
//   vd_knn_dev_noalloc               the same kNN on a static device scratch array
//   vd_kn_prepare / solve / free ... <- kn_prepare, kn_solve, kn_free, kn_get_*
//                                       (knearests.cu:230-432)
//   vd_compute_clipped_voro_diagram  <- compute_clipped_voro_diagram_GPU
//                                       (voronoi.cu:925-1188)
// All of them run on the engine's kernels; none exits the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "vd_api.h"
#include "vd_knn.hpp"

namespace {

__global__ void k_soa_to_f4(const float* src, size_t pitch, int n, float4* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = make_float4(src[i], src[i + pitch], src[i + 2 * pitch], 0.f);
}
__global__ void k_f4_to_xyz(const float4* src, int n, float* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = src[i];
  dst[3 * i] = p.x;
  dst[3 * i + 1] = p.y;
  dst[3 * i + 2] = p.z;
}
__global__ void k_inverse(const int* sorted_ids, int n, int* inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[sorted_ids[i]] = i;
}
// Neighbour lists by original index -> rows by grid position, entries as grid
// positions (the reference's kn_solve output convention, knearests.cu:138-144).
__global__ void k_to_sorted(const int* sorted_ids, const int* inv, const int32_t* by_orig, int n, int k,
                            unsigned int* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * k) return;
  const int pos = (int)(i / k), j = (int)(i % k);
  const int32_t nb = by_orig[(long long)sorted_ids[pos] * k + j];
  out[i] = nb < 0 ? 0xffffffffu : (unsigned int)inv[nb];
}

template <class T>
struct Tmp {
  T* p = nullptr;
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

struct GridTmp {
  Tmp<vd::GridShape> shape;
  Tmp<int> count, start, cursor, cell_of, ids, scan_tmp;
  Tmp<float4> sorted;
  Tmp<int4> offsets;
  int n_offsets = 0;
  int max_cells = 0;
  hipError_t alloc(int n, float density) {
    max_cells = vd::grid_max_cells(n, density);
    hipError_t e;
    if ((e = shape.alloc(1)) || (e = count.alloc(max_cells)) || (e = start.alloc((size_t)max_cells + 1)) ||
        (e = cursor.alloc(max_cells)) || (e = cell_of.alloc(n)) || (e = ids.alloc(n)) || (e = sorted.alloc(n)) ||
        (e = scan_tmp.alloc(vd::scan_tmp_ints(max_cells))))
      return e;
    const std::vector<int4> t = vd::knn_offset_table();
    if ((e = offsets.alloc(t.size()))) return e;
    n_offsets = (int)t.size();
    return hipMemcpy(offsets.p, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice);
  }
  vd::GridBufs bufs() const {
    vd::GridBufs g;
    g.shape = shape.p;
    g.max_cells = max_cells;
    g.count = count.p;
    g.start = start.p;
    g.cursor = cursor.p;
    g.cell_of = cell_of.p;
    g.sorted_ids = ids.p;
    g.sorted_pts = sorted.p;
    g.scan_tmp = scan_tmp.p;
    return g;
  }
};

// Arguments of vd_knn_dev and vd_knn_dev_noalloc (knncuda.cu:271-349: pitched
// SoA rows, dim 3; the index pitch must hold every query).
bool knn_dev_args_ok(const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
                     size_t query_pitch, int dim, int k, const int* knn_index_dev, size_t index_pitch) {
  return !(dim != 3 || k < 1 || k > 256 || ref_nb < 1 || query_nb < 0 || !ref_dev || !knn_index_dev ||
           ref_pitch < (size_t)ref_nb || (query_nb && (!query_dev || query_pitch < (size_t)query_nb)) ||
           index_pitch < (size_t)query_nb);
}

// The kNN of vd_knn_dev on given device buffers: points to float4, grid over
// the reference points, one query per query point.  Enqueues only.
hipError_t knn_dev_enqueue(const vd::GridBufs& gb, const int4* offsets, int n_offsets, float4* ref4, float4* q4,
                           const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
                           size_t query_pitch, int k, int* knn_index_dev, size_t index_pitch, hipStream_t st) {
  hipLaunchKernelGGL(k_soa_to_f4, dim3((ref_nb + 255) / 256), dim3(256), 0, st, ref_dev, ref_pitch, ref_nb, ref4);
  if (query_nb)
    hipLaunchKernelGGL(k_soa_to_f4, dim3((query_nb + 255) / 256), dim3(256), 0, st, query_dev, query_pitch, query_nb,
                       q4);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = vd::grid_build(gb, ref4, ref_nb, nullptr, nullptr, 3.1f, st);
  if (e == hipSuccess)
    e = vd::knn_query(gb, ref_nb, offsets, n_offsets, k, 2, query_nb, nullptr, nullptr, q4, nullptr, knn_index_dev,
                      (long long)index_pitch, st);
  return e;
}

// Device scratch of vd_knn_dev_noalloc, which allocates nothing: one array per
// device, held until the process ends.  Host code never uses the array's name
// as an address: each device's address comes from hipGetSymbolAddress at the
// first call on that device (noalloc_slot).  Layout (scratch_layout): the cell
// offset table, the grid shape, the grid's per-cell and per-point tables for
// up to VD_KNN_NOALLOC_MAX_REF reference points, the reference points and the
// queries as float4.
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchLayout {
  vd::GridBufs grid;
  int4* offsets = nullptr;
  float4* ref4 = nullptr;
  float4* q4 = nullptr;
  size_t bytes = 0;
};

// Carves the scratch at `base` (a device address); the grid's cell count is
// that of n_ref points, its buffers are sized for the maximum.
ScratchLayout scratch_layout(unsigned char* base, int n_ref) {
  ScratchLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const int max_ref = VD_KNN_NOALLOC_MAX_REF;
  const int max_cells = vd::grid_max_cells(max_ref, 3.1f);
  const int n_offsets = (2 * vd::kOffsetRadius + 1) * (2 * vd::kOffsetRadius + 1) * (2 * vd::kOffsetRadius + 1);
  L.offsets = (int4*)take(sizeof(int4) * n_offsets);
  L.grid.shape = (vd::GridShape*)take(sizeof(vd::GridShape));
  L.grid.count = (int*)take(sizeof(int) * max_cells);
  L.grid.start = (int*)take(sizeof(int) * ((size_t)max_cells + 1));
  L.grid.cursor = (int*)take(sizeof(int) * max_cells);
  L.grid.scan_tmp = (int*)take(sizeof(int) * vd::scan_tmp_ints(max_cells));
  L.grid.cell_of = (int*)take(sizeof(int) * max_ref);
  L.grid.sorted_ids = (int*)take(sizeof(int) * max_ref);
  L.grid.sorted_pts = (float4*)take(sizeof(float4) * max_ref);
  L.ref4 = (float4*)take(sizeof(float4) * max_ref);
  L.q4 = (float4*)take(sizeof(float4) * (size_t)VD_KNN_NOALLOC_MAX_QUERY);
  L.grid.max_cells = vd::grid_max_cells(std::min(n_ref, max_ref), 3.1f);
  L.bytes = off;
  return L;
}

struct NoAllocSlot {
  unsigned char* ptr = nullptr;  // this device's address of `scratch`
  int n_offsets = 0;
};
std::mutex g_noalloc_mu;
std::map<int, NoAllocSlot> g_noalloc;

// Set-up of the current device `dev` at its first call: the scratch's
// address, and the cell offset table uploaded into it (a synchronous copy of
// a constant table; later calls only enqueue kernels).
int noalloc_slot(int dev, const NoAllocSlot** out) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  auto it = g_noalloc.find(dev);
  if (it != g_noalloc.end()) {
    *out = &it->second;
    return VD_OK;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "vd_knn_dev_noalloc: no address for the scratch on device %d: %s\n", dev, hipGetErrorString(e));
    return VD_ERR_HIP;
  }
  NoAllocSlot slot;
  slot.ptr = (unsigned char*)p;
  const ScratchLayout L = scratch_layout(slot.ptr, VD_KNN_NOALLOC_MAX_REF);
  const std::vector<int4> t = vd::knn_offset_table();
  if (L.bytes > kScratchBytes || (size_t)((unsigned char*)L.grid.shape - (unsigned char*)L.offsets) <
                                     sizeof(int4) * t.size()) {
    fprintf(stderr, "vd_knn_dev_noalloc: scratch layout of %zu bytes does not fit %zu\n", L.bytes, kScratchBytes);
    return VD_ERR_HIP;
  }
  e = hipMemcpy(L.offsets, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fprintf(stderr, "vd_knn_dev_noalloc: cannot set up device %d: %s\n", dev, hipGetErrorString(e));
    return VD_ERR_HIP;
  }
  slot.n_offsets = (int)t.size();
  auto r = g_noalloc.emplace(dev, slot);
  *out = &r.first->second;
  return VD_OK;
}

}  // namespace

struct vd_kn {
  int n = 0, k = 0;
  hipStream_t stream = nullptr;
  GridTmp grid;
  Tmp<float4> pts;
  Tmp<float> stored_xyz;
  Tmp<int32_t> by_orig;
  Tmp<int> inv;
  Tmp<unsigned int> knearests;
  vd::GridShape shape{};
};

extern "C" {

int vd_knn_dev(const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
               size_t query_pitch, int dim, int k, int* knn_index_dev, size_t index_pitch, void* stream) {
  if (!knn_dev_args_ok(ref_dev, ref_nb, ref_pitch, query_dev, query_nb, query_pitch, dim, k, knn_index_dev,
                       index_pitch))
    return VD_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  GridTmp g;
  Tmp<float4> ref4, q4;
  hipError_t e;
  if ((e = g.alloc(ref_nb, 3.1f)) || (e = ref4.alloc(ref_nb)) || (e = q4.alloc(query_nb)))
    return e == hipErrorOutOfMemory ? VD_ERR_NOMEM : VD_ERR_HIP;
  if (knn_dev_enqueue(g.bufs(), g.offsets.p, g.n_offsets, ref4.p, q4.p, ref_dev, ref_nb, ref_pitch, query_dev,
                      query_nb, query_pitch, k, knn_index_dev, index_pitch, st))
    return VD_ERR_HIP;
  if ((e = hipStreamSynchronize(st))) return VD_ERR_HIP;
  return VD_OK;
}

int vd_knn_dev_noalloc(const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
                       size_t query_pitch, int dim, int k, int* knn_index_dev, size_t index_pitch, void* stream) {
  if (!knn_dev_args_ok(ref_dev, ref_nb, ref_pitch, query_dev, query_nb, query_pitch, dim, k, knn_index_dev,
                       index_pitch))
    return VD_ERR_ARG;
  if (ref_nb > VD_KNN_NOALLOC_MAX_REF || query_nb > VD_KNN_NOALLOC_MAX_QUERY) {
    fprintf(stderr,
            "vd_knn_dev_noalloc: %d reference points and %d queries exceed its scratch limits "
            "(%d and %d); use vd_knn_dev\n",
            ref_nb, query_nb, VD_KNN_NOALLOC_MAX_REF, VD_KNN_NOALLOC_MAX_QUERY);
    return VD_ERR_ARG;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return VD_ERR_HIP;
  const NoAllocSlot* slot = nullptr;
  const int rc = noalloc_slot(dev, &slot);
  if (rc != VD_OK) return rc;
  const ScratchLayout L = scratch_layout(slot->ptr, ref_nb);
  if (knn_dev_enqueue(L.grid, L.offsets, slot->n_offsets, L.ref4, L.q4, ref_dev, ref_nb, ref_pitch, query_dev,
                      query_nb, query_pitch, k, knn_index_dev, index_pitch, (hipStream_t)stream))
    return VD_ERR_HIP;
  return VD_OK;
}

int vd_kn_prepare(vd_kn** out, const float* points_xyz, int n, int k) {
  if (!out || n < 2 || !points_xyz || k < 1 || k > 256) return VD_ERR_ARG;
  *out = nullptr;
  vd_kn* kn = new vd_kn();
  kn->n = n;
  kn->k = k;
  hipError_t e = hipStreamCreateWithFlags(&kn->stream, hipStreamNonBlocking);
  std::vector<float4> h(n);
  for (int i = 0; i < n; ++i) h[i] = make_float4(points_xyz[3 * i], points_xyz[3 * i + 1], points_xyz[3 * i + 2], 0.f);
  if (!e) e = kn->grid.alloc(n, 3.1f);
  if (!e) e = kn->pts.alloc(n);
  if (!e) e = kn->stored_xyz.alloc(3 * (size_t)n);
  if (!e) e = kn->by_orig.alloc((size_t)n * k);
  if (!e) e = kn->inv.alloc(n);
  if (!e) e = kn->knearests.alloc((size_t)n * k);
  if (!e) e = hipMemcpyAsync(kn->pts.p, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, kn->stream);
  const vd::GridBufs gb = kn->grid.bufs();
  if (!e) e = vd::grid_build(gb, kn->pts.p, n, nullptr, nullptr, 3.1f, kn->stream);
  if (!e) {
    hipLaunchKernelGGL(k_f4_to_xyz, dim3((n + 255) / 256), dim3(256), 0, kn->stream, gb.sorted_pts, n,
                       kn->stored_xyz.p);
    hipLaunchKernelGGL(k_inverse, dim3((n + 255) / 256), dim3(256), 0, kn->stream, gb.sorted_ids, n, kn->inv.p);
    e = hipGetLastError();
  }
  if (!e) e = hipMemcpyAsync(&kn->shape, gb.shape, sizeof kn->shape, hipMemcpyDeviceToHost, kn->stream);
  if (!e) e = hipStreamSynchronize(kn->stream);
  if (e) {
    vd_kn_free(kn);
    return e == hipErrorOutOfMemory ? VD_ERR_NOMEM : VD_ERR_HIP;
  }
  *out = kn;
  return VD_OK;
}

int vd_kn_solve(vd_kn* kn) {
  if (!kn) return VD_ERR_ARG;
  const vd::GridBufs gb = kn->grid.bufs();
  const int k = std::min(kn->k, kn->n - 1);
  hipError_t e = hipSuccess;
  if (k < kn->k) e = vd::fill_u32(kn->by_orig.p, (long long)kn->n * kn->k, 0xffffffffu, kn->stream);
  // neighbour lists of width kn->k; entries past n-1 stay -1
  if (!e)
    e = vd::knn_query(gb, kn->n, kn->grid.offsets.p, kn->grid.n_offsets, kn->k, 0, kn->n, nullptr, nullptr, nullptr,
                      nullptr, kn->by_orig.p, 0, kn->stream);
  if (!e) {
    const long long tot = (long long)kn->n * kn->k;
    hipLaunchKernelGGL(k_to_sorted, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, kn->stream, gb.sorted_ids,
                       kn->inv.p, kn->by_orig.p, kn->n, kn->k, kn->knearests.p);
    e = hipGetLastError();
  }
  if (!e) e = hipStreamSynchronize(kn->stream);
  return e ? VD_ERR_HIP : VD_OK;
}

void vd_kn_free(vd_kn* kn) {
  if (!kn) return;
  if (kn->stream) {
    (void)hipStreamSynchronize(kn->stream);
    (void)hipStreamDestroy(kn->stream);
  }
  delete kn;
}

int vd_kn_info_get(const vd_kn* kn, vd_kn_info* info) {
  if (!kn || !info) return VD_ERR_ARG;
  info->allocated_points = kn->n;
  info->dimx = kn->shape.dim[0];
  info->dimy = kn->shape.dim[1];
  info->dimz = kn->shape.dim[2];
  info->num_cell_offsets = kn->grid.n_offsets;
  info->d_stored_points = kn->stored_xyz.p;
  info->d_knearests = kn->knearests.p;
  info->d_permutation = (unsigned int*)kn->grid.ids.p;
  info->d_counters = kn->grid.count.p;
  info->d_ptrs = kn->grid.start.p;
  return VD_OK;
}

int vd_kn_get_points(vd_kn* kn, float* out_xyz) {
  if (!kn || !out_xyz) return VD_ERR_ARG;
  return hipMemcpy(out_xyz, kn->stored_xyz.p, sizeof(float) * 3 * (size_t)kn->n, hipMemcpyDeviceToHost) ? VD_ERR_HIP
                                                                                                     : VD_OK;
}

int vd_kn_get_knearests(vd_kn* kn, unsigned int* out) {
  if (!kn || !out) return VD_ERR_ARG;
  return hipMemcpy(out, kn->knearests.p, sizeof(unsigned int) * (size_t)kn->n * kn->k, hipMemcpyDeviceToHost)
             ? VD_ERR_HIP
             : VD_OK;
}

int vd_kn_get_cell_counts(vd_kn* kn, int* out) {
  if (!kn || !out) return VD_ERR_ARG;
  const size_t cells = (size_t)kn->shape.dim[0] * kn->shape.dim[1] * kn->shape.dim[2];
  return hipMemcpy(out, kn->grid.count.p, sizeof(int) * cells, hipMemcpyDeviceToHost) ? VD_ERR_HIP : VD_OK;
}

int vd_kn_get_permutation(vd_kn* kn, unsigned int* out) {
  if (!kn || !out) return VD_ERR_ARG;
  return hipMemcpy(out, kn->grid.ids.p, sizeof(unsigned int) * (size_t)kn->n, hipMemcpyDeviceToHost) ? VD_ERR_HIP
                                                                                                  : VD_OK;
}

int vd_compute_clipped_voro_diagram(const float* vertices, int n_verts, const int32_t* indices, int n_tets,
                                    float* site, int n_site, int site_is_transposed, int nb_lloyd_iter,
                                    int preferred_tet_k, const vd_config* cfg_in, const char* out_dir) {
  if (!vertices || !indices || !site || n_site < 1 || nb_lloyd_iter < 0) return VD_ERR_ARG;
  vd_config cfg;
  if (cfg_in)
    cfg = *cfg_in;
  else
    vd_config_default(&cfg);
  cfg.tet_k = preferred_tet_k;
  vd_ctx* ctx = nullptr;
  int rc = vd_ctx_create(&ctx, 0, &cfg);
  if (rc != VD_OK) return rc;
  std::vector<float> xyz(3 * (size_t)n_site);
  for (int i = 0; i < n_site; ++i)
    for (int a = 0; a < 3; ++a)
      xyz[3 * i + a] = site_is_transposed ? site[i + (size_t)a * n_site] : site[3 * i + a];
  rc = vd_set_mesh(ctx, vertices, n_verts, indices, n_tets);
  if (rc == VD_OK) rc = vd_set_sites(ctx, xyz.data(), n_site);
  const std::string dir = (out_dir && *out_dir) ? std::string(out_dir) + "/" : std::string();
  FILE* rec = rc == VD_OK ? fopen((dir + "record.csv").c_str(), "a") : nullptr;
  if (rc == VD_OK && !rec) rc = VD_ERR_ARG;
  if (rec) fprintf(rec, "n_site, n_tet, TET_K, SITE_K, Iter, k-NN, Clip_tet-cell, Clip_cell-tet, Compute_new_site, GPU2CPU\n");
  for (int it = 0; rc == VD_OK && it < nb_lloyd_iter; ++it) {
    printf("--------------------Iter %d--------------------\n", it);
    const auto t0 = std::chrono::steady_clock::now();
    rc = vd_knn_step(ctx);
    if (rc == VD_OK) rc = vd_clip_step(ctx);
    if (rc == VD_OK) rc = vd_update_step(ctx);
    vd_timing tm{};
    if (rc == VD_OK) rc = vd_get_timing(ctx, &tm);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc == VD_OK) rc = vd_get_sites(ctx, xyz.data());
    const auto t2 = std::chrono::steady_clock::now();
    if (rc != VD_OK) break;
    int K = 0, TK = 0;
    vd_get_dims(ctx, &K, &TK, nullptr, nullptr);
    // columns as in voronoi.cu:985-1175, in seconds.  The reference runs both
    // formulations every iteration and times kernel A under Clip_tet-cell and
    // kernels B+C under Clip_cell-tet (voronoi.cu:1034-1116); this engine runs
    // the configured one, whose time goes under its column, and 0 under the other
    const double clip_s = (tm.pairs_ms + tm.clip_ms) / 1e3;
    const bool cells = cfg.form == VD_FORM_CELLS;
    fprintf(rec, "%d, %d, %d, %d, %d, %.5f, %.5f, %.5f, %.5f, %.5f, \n", n_site, n_tets, TK, cfg.site_k, it,
            (tm.knn_ms) / 1e3, cells ? 0.0 : clip_s, cells ? clip_s : 0.0, tm.update_ms / 1e3,
            std::chrono::duration<double>(t2 - t1).count());
    printf("Iteration: %gs\n", std::chrono::duration<double>(t2 - t0).count());
    FILE* f = fopen((dir + "iter_" + std::to_string(it) + ".xyz").c_str(), "w");
    if (f) {
      fprintf(f, "%d\n", n_site);
      for (int i = 0; i < n_site; ++i) fprintf(f, "%g  %g  %g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      fclose(f);
    }
  }
  if (rec) fclose(rec);
  if (rc == VD_OK)
    for (int i = 0; i < n_site; ++i)
      for (int a = 0; a < 3; ++a) {
        if (site_is_transposed)
          site[i + (size_t)a * n_site] = xyz[3 * i + a];
        else
          site[3 * i + a] = xyz[3 * i + a];
      }
  if (rc != VD_OK) fprintf(stderr, "vd_compute_clipped_voro_diagram: %s\n", vd_ctx_error(ctx));
  vd_ctx_destroy(ctx);
  return rc;
}

}  // extern "C"
