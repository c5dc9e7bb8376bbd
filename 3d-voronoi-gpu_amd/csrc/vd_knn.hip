// Exact k-nearest-neighbour search on gfx950.
//
// Replaces the reference's grid kNN (knearests.cu:19-145, 230-386) and its
// brute-force kNN (extern/kNN-CUDA/knncuda.cu:20-91, 143-189, 271-349).
//
// Result contract: for every query, the k reference points with the smallest
// (d2, index) keys, nearest first, where d2 = fma(dz,dz, fma(dy,dy, dx*dx)) in
// FP32 with d = ref - query.  The output does not depend on the grid, the
// launch shape or the order of work, so it is bit-exact against any exact
// implementation that uses the same key.
//
// Grid: a uniform grid over the bounding box of the reference points (and an
// optional extra box), sized on the device so that a whole Lloyd iteration can
// be captured in a graph.  Binning is a counting sort (count, scan, scatter)
// followed by an in-cell sort by index, so the sorted order is deterministic.
//
// Query: one wavefront per query.  The wave walks a table of cell offsets
// ordered by a query-independent lower bound of their distance, loads the
// points of 64 cells at a time into its 64 lanes, filters them against the
// current k-th key, stages survivors in LDS and merges every 64 of them into a
// sorted top-k held in registers (64*M slots, bitonic sort + bitonic merge
// across lanes).  It stops when the lower bound of the next offsets exceeds
// the k-th distance; offsets beyond the table fall back to a full scan.
#include "vd_knn.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace vd {

namespace {

constexpr int kScanThreads = 1024;

__device__ __forceinline__ int cell_coord(float x, float lo, float inv_h, int dim) {
  int c = (int)floorf((x - lo) * inv_h);
  return min(max(c, 0), dim - 1);
}

__device__ __forceinline__ uint32_t bperm32(uint32_t v, int src_lane) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src_lane << 2, (int)v);
}
__device__ __forceinline__ uint64_t bperm64(uint64_t v, int src_lane) {
  const uint32_t lo = bperm32((uint32_t)v, src_lane);
  const uint32_t hi = bperm32((uint32_t)(v >> 32), src_lane);
  return ((uint64_t)hi << 32) | lo;
}

// Value of lane (lane ^ j) for a compile-time j: within a row of 16 lanes
// through DPP moves (quad permutes for 1 and 2; row rotations, where lane i
// reads lane i - n of its row, for 4 and 8), which stay in the VALU, and
// through the LDS crossbar (ds_bpermute) for 16 and 32.
__device__ __forceinline__ uint32_t xor_lane32(uint32_t v, int j, int lane) {
  const int x = (int)v;
  switch (j) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    case 4: {
      const int from_below = __builtin_amdgcn_mov_dpp(x, 0x124, 0xF, 0xF, false);  // row_ror:4, lane i - 4
      const int from_above = __builtin_amdgcn_mov_dpp(x, 0x12C, 0xF, 0xF, false);  // row_ror:12, lane i + 4
      return (uint32_t)((lane & 4) ? from_below : from_above);
    }
    case 8: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror:8
    default: return bperm32(v, lane ^ j);
  }
}
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v, int j, int lane) {
  const uint32_t lo = xor_lane32((uint32_t)v, j, lane);
  const uint32_t hi = xor_lane32((uint32_t)(v >> 32), j, lane);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum across the wave through DPP moves: row shifts by 1, 2,
// 4 and 8 within each row of 16 lanes (lanes without a source add 0), then
// the last lane of row 0 into row 1 and of row 2 into row 3 (row_bcast:15),
// then the last lane of row 1 into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ int wave_incl_sum(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1 and 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2 and 3
  return x;
}
// Inclusive prefix maximum across the wave, the same DPP steps (lanes
// without a source contribute -1).
__device__ __forceinline__ int wave_incl_max(int x) {
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x111, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x112, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x114, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x118, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x142, 0xA, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x143, 0xC, 0xF, false));
  return x;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Owner lane of flat index f0 + lane of the cells' concatenated point ranges
// (the largest lane whose range starts at or before it): every lane with
// points whose range starts in [f0, f0 + 64) marks its start in `mark`, a
// prefix maximum over the marks finds the latest start, and `carry` (the
// owner of f0 - 1, 0 at first) covers ranges begun in earlier batches.
__device__ __forceinline__ int batch_owner(int* mark, int lane, int cnt, int excl, int f0, int& carry) {
  mark[lane] = -1;
  wave_sync();
  if (cnt > 0 && excl >= f0 && excl < f0 + 64) mark[excl - f0] = lane;
  wave_sync();
  const int own = max(wave_incl_max(mark[lane]), carry);
  carry = __builtin_amdgcn_readlane(own, 63);
  wave_sync();
  return own;
}
__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a < b ? b : a; }

__device__ __forceinline__ uint64_t knn_key(float d2, uint32_t id) {
  return ((uint64_t)__float_as_uint(d2) << 32) | id;
}
__device__ __forceinline__ float key_d2(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }

__device__ __forceinline__ float dist2(const float4& r, const float4& q) {
  const float dx = r.x - q.x, dy = r.y - q.y, dz = r.z - q.z;
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// ---------------------------------------------------------------- grid ----

// Bounding box of the points (and of `extra`, a box given by the host), then
// the grid shape: about n/density cells of equal size, at most max_cells.
// One block per instance.
__global__ __launch_bounds__(kScanThreads) void k_grid_shape(const float4* pts, int n, const int* pt_begin,
                                                              const int* cell_base, const float* box, float3 extra_lo,
                                                              float3 extra_hi, int use_extra, float density,
                                                              int max_cells, GridShape* shapes) {
  __shared__ float red[6][kScanThreads / 64];
  const int inst = blockIdx.x;
  const int p0 = pt_begin ? pt_begin[inst] : 0, p1 = pt_begin ? pt_begin[inst + 1] : n;
  if (cell_base) max_cells = cell_base[inst + 1] - cell_base[inst];
  if (box) {
    extra_lo = make_float3(box[6 * inst], box[6 * inst + 1], box[6 * inst + 2]);
    extra_hi = make_float3(box[6 * inst + 3], box[6 * inst + 4], box[6 * inst + 5]);
    use_extra = 1;
  }
  n = p1 - p0;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
    const float4 p = pts[i];
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off));
    }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      red[a][w] = lo[a];
      red[3 + a][w] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 1; j < (int)(blockDim.x >> 6); ++j)
      for (int a = 0; a < 3; ++a) {
        red[a][0] = fminf(red[a][0], red[a][j]);
        red[3 + a][0] = fmaxf(red[3 + a][0], red[3 + a][j]);
      }
    double l[3], h[3];
    const float elo[3] = {extra_lo.x, extra_lo.y, extra_lo.z}, ehi[3] = {extra_hi.x, extra_hi.y, extra_hi.z};
    for (int a = 0; a < 3; ++a) {
      l[a] = red[a][0];
      h[a] = red[3 + a][0];
      if (use_extra) {
        l[a] = fmin(l[a], (double)elo[a]);
        h[a] = fmax(h[a], (double)ehi[a]);
      }
      if (!(l[a] <= h[a])) { l[a] = 0.0; h[a] = 0.0; }
    }
    double ext = fmax(fmax(h[0] - l[0], h[1] - l[1]), fmax(h[2] - l[2], 1e-20));
    // pad the box so that no point sits on the upper face
    for (int a = 0; a < 3; ++a) {
      l[a] -= 1e-4 * ext;
      h[a] += 1e-4 * ext;
    }
    ext *= 1.0002;
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= fmax(h[a] - l[a], ext * 1e-3);
    const double cells = fmax(1.0, (double)n / (double)density);
    double cs = cbrt(vol / cells);
    int dim[3];
    for (;;) {
      long long tot = 1;
      for (int a = 0; a < 3; ++a) {
        dim[a] = (int)fmin(1024.0, fmax(1.0, ceil((h[a] - l[a]) / cs)));
        tot *= dim[a];
      }
      if (tot <= max_cells) break;
      cs *= 1.05;
    }
    GridShape g;
    for (int a = 0; a < 3; ++a) {
      g.lo[a] = (float)l[a];
      g.dim[a] = dim[a];
    }
    g.h = (float)cs;
    g.inv_h = (float)(1.0 / cs);
    g.ncell = dim[0] * dim[1] * dim[2];
    // absolute slack for the cell-box distance bounds (rounding of the cell
    // assignment and of lo + c*h)
    g.slack = (float)(1e-3 * cs + 1e-6 * (fabs(l[0]) + fabs(l[1]) + fabs(l[2]) + 1024.0 * cs));
    shapes[inst] = g;
  }
}

__global__ void k_grid_count(const float4* pts, int n, const GridShape* gs, const int* pt_inst, const int* cell_base,
                             int* cell_of, int* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int inst = pt_inst ? pt_inst[i] : 0;
  const GridShape g = gs[inst];
  const float4 p = pts[i];
  const int cx = cell_coord(p.x, g.lo[0], g.inv_h, g.dim[0]);
  const int cy = cell_coord(p.y, g.lo[1], g.inv_h, g.dim[1]);
  const int cz = cell_coord(p.z, g.lo[2], g.inv_h, g.dim[2]);
  const int c = (cell_base ? cell_base[inst] : 0) + cx + g.dim[0] * (cy + g.dim[1] * cz);
  cell_of[i] = c;
  atomicAdd(count + c, 1);
}

__global__ void k_grid_scatter(int n, const int* cell_of, const int* start, int* cursor, int* sorted_ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  sorted_ids[start[c] + atomicAdd(cursor + c, 1)] = i;
}

// One thread per cell: order the cell's ids, then gather the points.
// n_cells < 0: the cells of the single instance gs[0].
__global__ void k_grid_sort(const GridShape* gs, int n_cells, const float4* pts, const int* start, int* sorted_ids,
                            float4* sorted_pts) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (n_cells >= 0 ? n_cells : gs->ncell)) return;
  const int b = start[c], e = start[c + 1];
  for (int i = b + 1; i < e; ++i) {
    const int v = sorted_ids[i];
    int j = i - 1;
    while (j >= b && sorted_ids[j] > v) {
      sorted_ids[j + 1] = sorted_ids[j];
      --j;
    }
    sorted_ids[j + 1] = v;
  }
  for (int i = b; i < e; ++i) sorted_pts[i] = pts[sorted_ids[i]];
}

// ---------------------------------------------------------------- query ---

struct QueryArgs {
  const GridShape* grid;  // per instance
  const int* cell_base;   // per instance, or null (one instance)
  const int* pt_inst;     // instance of every point, or null
  const int* q_inst;      // modes 1 and 2: instance of every query, or null
  const float4* sorted_pts;
  const int* sorted_ids;
  const int* start;
  int n_ref;
  const int4* offsets;  // (dx, dy, dz, lb) sorted by lb = sum max(|d|-1,0)^2
  int n_offsets;
  int k;
  int n_query;
  int mode;  // 0: queries are the reference points themselves (self excluded)
             // 1: queries are tet centroids
             // 2: queries are given points (self not excluded)
             // 3: queries are the points qpts[q_ids[q]] with ids q_ids[q] (self excluded)
  const int* q_ids;
  const unsigned* n_query_dev;  // optional device count, capping n_query
  const float4* verts;
  const int4* tets;
  const float4* qpts;
  int32_t* out;        // row-major n_query x k, or pitched rows when pitch > 0
  long long out_pitch; // > 0: out[j * pitch + q] (column per query)
  // modes 0 and 1 with row-major output: when hint_pts is given, every row of
  // `out` holds a previous result of this problem (k distinct reference ids of
  // the query's instance, the query itself excluded) and hint_pts holds the
  // reference points by original id
  const float4* hint_pts;
  const int* pt_begin;  // first point of every instance (n_inst + 1), or null
};

template <int M>
struct TopK {
  uint64_t s[M];
};

// Bitonic sort of one key per lane across the wave, ascending.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t x, int lane) {
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = xor_lane64(x, j, lane);
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      x = (lower == up) ? umin64(x, o) : umax64(x, o);
    }
  }
  return x;
}

// Sorted merge of 64 new keys (one per lane) into the top list; into an
// empty list (first) the sorted keys are the list.
template <int M>
__device__ __forceinline__ void topk_merge(TopK<M>& L, uint64_t x, int lane, bool first = false) {
  x = wave_sort64(x, lane);
  if (first) {
    L.s[0] = x;
    return;
  }
  // min against the reversed batch gives the 64*M smallest as a bitonic sequence
  const uint64_t r = bperm64(x, 63 - lane);
  L.s[M - 1] = umin64(L.s[M - 1], r);
#pragma unroll
  for (int js = M / 2; js >= 1; js >>= 1) {
#pragma unroll
    for (int s = 0; s < M; ++s) {
      if ((s & js) == 0) {
        const uint64_t a = L.s[s], b = L.s[s + js];
        L.s[s] = umin64(a, b);
        L.s[s + js] = umax64(a, b);
      }
    }
  }
#pragma unroll
  for (int j = 32; j > 0; j >>= 1) {
    const bool lower = (lane & j) == 0;
#pragma unroll
    for (int s = 0; s < M; ++s) {
      const uint64_t o = xor_lane64(L.s[s], j, lane);
      L.s[s] = lower ? umin64(L.s[s], o) : umax64(L.s[s], o);
    }
  }
}

template <int M>
__device__ __forceinline__ uint64_t topk_kth(const TopK<M>& L, int k) {
  const int slot = (k - 1) >> 6, ln = (k - 1) & 63;
  uint64_t v = 0;
#pragma unroll
  for (int s = 0; s < M; ++s)
    if (s == slot) v = L.s[s];
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, ln);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), ln);
  return ((uint64_t)hi << 32) | lo;
}

// Inserts one key (wave-uniform) into the sorted list: the keys after its
// place move up one slot and the largest drops out.
template <int M>
__device__ __forceinline__ void topk_insert(TopK<M>& L, uint64_t x, int lane) {
  int pos = 0;
#pragma unroll
  for (int s = 0; s < M; ++s) pos += __popcll(__ballot(L.s[s] < x));
  if (pos >= 64 * M) return;
  uint64_t carry = x;  // last key of the previous slot (x before slot 0)
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const uint64_t v = L.s[s];
    const uint64_t up = bperm64(v, lane == 0 ? 0 : lane - 1);
    const int e = s * 64 + lane;
    const uint64_t prev = lane == 0 ? carry : up;
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    carry = ((uint64_t)hi << 32) | lo;
    L.s[s] = e < pos ? v : (e == pos ? x : prev);
  }
}

// Largest squared distance from qp to the k points of a previous result row
// (hint_pts by id), or -1 when the row is not k distinct points of the
// instance [lo, hi) other than the query itself.  Any k such points bound the
// query's k-th squared distance from above.
template <int M>
__device__ __forceinline__ float hint_bound(const QueryArgs& a, int out_row, const float4& qp, int self, int lo,
                                            int hi, int lane) {
  float mx = 0.f;
  bool bad = false;
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int j = s * 64 + lane;
    if (j < a.k) {
      const int id = a.out[(long long)out_row * a.k + j];
      if (id < lo || id >= hi || id == self)
        bad = true;
      else
        mx = fmaxf(mx, dist2(a.hint_pts[id], qp));
    }
  }
  if (__ballot(bad) != 0) return -1.f;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  return mx;
}

// Squared distance from q to the box of cell (cx,cy,cz), minus the slack,
// scaled down so that any point binned in that cell has a larger FP32 d2.
__device__ __forceinline__ float cell_bound(const GridShape& g, const float4& q, int cx, int cy, int cz) {
  const int c[3] = {cx, cy, cz};
  const float qq[3] = {q.x, q.y, q.z};
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = g.lo[a] + (float)c[a] * g.h;
    const float hi = lo + g.h;
    float d = 0.f;
    if (qq[a] < lo) d = lo - qq[a];
    if (qq[a] > hi) d = qq[a] - hi;
    d = fmaxf(0.f, d - g.slack);
    s = fmaf(d, d, s);
  }
  return s * (1.f - 1e-5f);
}

// Waves per workgroup of the query kernel: 2, 4 or 8 (knn_query's `waves`);
// each wave works alone, the workgroup size only changes how waves share a
// CU's LDS and are scheduled.
constexpr int kQueryWavesDefault = 4;

// Consecutive queries (tets of one region, or sites of one grid cell) are
// close to each other: a wave takes a run of kQueryRun of them, and each
// query starts from the previous one's k-th distance r and their distance d:
// the previous query's k neighbours lie within r + d of it (with itself in
// place of one that it excludes, mode 0 and 3), so its k-th key is below that
// bound (rounded up by a 1e-5 margin) and candidates beyond it are never
// staged, and the cell walk stops there.
#ifndef VD_KNN_RUN
#define VD_KNN_RUN 6
#endif
constexpr int kQueryRun = VD_KNN_RUN;

// After the first query of a run, the points that the run's other queries can
// need are gathered once into LDS (at most kUnionCap): query i's k-th distance
// is at most r0 + |q_i - q_0| (the first query's k neighbours, with q_0 in
// place of q_i itself where the query excludes itself), so every point such a
// query can keep lies within R = max_i(|q_i - c| + r0 + |q_i - q_0|) of the
// centre c of the run.  Those queries scan the list instead of walking the
// grid, with the same keys and the same bounds, so their results are the
// same; a run whose list would not fit walks as before.
#ifndef VD_KNN_UNION_CAP
#define VD_KNN_UNION_CAP 512
#endif
constexpr int kUnionCap = VD_KNN_UNION_CAP;

// At most this many keys staged after the last merge are inserted one at a
// time instead of sorted and merged.
#ifndef VD_KNN_INSERT_MAX
#define VD_KNN_INSERT_MAX 12
#endif
constexpr int kInsertMax = VD_KNN_INSERT_MAX;

// Point of query q, its excluded reference id (or -1), output row and instance.
__device__ __forceinline__ void query_point(const QueryArgs& a, int q, float4& qp, int& self, int& out_row, int& inst) {
  self = -1;
  out_row = q;
  inst = 0;
  if (a.mode == 0) {
    qp = a.sorted_pts[q];
    self = a.sorted_ids[q];
    out_row = self;
    if (a.pt_inst) inst = a.pt_inst[self];
  } else if (a.mode == 3) {
    self = a.q_ids[q];
    qp = a.qpts[self];
    if (a.pt_inst) inst = a.pt_inst[self];
  } else {
    if (a.q_inst) inst = a.q_inst[q];
    if (a.mode == 1) {
      const int4 t = a.tets[q];
      const float4 v0 = a.verts[t.x], v1 = a.verts[t.y], v2 = a.verts[t.z], v3 = a.verts[t.w];
      qp.x = (((v0.x + v1.x) + v2.x) + v3.x) * 0.25f;
      qp.y = (((v0.y + v1.y) + v2.y) + v3.y) * 0.25f;
      qp.z = (((v0.z + v1.z) + v2.z) + v3.z) * 0.25f;
      qp.w = 0.f;
    } else {
      qp = a.qpts[q];
    }
  }
}

template <int M, int kQueryWaves>
__global__ __launch_bounds__(64 * kQueryWaves) void k_knn_query(QueryArgs a) {
  __shared__ uint64_t stage_all[kQueryWaves][128];
  __shared__ int mark_all[kQueryWaves][64];  // batch_owner's marks
  __shared__ float4 uni_all[kQueryWaves][kUnionCap];  // (x, y, z, id bits)
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  uint64_t* stage = stage_all[wv];
  int* mark = mark_all[wv];
  float4* uni = uni_all[wv];
  const int q0 = (blockIdx.x * kQueryWaves + wv) * kQueryRun;
  int nq = a.n_query;
  if (a.n_query_dev) nq = min(nq, (int)min(*a.n_query_dev, (unsigned)a.n_query));
  float prev_r = -1.f;
  float4 prev_qp = {0.f, 0.f, 0.f, 0.f};
  int prev_inst = -1;
  // the run's point list: n_uni points (-1: none), valid for instance uni_inst
  int n_uni = -1, uni_inst = -1;
  float r0 = 0.f;
  float4 q0p = {0.f, 0.f, 0.f, 0.f};
  // with previous results (hint_pts): lane j holds the hint bound of query
  // q0 + j once it is known (-2 before; -1: that row gives none)
  const bool use_hint = a.hint_pts && a.out_pitch <= 0 && (a.mode == 0 || a.mode == 1);
  float run_hb = -2.f;
  // The run's point list: every point of the instance within sqrt(R2) of c,
  // with its id in w; returns their number, or -1 when they do not fit in
  // kUnionCap or the offset table does not reach that far (beyond it the
  // walk scans the whole instance: no list).
  auto gather = [&](const float4& c, float R2, const GridShape& g, int cbase) -> int {
    const int cc[3] = {cell_coord(c.x, g.lo[0], g.inv_h, g.dim[0]), cell_coord(c.y, g.lo[1], g.inv_h, g.dim[1]),
                       cell_coord(c.z, g.lo[2], g.inv_h, g.dim[2])};
    int n = 0;
    bool fits = true, covered = false;
    for (int ob = 0; ob < a.n_offsets && fits; ob += 64) {
      const float lbd = fmaxf(0.f, sqrtf((float)a.offsets[ob].w) * g.h - 2.f * g.slack);
      if (lbd * lbd * (1.f - 1e-5f) > R2) {
        covered = true;  // offsets are sorted by lb: the rest are beyond R too
        break;
      }
      int cnt = 0, st = 0;
      const int oi = ob + lane;
      if (oi < a.n_offsets) {
        const int4 o = a.offsets[oi];
        const int ox = cc[0] + o.x, oy = cc[1] + o.y, oz = cc[2] + o.z;
        if (ox >= 0 && ox < g.dim[0] && oy >= 0 && oy < g.dim[1] && oz >= 0 && oz < g.dim[2] &&
            !(cell_bound(g, c, ox, oy, oz) > R2)) {
          const int cell = cbase + ox + g.dim[0] * (oy + g.dim[1] * oz);
          st = a.start[cell];
          cnt = a.start[cell + 1] - st;
        }
      }
      const int incl = wave_incl_sum(cnt);
      const int excl = incl - cnt;
      const int total = __shfl(incl, 63);
      int carry = 0;
      for (int f0 = 0; f0 < total && fits; f0 += 64) {
        const int f = f0 + lane;
        const int own = batch_owner(mark, lane, cnt, excl, f0, carry);
        const int o_st = __shfl(st, own), o_ex = __shfl(excl, own);
        bool keep = false;
        float4 p = {0.f, 0.f, 0.f, 0.f};
        if (f < total) {
          const int idx = o_st + (f - o_ex);
          p = a.sorted_pts[idx];
          p.w = __uint_as_float((uint32_t)a.sorted_ids[idx]);
          keep = !(dist2(p, c) > R2);
        }
        const uint64_t m = __ballot(keep);
        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        const int add = __popcll(m);
        if (n + add > kUnionCap) {
          fits = false;
        } else {
          if (keep) uni[n + below] = p;
          n += add;
        }
      }
    }
    if (!(fits && covered)) return -1;
    wave_sync();
    return n;
  };
  // With previous results (hint_pts) every query of the run has a bound on its
  // k-th distance before any walk: the run's list is then gathered first,
  // within max_j(|q_j - c| + r_j) of the centre c of the run's queries (r_j:
  // query j's bound), and every query of the run, the first one included,
  // scans it instead of walking the grid, with the walk's keys and bounds.
  if (use_hint) {
    const int n_run = min(q0 + kQueryRun, nq) - q0;
    float4 pj = {0.f, 0.f, 0.f, 0.f};
    int sj = -1, rj = 0, ij = 0;
    const bool in = lane < n_run;
    if (in) query_point(a, q0 + lane, pj, sj, rj, ij);
    const int inst0 = __shfl(ij, 0);
    if (n_run > 0 && __ballot(in && ij != inst0) == 0) {
      const int lo = a.pt_begin ? a.pt_begin[inst0] : 0, hi = a.pt_begin ? a.pt_begin[inst0 + 1] : a.n_ref;
      for (int j = 0; j < n_run; ++j) {
        const float4 qj4 = make_float4(__shfl(pj.x, j), __shfl(pj.y, j), __shfl(pj.z, j), 0.f);
        const float hb = hint_bound<M>(a, __shfl(rj, j), qj4, __shfl(sj, j), lo, hi, lane);
        if (lane == j) run_hb = hb;
      }
      if (__ballot(in && !(run_hb >= 0.f)) == 0) {
        float cx = in ? pj.x : 0.f, cy = in ? pj.y : 0.f, cz = in ? pj.z : 0.f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          cx += __shfl_xor(cx, o);
          cy += __shfl_xor(cy, o);
          cz += __shfl_xor(cz, o);
        }
        const float inv = 1.f / (float)n_run;
        const float4 c = make_float4(cx * inv, cy * inv, cz * inv, 0.f);
        float need = 0.f;
        if (in) {
          const float ax = pj.x - c.x, ay = pj.y - c.y, az = pj.z - c.z;
          need = sqrtf(fmaf(az, az, fmaf(ay, ay, ax * ax))) + sqrtf(run_hb);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) need = fmaxf(need, __shfl_xor(need, o));
        const GridShape g = a.grid[inst0];
        // rounding margins of the FP32 distances on either side
        const float R = need * (1.f + 1e-4f) + 1e-3f * g.h;
        const int n = gather(c, R * R, g, a.cell_base ? a.cell_base[inst0] : 0);
        if (n >= 0) {
          n_uni = n;
          uni_inst = inst0;
          r0 = INFINITY;  // the queries' own bounds hold
        }
      }
    }
  }
  for (int q = q0; q < q0 + kQueryRun && q < nq; ++q) {  // wave-uniform

  float4 qp;
  int self, out_row, inst;
  query_point(a, q, qp, self, out_row, inst);
  const GridShape g = a.grid[inst];
  const int cbase = a.cell_base ? a.cell_base[inst] : 0;
  const bool from_list = n_uni >= 0 && inst == uni_inst;
  const int qc[3] = {cell_coord(qp.x, g.lo[0], g.inv_h, g.dim[0]), cell_coord(qp.y, g.lo[1], g.inv_h, g.dim[1]),
                     cell_coord(qp.z, g.lo[2], g.inv_h, g.dim[2])};

  TopK<M> L;
#pragma unroll
  for (int s = 0; s < M; ++s) L.s[s] = ~0ull;
  uint64_t bound = ~0ull;
  if (prev_r >= 0.f && prev_inst == inst) {
    const float dx = qp.x - prev_qp.x, dy = qp.y - prev_qp.y, dz = qp.z - prev_qp.z;
    const float b = prev_r + sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    const float b2 = b * b * (1.f + 1e-5f);
    if (b2 < INFINITY) bound = ((uint64_t)__float_as_uint(b2) << 32) | 0xffffffffull;
  }
  if (from_list) {
    // the bound the list was gathered for: r0 + |q - q0|
    const float dx = qp.x - q0p.x, dy = qp.y - q0p.y, dz = qp.z - q0p.z;
    const float b = r0 + sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    const float b2 = b * b * (1.f + 1e-5f);
    bound = umin64(bound, ((uint64_t)__float_as_uint(b2) << 32) | 0xffffffffull);
  }
  if (use_hint) {
    // the query's previous result: its k points' current distances
    float hb = __shfl(run_hb, q - q0);
    if (hb == -2.f) {
      const int lo = a.pt_begin ? a.pt_begin[inst] : 0, hi = a.pt_begin ? a.pt_begin[inst + 1] : a.n_ref;
      hb = hint_bound<M>(a, out_row, qp, self, lo, hi, lane);
    }
    if (hb >= 0.f) bound = umin64(bound, ((uint64_t)__float_as_uint(hb) << 32) | 0xffffffffull);
  }
  uint64_t thr = bound;
  int staged = 0;
  bool done = false;
  bool empty = true;  // no key merged yet (wave-uniform)

  auto push = [&](uint64_t key, bool ok) {
    const uint64_t mask = __ballot(ok);
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
    if (ok) stage[staged + below] = key;
    staged += __popcll(mask);
    wave_sync();
    if (staged >= 64) {
      const uint64_t x = stage[lane];
      const uint64_t rest = (lane + 64 < staged) ? stage[lane + 64] : ~0ull;
      wave_sync();
      if (lane + 64 < staged) stage[lane] = rest;
      staged -= 64;
      wave_sync();
      topk_merge<M>(L, x, lane, empty);
      empty = false;
      thr = umin64(bound, topk_kth<M>(L, a.k));
    }
  };

  if (from_list) {
    for (int f0 = 0; f0 < n_uni; f0 += 64) {
      const int f = f0 + lane;
      uint64_t key = ~0ull;
      bool ok = false;
      if (f < n_uni) {
        const float4 p = uni[f];
        const uint32_t id = __float_as_uint(p.w);
        key = knn_key(dist2(p, qp), id);
        ok = ((int)id != self) && key < thr;
      }
      push(key, ok);
    }
    done = true;
  }
  for (int ob = 0; ob < a.n_offsets && !done; ob += 64) {
    const float kth = key_d2(thr);
    // offsets are sorted by lb: once the first of the batch is beyond the k-th
    // distance, so are all later ones
    const float lbd = fmaxf(0.f, sqrtf((float)a.offsets[ob].w) * g.h - 2.f * g.slack);
    if (thr != ~0ull && lbd * lbd * (1.f - 1e-5f) > kth) {
      done = true;
      break;
    }
    int cnt = 0, st = 0;
    const int oi = ob + lane;
    if (oi < a.n_offsets) {
      const int4 o = a.offsets[oi];
      const int cx = qc[0] + o.x, cy = qc[1] + o.y, cz = qc[2] + o.z;
      if (cx >= 0 && cx < g.dim[0] && cy >= 0 && cy < g.dim[1] && cz >= 0 && cz < g.dim[2]) {
        if (thr == ~0ull || !(cell_bound(g, qp, cx, cy, cz) > kth)) {
          const int c = cbase + cx + g.dim[0] * (cy + g.dim[1] * cz);
          st = a.start[c];
          cnt = a.start[c + 1] - st;
        }
      }
    }
    // exclusive prefix of the counts across the wave
    const int incl = wave_incl_sum(cnt);
    const int excl = incl - cnt;
    const int total = __shfl(incl, 63);
    int carry = 0;
    for (int f0 = 0; f0 < total; f0 += 64) {
      const int f = f0 + lane;
      const int own = batch_owner(mark, lane, cnt, excl, f0, carry);
      const int o_st = __shfl(st, own), o_ex = __shfl(excl, own);
      uint64_t key = ~0ull;
      bool ok = false;
      if (f < total) {
        const int idx = o_st + (f - o_ex);
        const int id = a.sorted_ids[idx];
        key = knn_key(dist2(a.sorted_pts[idx], qp), (uint32_t)id);
        ok = (id != self) && key < thr;
      }
      push(key, ok);
    }
  }
  if (!done) {
    // offsets table exhausted without the stop test passing: scan all points
    // of the instance
    const int r0 = a.start[cbase], r1 = a.start[cbase + g.ncell];
    for (int f0 = r0; f0 < r1; f0 += 64) {
      const int f = f0 + lane;
      uint64_t key = ~0ull;
      bool ok = false;
      if (f < r1) {
        // skip cells already visited: those inside the offset table's reach
        const float4 p = a.sorted_pts[f];
        const int id = a.sorted_ids[f];
        const int cx = cell_coord(p.x, g.lo[0], g.inv_h, g.dim[0]) - qc[0];
        const int cy = cell_coord(p.y, g.lo[1], g.inv_h, g.dim[1]) - qc[1];
        const int cz = cell_coord(p.z, g.lo[2], g.inv_h, g.dim[2]) - qc[2];
        const int r = max(abs(cx), max(abs(cy), abs(cz)));
        key = knn_key(dist2(p, qp), (uint32_t)id);
        ok = (r > kOffsetRadius) && (id != self) && key < thr;
      }
      push(key, ok);
    }
  }
  if (staged > 0 && !empty && staged <= kInsertMax) {
    // a few keys past the last merge: inserted one by one
    for (int i = 0; i < staged; ++i) {
      const uint64_t x = stage[i];
      if (x < thr) {
        topk_insert<M>(L, x, lane);
        thr = umin64(bound, topk_kth<M>(L, a.k));
      }
    }
  } else if (staged > 0) {
    const uint64_t x = lane < staged ? stage[lane] : ~0ull;
    topk_merge<M>(L, x, lane, empty);
  }
#pragma unroll
  for (int s = 0; s < M; ++s) {
    const int j = s * 64 + lane;
    if (j < a.k) {
      const int32_t id = L.s[s] == ~0ull ? -1 : (int32_t)(uint32_t)L.s[s];
      if (a.out_pitch > 0)
        a.out[(long long)j * a.out_pitch + out_row] = id;
      else
        a.out[(long long)out_row * a.k + j] = id;
    }
  }
  const uint64_t kth = topk_kth<M>(L, a.k);
  prev_r = kth == ~0ull ? -1.f : sqrtf(key_d2(kth));
  prev_qp = qp;
  prev_inst = inst;

  const int q_end = min(q0 + kQueryRun, nq);
  if (q == q0 && n_uni < 0 && prev_r >= 0.f && q + 1 < q_end) {
    // the run's other queries: their points (lane j takes query q0 + j)
    float4 pj = qp;
    int sj = self, rj = out_row, ij = inst;
    const int qj = q0 + lane;
    if (lane >= 1 && qj < q_end) query_point(a, qj, pj, sj, rj, ij);
    const bool in = lane >= 1 && qj < q_end;
    const bool same = __ballot(in && ij != inst) == 0;
    if (same) {
      const int n_in = q_end - q0 - 1;
      float cx = in ? pj.x : 0.f, cy = in ? pj.y : 0.f, cz = in ? pj.z : 0.f;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        cx += __shfl_xor(cx, o);
        cy += __shfl_xor(cy, o);
        cz += __shfl_xor(cz, o);
      }
      const float inv = 1.f / (float)n_in;
      const float4 c = make_float4(cx * inv, cy * inv, cz * inv, 0.f);
      if (use_hint) {
        // the other queries' bounds from their previous results (lane j: query q0 + j)
        const int lo = a.pt_begin ? a.pt_begin[inst] : 0, hi = a.pt_begin ? a.pt_begin[inst + 1] : a.n_ref;
        for (int j = 1; j <= n_in; ++j) {
          const float4 qj4 = make_float4(__shfl(pj.x, j), __shfl(pj.y, j), __shfl(pj.z, j), 0.f);
          const int sj_row = __shfl(rj, j), sj_self = __shfl(sj, j);
          const float hb = hint_bound<M>(a, sj_row, qj4, sj_self, lo, hi, lane);
          if (lane == j) run_hb = hb;
        }
      }
      float rj_need = 0.f;
      if (in) {
        const float ax = pj.x - c.x, ay = pj.y - c.y, az = pj.z - c.z;
        const float bx = pj.x - qp.x, by = pj.y - qp.y, bz = pj.z - qp.z;
        float rq = prev_r + sqrtf(fmaf(bz, bz, fmaf(by, by, bx * bx)));
        if (run_hb >= 0.f) rq = fminf(rq, sqrtf(run_hb));
        rj_need = sqrtf(fmaf(az, az, fmaf(ay, ay, ax * ax))) + rq;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) rj_need = fmaxf(rj_need, __shfl_xor(rj_need, o));
      // rounding margins of the FP32 distances on either side
      const float R = rj_need * (1.f + 1e-4f) + 1e-3f * g.h;
      const float R2 = R * R;
      const int n = gather(c, R2, g, cbase);
      if (n >= 0) {
        n_uni = n;
        uni_inst = inst;
        r0 = prev_r;
        q0p = qp;
      }
    }
  }
  }  // query loop
}

}  // namespace

__global__ void k_fill_u32(uint32_t* p, long long n, uint32_t v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, n, v);
  return hipGetLastError();
}

// Exclusive scan, three passes: per-block scans of kScanTile items, a scan of
// the block totals, then the block offsets added back.  Also clears `cursor`.
constexpr int kScanItems = 8;
constexpr int kScanTile = 1024 * kScanItems;

__device__ __forceinline__ int scan_n(const int* n_dev, int n_max) { return n_dev ? min(*n_dev, n_max) : n_max; }

// T: int, or unsigned for totals up to 2^32 - 1.
template <class T>
__global__ __launch_bounds__(1024) void k_scan_tiles(const int* in, const int* n_dev, int n_max, T* out,
                                                      int* cursor, T* tile_sum) {
  __shared__ T part[1024];
  const int n = scan_n(n_dev, n_max);
  const long long base = (long long)blockIdx.x * kScanTile;
  if (base >= n && blockIdx.x > 0) return;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const long long i = base + (long long)j * 1024 + threadIdx.x;
    v[j] = i < n ? (T)in[i] : (T)0;
  }
  // thread t owns items base + t*kScanItems .. ; reorder through LDS for that
  __shared__ T tile[kScanTile];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) tile[j * 1024 + threadIdx.x] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = tile[threadIdx.x * kScanItems + j];
    s += v[j];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const T x = threadIdx.x >= off ? part[threadIdx.x - off] : (T)0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  T run = part[threadIdx.x] - s;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    tile[threadIdx.x * kScanItems + j] = run;
    run += v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const long long i = base + (long long)j * 1024 + threadIdx.x;
    if (i < n) {
      out[i] = tile[j * 1024 + threadIdx.x];
      if (cursor) cursor[i] = 0;
    }
  }
  if (threadIdx.x == 1023) tile_sum[blockIdx.x] = part[1023];
}

template <class T>
__global__ __launch_bounds__(1024) void k_scan_top(const int* n_dev, int n_max, T* tile_sum) {
  __shared__ T part[1024];
  const int n = scan_n(n_dev, n_max);
  const int tiles = (n + kScanTile - 1) / kScanTile;
  const int per = (tiles + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(tiles, b + per);
  T s = 0;
  for (int i = b; i < e; ++i) s += tile_sum[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const T x = threadIdx.x >= off ? part[threadIdx.x - off] : (T)0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  T run = part[threadIdx.x] - s;
  for (int i = b; i < e; ++i) {
    const T t = tile_sum[i];
    tile_sum[i] = run;
    run += t;
  }
  if (threadIdx.x == 1023) tile_sum[tiles] = part[1023];
}

template <class T>
__global__ void k_scan_add(const int* n_dev, int n_max, T* out, const T* tile_sum) {
  const int n = scan_n(n_dev, n_max);
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sum[i / kScanTile];
  if (i == n) out[n] = tile_sum[(n + kScanTile - 1) / kScanTile];
}

int scan_tmp_ints(int n_max) { return (n_max + kScanTile - 1) / kScanTile + 2; }

template <class T>
static hipError_t scan_excl_t(const int* in, const int* n_dev, int n_max, T* out, int* cursor, T* tmp,
                              hipStream_t st) {
  const int tiles = std::max(1, (n_max + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3(tiles), dim3(1024), 0, st, in, n_dev, n_max, out, cursor, tmp);
  hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(1024), 0, st, n_dev, n_max, tmp);
  hipLaunchKernelGGL(k_scan_add<T>, dim3((unsigned)(((long long)n_max + 1 + 255) / 256)), dim3(256), 0, st, n_dev,
                     n_max, out, tmp);
  return hipGetLastError();
}
hipError_t scan_excl(const int* in, const int* n_dev, int n_max, int* out, int* cursor, int* tmp, hipStream_t st) {
  return scan_excl_t<int>(in, n_dev, n_max, out, cursor, tmp, st);
}
hipError_t scan_excl(const int* in, const int* n_dev, int n_max, unsigned* out, int* cursor, unsigned* tmp,
                     hipStream_t st) {
  return scan_excl_t<unsigned>(in, n_dev, n_max, out, cursor, tmp, st);
}

std::vector<int4> knn_offset_table() {
  std::vector<int4> t;
  const int R = kOffsetRadius;
  for (int z = -R; z <= R; ++z)
    for (int y = -R; y <= R; ++y)
      for (int x = -R; x <= R; ++x) {
        auto m = [](int d) { return std::max(std::abs(d) - 1, 0); };
        t.push_back(make_int4(x, y, z, m(x) * m(x) + m(y) * m(y) + m(z) * m(z)));
      }
  std::stable_sort(t.begin(), t.end(), [](const int4& a, const int4& b) {
    if (a.w != b.w) return a.w < b.w;
    const int ca = std::max(std::abs(a.x), std::max(std::abs(a.y), std::abs(a.z)));
    const int cb = std::max(std::abs(b.x), std::max(std::abs(b.y), std::abs(b.z)));
    return ca < cb;
  });
  return t;
}

int grid_max_cells(int n, float density) {
  const double c = std::max(1.0, n / (double)density);
  return (int)std::min(64.0 * 1024 * 1024, std::max(64.0, 1.25 * c + 64.0));
}

hipError_t grid_build(const GridBufs& gb, const float4* pts, int n, const float* extra_lo, const float* extra_hi,
                      float density, hipStream_t st) {
  const float3 elo = extra_lo ? make_float3(extra_lo[0], extra_lo[1], extra_lo[2]) : make_float3(0, 0, 0);
  const float3 ehi = extra_hi ? make_float3(extra_hi[0], extra_hi[1], extra_hi[2]) : make_float3(0, 0, 0);
  if (gb.n_inst < 1 || (gb.n_inst > 1 && (!gb.cell_base || !gb.pt_begin || !gb.pt_inst))) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_grid_shape, dim3((unsigned)gb.n_inst), dim3(kScanThreads), 0, st, pts, n, gb.pt_begin,
                     gb.cell_base, gb.box, elo, ehi, extra_lo ? 1 : 0, density, gb.max_cells, gb.shape);
  hipError_t e = fill_u32(gb.count, gb.max_cells, 0u, st);
  if (e != hipSuccess) return e;
  const int nb = (n + 255) / 256;
  if (n > 0)
    hipLaunchKernelGGL(k_grid_count, dim3(nb), dim3(256), 0, st, pts, n, gb.shape, gb.pt_inst, gb.cell_base,
                       gb.cell_of, gb.count);
  // one instance: scan only the cells in use; a batch: all cells (unused
  // ones are empty)
  e = scan_excl(gb.count, gb.n_inst == 1 ? &gb.shape->ncell : nullptr, gb.max_cells, gb.start, gb.cursor,
                gb.scan_tmp, st);
  if (e != hipSuccess) return e;
  if (n > 0)
    hipLaunchKernelGGL(k_grid_scatter, dim3(nb), dim3(256), 0, st, n, gb.cell_of, gb.start, gb.cursor,
                       gb.sorted_ids);
  hipLaunchKernelGGL(k_grid_sort, dim3((gb.max_cells + 255) / 256), dim3(256), 0, st, gb.shape,
                     gb.n_inst == 1 ? -1 : gb.max_cells, pts, gb.start, gb.sorted_ids, gb.sorted_pts);
  return hipGetLastError();
}

namespace {
template <int W>
hipError_t launch_query_w(const QueryArgs& a, int n_query, hipStream_t st) {
  const dim3 grid((n_query + W * kQueryRun - 1) / (W * kQueryRun)), block(64 * W);
  if (a.k <= 64)
    hipLaunchKernelGGL((k_knn_query<1, W>), grid, block, 0, st, a);
  else if (a.k <= 128)
    hipLaunchKernelGGL((k_knn_query<2, W>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_knn_query<4, W>), grid, block, 0, st, a);
  return hipGetLastError();
}
hipError_t launch_query(const QueryArgs& a, int n_query, int waves, hipStream_t st) {
  if (waves == 2) return launch_query_w<2>(a, n_query, st);
  if (waves == 8) return launch_query_w<8>(a, n_query, st);
  return launch_query_w<kQueryWavesDefault>(a, n_query, st);
}
}  // namespace

hipError_t knn_query(const GridBufs& gb, int n_ref, const int4* offsets, int n_offsets, int k, int mode,
                     int n_query, const float4* verts, const int4* tets, const float4* qpts, const int* q_inst,
                     int32_t* out, long long out_pitch, hipStream_t st, const float4* hint_pts, int waves) {
  if (k < 1 || k > 256) return hipErrorInvalidValue;
  QueryArgs a;
  a.hint_pts = hint_pts;
  a.pt_begin = gb.pt_begin;
  a.grid = gb.shape;
  a.cell_base = gb.cell_base;
  a.pt_inst = gb.pt_inst;
  a.q_inst = q_inst;
  a.sorted_pts = gb.sorted_pts;
  a.sorted_ids = gb.sorted_ids;
  a.start = gb.start;
  a.n_ref = n_ref;
  a.offsets = offsets;
  a.n_offsets = n_offsets;
  a.k = k;
  a.n_query = n_query;
  a.mode = mode;
  a.verts = verts;
  a.tets = tets;
  a.qpts = qpts;
  a.out = out;
  a.out_pitch = out_pitch;
  a.q_ids = nullptr;
  a.n_query_dev = nullptr;
  if (n_query <= 0) return hipSuccess;
  return launch_query(a, n_query, waves, st);
}

hipError_t knn_query_ids(const GridBufs& gb, const int4* offsets, int n_offsets, int k, const float4* pts,
                         const int* ids, const unsigned* n_ids_dev, int n_ids_max, int32_t* out, hipStream_t st) {
  if (k < 1 || k > 256) return hipErrorInvalidValue;
  if (n_ids_max <= 0) return hipSuccess;
  QueryArgs a;
  a.hint_pts = nullptr;
  a.pt_begin = gb.pt_begin;
  a.grid = gb.shape;
  a.cell_base = gb.cell_base;
  a.pt_inst = gb.pt_inst;
  a.q_inst = nullptr;
  a.sorted_pts = gb.sorted_pts;
  a.sorted_ids = gb.sorted_ids;
  a.start = gb.start;
  a.n_ref = 0;
  a.offsets = offsets;
  a.n_offsets = n_offsets;
  a.k = k;
  a.n_query = n_ids_max;
  a.mode = 3;
  a.verts = nullptr;
  a.tets = nullptr;
  a.qpts = pts;
  a.out = out;
  a.out_pitch = 0;
  a.q_ids = ids;
  a.n_query_dev = n_ids_dev;
  return launch_query(a, n_ids_max, kQueryWavesDefault, st);
}

}  // namespace vd
