// Internal interface of the exact kNN kernels (vd_knn.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace vd {

// Chebyshev radius of the offset table; cells farther away are reached by a
// full scan, which only sparse neighbourhoods need.
constexpr int kOffsetRadius = 12;

struct GridShape {
  float lo[3];
  float h, inv_h, slack;
  int dim[3];
  int ncell;
};

// Device buffers of a grid over one point set, or over a batch of n_inst
// independent point sets (instances) that share the buffers: instance i owns
// the points pt_begin[i] .. pt_begin[i+1] and the cells cell_base[i] ..
// cell_base[i+1], and its grid is shape[i].  A query only sees the points of
// its own instance.  With n_inst == 1 the three tables may be null (one
// instance owning all points and max_cells cells).
struct GridBufs {
  GridShape* shape = nullptr;      // n_inst
  int n_inst = 1;
  const int* cell_base = nullptr;  // n_inst + 1, or null
  const int* pt_begin = nullptr;   // n_inst + 1, or null
  const int* pt_inst = nullptr;    // instance of every point, or null
  const float* box = nullptr;      // n_inst x (lo xyz, hi xyz): box the grid must cover besides its points, or null
  int max_cells = 0;               // total cells (cell_base[n_inst])
  int* count = nullptr;            // max_cells
  int* start = nullptr;            // max_cells + 1
  int* cursor = nullptr;           // max_cells
  int* cell_of = nullptr;          // n
  int* sorted_ids = nullptr;       // n
  float4* sorted_pts = nullptr;    // n
  int* scan_tmp = nullptr;         // scan_tmp_ints(max_cells)
};

// Exclusive scan of in[0..n) into out[0..n] (out[n] = total), n = *n_dev when
// n_dev is given (capped at n_max), else n_max; clears cursor[0..n) if given.
// tmp holds scan_tmp_ints(n_max) ints.
int scan_tmp_ints(int n_max);
hipError_t scan_excl(const int* in, const int* n_dev, int n_max, int* out, int* cursor, int* tmp, hipStream_t st);
// The same with unsigned sums (totals up to 2^32 - 1).
hipError_t scan_excl(const int* in, const int* n_dev, int n_max, unsigned* out, int* cursor, unsigned* tmp,
                     hipStream_t st);

std::vector<int4> knn_offset_table();
int grid_max_cells(int n, float density);

// Fills (on the device) a buffer of n 32-bit words with v.
hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t st);

// Bins n points (x,y,z,_) into a grid over their bounding box, extended by the
// box [extra_lo, extra_hi] when given (single instance) or by gb.box (per
// instance).  Deterministic.
hipError_t grid_build(const GridBufs& gb, const float4* pts, int n, const float* extra_lo, const float* extra_hi,
                      float density, hipStream_t st);

// k nearest binned points of every query, nearest first, (d2, index) order.
//   mode 0: queries are the binned points, in binned order, self excluded;
//           row = original index.
//   mode 1: queries are the centroids of tets[0..n_query).
//   mode 2: queries are qpts[0..n_query).
// out: row-major n_query x k (out_pitch == 0) or out[j*out_pitch + q].
// Missing neighbours (k larger than the point count) are -1.  q_inst: the
// instance of every query of mode 1 or 2 (null: instance 0).  hint_pts
// (modes 0 and 1, row-major out): out already holds a complete result of this
// problem for earlier positions of the same points (k distinct ids of the
// query's instance per row, itself excluded), and hint_pts holds the points by
// original id; the query's k-th distance is then bounded from the start by
// those k points' current distances.  The result does not depend on it.
// waves: wavefronts per workgroup, 2, 4 or 8 (other values: 4); the result
// does not depend on it either.
hipError_t knn_query(const GridBufs& gb, int n_ref, const int4* offsets, int n_offsets, int k, int mode,
                     int n_query, const float4* verts, const int4* tets, const float4* qpts, const int* q_inst,
                     int32_t* out, long long out_pitch, hipStream_t st, const float4* hint_pts = nullptr,
                     int waves = 4);

// k nearest binned points of the points pts[ids[q]], q < min(*n_ids_dev,
// n_ids_max), each excluding itself; row q of out (row-major, k per row).
hipError_t knn_query_ids(const GridBufs& gb, const int4* offsets, int n_offsets, int k, const float4* pts,
                         const int* ids, const unsigned* n_ids_dev, int n_ids_max, int32_t* out, hipStream_t st);

}  // namespace vd
