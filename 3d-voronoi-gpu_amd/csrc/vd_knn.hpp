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

// Device buffers of one grid (all sized by the caller).
struct GridBufs {
  GridShape* shape;   // 1
  int max_cells;
  int* count;         // max_cells
  int* start;         // max_cells + 1
  int* cursor;        // max_cells
  int* cell_of;       // n
  int* sorted_ids;    // n
  float4* sorted_pts; // n
  int* scan_tmp;      // scan_tmp_ints(max_cells)
};

// Exclusive scan of in[0..n) into out[0..n] (out[n] = total), n = *n_dev when
// n_dev is given (capped at n_max), else n_max; clears cursor[0..n) if given.
// tmp holds scan_tmp_ints(n_max) ints.
int scan_tmp_ints(int n_max);
hipError_t scan_excl(const int* in, const int* n_dev, int n_max, int* out, int* cursor, int* tmp, hipStream_t st);

std::vector<int4> knn_offset_table();
int grid_max_cells(int n, float density);

// Fills (on the device) a buffer of n 32-bit words with v.
hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t st);

// Bins n points (x,y,z,_) into a grid over their bounding box, extended by the
// box [extra_lo, extra_hi] when given.  Deterministic.
hipError_t grid_build(const GridBufs& gb, const float4* pts, int n, const float* extra_lo, const float* extra_hi,
                      float density, hipStream_t st);

// k nearest binned points of every query, nearest first, (d2, index) order.
//   mode 0: queries are the binned points, in binned order, self excluded;
//           row = original index.
//   mode 1: queries are the centroids of tets[0..n_query).
//   mode 2: queries are qpts[0..n_query).
// out: row-major n_query x k (out_pitch == 0) or out[j*out_pitch + q].
// Missing neighbours (k larger than the point count) are -1.
hipError_t knn_query(const GridBufs& gb, int n_ref, const int4* offsets, int n_offsets, int k, int mode,
                     int n_query, const float4* verts, const int4* tets, const float4* qpts, int32_t* out,
                     long long out_pitch, hipStream_t st);

}  // namespace vd
