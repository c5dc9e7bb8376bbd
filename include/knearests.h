/* The reference's grid-kNN interface (knearests.h:1-27 of
 * xh-liu-tech/3D-Voronoi-GPU): same struct layout, function names and
 * arguments, as header-only C++ shims over the C ABI of libvoronoi_hip.so
 * (vd_kn_* in vd_api.h).
 *
 * Differences, by design:
 *   - neighbour lists are exact, (d2, index)-ordered, with no ring limit;
 *   - the grid order of d_stored_points is deterministic (cell, then index);
 *   - d_cell_offsets, d_cell_offset_dists, d_cell_max and d_globcounter are
 *     not used by this implementation and are NULL;
 *   - errors print a message and exit(EXIT_FAILURE), as the reference does.
 */
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "params.h"
#include "vd_api.h"

typedef struct {
  int allocated_points;        /* number of input points */
  int dimx, dimy, dimz;        /* grid resolution */
  int num_cell_offsets;        /* size of the cell-offset table */
  int* d_cell_offsets;         /* unused (NULL) */
  float* d_cell_offset_dists;  /* unused (NULL) */
  float* d_cell_max;           /* unused (NULL) */
  unsigned int* d_permutation; /* device: grid position -> input index */
  int* d_counters;             /* device: points per cell */
  int* d_ptrs;                 /* device: first grid position of each cell */
  int* d_globcounter;          /* unused (NULL) */
  float3* d_stored_points;     /* device: points in grid order */
  unsigned int* d_knearests;   /* device: allocated_points x _K_, grid positions, nearest first */
} kn_problem;

namespace vd_compat {
/* kn_problem as the first member, so a kn_problem* converts back. */
struct KnHolder {
  kn_problem pub;
  vd_kn* impl;
};
inline void kn_die(const char* what) {
  fprintf(stderr, "%s failed\n", what);
  exit(EXIT_FAILURE);
}
}  // namespace vd_compat

inline kn_problem* kn_prepare(float3* points, int numpoints) {
  vd_compat::KnHolder* h = new vd_compat::KnHolder();
  if (vd_kn_prepare(&h->impl, reinterpret_cast<const float*>(points), numpoints, _K_) != VD_OK)
    vd_compat::kn_die("kn_prepare");
  vd_kn_info i;
  vd_kn_info_get(h->impl, &i);
  kn_problem& p = h->pub;
  p.allocated_points = i.allocated_points;
  p.dimx = i.dimx;
  p.dimy = i.dimy;
  p.dimz = i.dimz;
  p.num_cell_offsets = i.num_cell_offsets;
  p.d_cell_offsets = nullptr;
  p.d_cell_offset_dists = nullptr;
  p.d_cell_max = nullptr;
  p.d_permutation = i.d_permutation;
  p.d_counters = i.d_counters;
  p.d_ptrs = i.d_ptrs;
  p.d_globcounter = nullptr;
  p.d_stored_points = reinterpret_cast<float3*>(i.d_stored_points);
  p.d_knearests = i.d_knearests;
  return &h->pub;
}

inline void kn_solve(kn_problem* kn) {
  if (vd_kn_solve(reinterpret_cast<vd_compat::KnHolder*>(kn)->impl) != VD_OK) vd_compat::kn_die("kn_solve");
}

inline void kn_free(kn_problem** kn) {
  vd_compat::KnHolder* h = reinterpret_cast<vd_compat::KnHolder*>(*kn);
  vd_kn_free(h->impl);
  delete h;
  *kn = nullptr;
}

/* Points per cell: min, max, average, total, then the histogram
 * (knearests.cu:434-460), on stderr. */
inline void kn_print_stats(kn_problem* kn) {
  const size_t cells = (size_t)kn->dimx * kn->dimy * kn->dimz;
  std::vector<int> counters(cells);
  if (vd_kn_get_cell_counts(reinterpret_cast<vd_compat::KnHolder*>(kn)->impl, counters.data()) != VD_OK)
    vd_compat::kn_die("kn_print_stats");
  long long tot = 0;
  int cmin = INT_MAX, cmax = 0;
  std::map<int, int> histo;
  for (int c : counters) {
    histo[c]++;
    cmin = std::min(cmin, c);
    cmax = std::max(cmax, c);
    tot += c;
  }
  fprintf(stderr, "Grid:  points per cell: %d (min), %d (max), %g avg, total %lld\n", cmin, cmax,
          kn->allocated_points / (double)cells, tot);
  for (const auto& h : histo) fprintf(stderr, "[%d] => %d\n", h.first, h.second);
}

/* Host copies; the caller free()s them (knearests.cu:404-432). */
inline float3* kn_get_points(kn_problem* kn) {
  float3* out = static_cast<float3*>(malloc(sizeof(float3) * kn->allocated_points));
  if (vd_kn_get_points(reinterpret_cast<vd_compat::KnHolder*>(kn)->impl, reinterpret_cast<float*>(out)) != VD_OK)
    vd_compat::kn_die("kn_get_points");
  return out;
}

inline unsigned int* kn_get_knearests(kn_problem* kn) {
  unsigned int* out = static_cast<unsigned int*>(malloc(sizeof(unsigned int) * (size_t)kn->allocated_points * _K_));
  if (vd_kn_get_knearests(reinterpret_cast<vd_compat::KnHolder*>(kn)->impl, out) != VD_OK)
    vd_compat::kn_die("kn_get_knearests");
  return out;
}

inline unsigned int* kn_get_permutation(kn_problem* kn) {
  unsigned int* out = static_cast<unsigned int*>(malloc(sizeof(unsigned int) * kn->allocated_points));
  if (vd_kn_get_permutation(reinterpret_cast<vd_compat::KnHolder*>(kn)->impl, out) != VD_OK)
    vd_compat::kn_die("kn_get_permutation");
  return out;
}
