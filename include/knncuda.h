/* The reference's device kNN entry point (extern/kNN-CUDA/knncuda.h:1-12 of
 * xh-liu-tech/3D-Voronoi-GPU) with the same signature, as a header-only shim
 * over vd_knn_dev (vd_api.h): exact k nearest reference points of every query,
 * ties to the lower reference index, index[j * index_pitch + q].  Returns
 * false on error (the reference's convention); dim must be 3. */
#pragma once

#include <cstddef>

#include "vd_api.h"

inline bool knn_cuda_global_dev(const float* ref_dev, const int ref_nb, const size_t ref_pitch,
                                const float* query_dev, const int query_nb, const size_t query_pitch, const int dim,
                                const int k, int* knn_index_dev, const size_t index_pitch) {
  return vd_knn_dev(ref_dev, ref_nb, ref_pitch, query_dev, query_nb, query_pitch, dim, k, knn_index_dev, index_pitch,
                    nullptr) == VD_OK;
}
