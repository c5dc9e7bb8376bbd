// VolumeVoronoiGPU <tet_mesh.tet/vtk> <sites.xyz> <nb_iter> <k (optional)>
//
// The reference's command line (test_voronoi.cu:190-227): reads a tet mesh
// and maps it into [0,1000]^3, reads the sites (already in that frame), runs
// nb_iter Lloyd iterations with k candidate sites per tet (0 = heuristic), and
// writes out.xyz, iter_<i>.xyz and record.csv in the working directory.
// Environment: VD_FORM=site|tet selects the tet -> cell formulation (default
// "cells", the reference's cell -> tet source of the new sites), VD_ACCUM=f64
// the FP64-accumulated form, VD_MOMENTS=robust the well-conditioned moment
// decomposition.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mesh_io.hpp"
#include "params.h"
#include "voronoi.h"

namespace {

// Reads the mesh and maps it into [0,1000]^3 (test_voronoi.cu:40-121).
bool load_mesh(const std::string& path, std::vector<float>& verts, std::vector<int>& tets) {
  if (!vdio::read_mesh(path, verts, tets)) return false;
  vdio::normalize(verts);
  float lo[3], hi[3];
  vdio::grown_bbox(verts, lo, hi);
  std::cerr << "bbox [" << lo[0] << ":" << hi[0] << "], [" << lo[1] << ":" << hi[1] << "], [" << lo[2] << ":"
            << hi[2] << "]" << std::endl;
  return true;
}

void print_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  printf("HIP devices: %d\n", n);
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, i) != hipSuccess) continue;
    printf("Device %d: %s (%s), %d CUs, %.1f GB, LDS/block %zu B, wavefront %d\n", i, p.name, p.gcnArchName,
           p.multiProcessorCount, p.totalGlobalMem / 1e9, (size_t)p.sharedMemPerBlock, p.warpSize);
  }
}

}  // namespace

int main(int argc, char** argv) {
  print_device();
  if (argc < 4) {
    std::cerr << "Usage: " << argv[0] << " <tet_mesh.tet/vtk> <sites_file.xyz> <nb_iter> <k (optional)> (e.g.: "
              << argv[0] << " ../data/joint.tet ../data/joint.xyz 120 0)" << std::endl;
    return 1;
  }
  std::vector<float> verts;
  std::vector<int> tets;
  if (!load_mesh(argv[1], verts, tets)) {
    std::cerr << argv[1] << ": could not load file" << std::endl;
    return 1;
  }
  // SoA when fewer than KNEARESTS_MIN_N sites, as the reference stores them
  // (test_voronoi.cu:135-147)
  int n_site = 0;
  {
    std::ifstream probe(argv[2]);
    if (!(probe >> n_site)) n_site = 0;
  }
  const bool soa = n_site < KNEARESTS_MIN_N;
  std::vector<float> site;
  if (!vdio::read_sites(argv[2], soa, site, n_site)) {
    std::cerr << argv[2] << ": could not load file" << std::endl;
    return 1;
  }
  // like the reference, k is read only when it is the 4th and last argument
  if (argc == 5)
    compute_clipped_voro_diagram_GPU(verts, tets, site, n_site, soa, atoi(argv[3]), atoi(argv[4]));
  else
    compute_clipped_voro_diagram_GPU(verts, tets, site, n_site, soa, atoi(argv[3]));
  std::ofstream out("out.xyz");
  vdio::write_sites(out, soa, site, n_site);
  return 0;
}
