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

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "params.h"
#include "voronoi.h"

namespace {

// Bounding box grown by 0.1% of its largest side (test_voronoi.cu:14-38).
void grown_bbox(const std::vector<float>& v, float lo[3], float hi[3]) {
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = v[a];
  for (size_t i = 1; i < v.size() / 3; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], v[3 * i + a]);
      hi[a] = std::max(hi[a], v[3 * i + a]);
    }
  float d = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  d = 0.001f * d;
  for (int a = 0; a < 3; ++a) {
    lo[a] -= d;
    hi[a] += d;
  }
}

// .tet: "<n> vertices", "<m> cells", coordinates, "4 a b c d" (0-based);
// .vtk: 4 header lines, "POINTS n type", coordinates, "CELLS m size",
// "4 a b c d" (1-based).  Vertices are mapped to [0,1000]^3
// (test_voronoi.cu:40-121).
bool load_mesh(const std::string& path, std::vector<float>& verts, std::vector<int>& tets) {
  std::ifstream in(path);
  if (!in) return false;
  const std::string ext = path.substr(path.find_last_of('.') + 1);
  std::string line, word;
  int nv = 0, nt = 0, four = 0;
  if (ext == "tet") {
    in >> nv;
    std::getline(in, line);
    in >> nt;
    std::getline(in, line);
  } else if (ext == "vtk") {
    for (int i = 0; i < 4; ++i) std::getline(in, line);
    in >> word >> nv >> word;
  } else {
    return false;
  }
  if (nv <= 0) return false;
  verts.resize(3 * (size_t)nv);
  for (int i = 0; i < nv; ++i) in >> verts[3 * i] >> verts[3 * i + 1] >> verts[3 * i + 2];
  if (ext == "vtk") in >> word >> nt >> word;
  if (nt <= 0) return false;
  tets.resize(4 * (size_t)nt);
  const int base = ext == "vtk" ? 1 : 0;
  for (int i = 0; i < nt; ++i) {
    in >> four >> tets[4 * i] >> tets[4 * i + 1] >> tets[4 * i + 2] >> tets[4 * i + 3];
    if (four != 4 || !in) return false;
    for (int j = 0; j < 4; ++j) tets[4 * i + j] -= base;
  }
  float lo[3], hi[3];
  grown_bbox(verts, lo, hi);
  const float side = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) verts[3 * i + a] = 1000.f * (verts[3 * i + a] - lo[a]) / side;
  grown_bbox(verts, lo, hi);
  std::cerr << "bbox [" << lo[0] << ":" << hi[0] << "], [" << lo[1] << ":" << hi[1] << "], [" << lo[2] << ":"
            << hi[2] << "]" << std::endl;
  return true;
}

// Count line then "x y z"; SoA when fewer than KNEARESTS_MIN_N sites, as the
// reference stores them (test_voronoi.cu:135-147).
bool load_sites(const char* path, bool& soa, std::vector<float>& s, int& n) {
  std::ifstream in(path);
  if (!in || !(in >> n) || n <= 0) return false;
  soa = n < KNEARESTS_MIN_N;
  s.resize(3 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) in >> (soa ? s[i + (size_t)a * n] : s[3 * i + a]);
  return (bool)in;
}

// "x  y  z" with the stream's default 6 significant digits
// (test_voronoi.cu:123-133).
void write_sites(const char* path, bool soa, const std::vector<float>& s, int n) {
  std::ofstream out(path);
  out << n << std::endl;
  for (int i = 0; i < n; ++i) {
    if (soa)
      out << s[i] << "  " << s[i + n] << "  " << s[i + 2 * (size_t)n] << std::endl;
    else
      out << s[3 * i] << "  " << s[3 * i + 1] << "  " << s[3 * i + 2] << std::endl;
  }
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
  bool soa = false;
  int n_site = 0;
  std::vector<float> site;
  if (!load_sites(argv[2], soa, site, n_site)) {
    std::cerr << argv[2] << ": could not load file" << std::endl;
    return 1;
  }
  // like the reference, k is read only when it is the 4th and last argument
  if (argc == 5)
    compute_clipped_voro_diagram_GPU(verts, tets, site, n_site, soa, atoi(argv[3]), atoi(argv[4]));
  else
    compute_clipped_voro_diagram_GPU(verts, tets, site, n_site, soa, atoi(argv[3]));
  write_sites("out.xyz", soa, site, n_site);
  return 0;
}
