// Test program (host only): runs the command line's loader on a mesh and a
// site file and writes what it read as raw little-endian arrays, so that the
// tests can pin the loader against the closed form of the [0,1000]^3 mapping.
//   loader_dump <mesh.tet|.vtk> <sites.xyz> <out_dir>
// writes <out_dir>/verts.f32 (normalised, n x 3), tets.i32 (0-based, m x 4),
// sites.f32 (n_sites x 3, as read), sites.txt (written back in the .xyz format).
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "mesh_io.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <mesh> <sites.xyz> <out_dir>\n", argv[0]);
    return 2;
  }
  std::vector<float> verts;
  std::vector<int> tets;
  if (!vdio::read_mesh(argv[1], verts, tets)) return 1;
  vdio::normalize(verts);
  std::vector<float> sites;
  int n = 0;
  if (!vdio::read_sites(argv[2], false, sites, n)) return 1;
  const std::string dir = argv[3];
  std::ofstream(dir + "/verts.f32", std::ios::binary).write((const char*)verts.data(), verts.size() * 4);
  std::ofstream(dir + "/tets.i32", std::ios::binary).write((const char*)tets.data(), tets.size() * 4);
  std::ofstream(dir + "/sites.f32", std::ios::binary).write((const char*)sites.data(), sites.size() * 4);
  std::ofstream txt(dir + "/sites.txt");
  vdio::write_sites(txt, false, sites, n);
  return 0;
}
