// Readers and writers of the reference's file formats, host C++ only (no HIP),
// shared by the VolumeVoronoiGPU command line and the loader test program.
//
//   .tet : "<n> vertices", "<m> cells", n lines "x y z", m lines "4 a b c d"
//          (0-based)                                   test_voronoi.cu:55-73
//   .vtk : 4 header lines, "POINTS n <type>", coordinates, "CELLS m size",
//          m lines "4 a b c d" (1-based)               test_voronoi.cu:74-93
//   .xyz : count line, then "x y z" per site           test_voronoi.cu:135-147
//          written as "x  y  z", 6 significant digits  test_voronoi.cu:123-133
//
// A loaded mesh is mapped into [0,1000]^3: every coordinate becomes
// 1000 * (x - lo) / side, in FP32 and in that order, where lo is the bounding
// box grown by 0.1 % of its largest side and side that box's largest side
// (test_voronoi.cu:14-38, 103-118).  For the unit cube that is
// 1000 * (x + 0.001) / 1.002.
#ifndef VD_MESH_IO_HPP
#define VD_MESH_IO_HPP

#include <algorithm>
#include <fstream>
#include <ostream>
#include <string>
#include <vector>

namespace vdio {

// Bounding box grown by 0.1 % of its largest side (test_voronoi.cu:14-38).
inline void grown_bbox(const std::vector<float>& v, float lo[3], float hi[3]) {
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

// Maps the vertices into [0,1000]^3 in place (test_voronoi.cu:103-118).
inline void normalize(std::vector<float>& verts) {
  if (verts.size() < 3) return;
  float lo[3], hi[3];
  grown_bbox(verts, lo, hi);
  const float side = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  for (size_t i = 0; i < verts.size() / 3; ++i)
    for (int a = 0; a < 3; ++a) verts[3 * i + a] = 1000.f * (verts[3 * i + a] - lo[a]) / side;
}

// Reads a .tet or .vtk mesh (by extension); tets come out 0-based.  The
// vertices are not normalised here.  False on a missing file, an unknown
// extension or a malformed record.
inline bool read_mesh(const std::string& path, std::vector<float>& verts, std::vector<int>& tets) {
  std::ifstream in(path);
  if (!in) return false;
  const size_t dot = path.find_last_of('.');
  const std::string ext = dot == std::string::npos ? std::string() : path.substr(dot + 1);
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
  if (nt <= 0 || !in) return false;
  tets.resize(4 * (size_t)nt);
  const int base = ext == "vtk" ? 1 : 0;
  for (int i = 0; i < nt; ++i) {
    in >> four >> tets[4 * i] >> tets[4 * i + 1] >> tets[4 * i + 2] >> tets[4 * i + 3];
    if (four != 4 || !in) return false;
    for (int j = 0; j < 4; ++j) tets[4 * i + j] -= base;
  }
  return true;
}

// Sites: count line then "x y z"; stored SoA (x row, y row, z row) when
// `soa`, else AoS.
inline bool read_sites(const std::string& path, bool soa, std::vector<float>& s, int& n) {
  std::ifstream in(path);
  if (!in || !(in >> n) || n <= 0) return false;
  s.resize(3 * (size_t)n);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) in >> (soa ? s[i + (size_t)a * n] : s[3 * i + a]);
  return (bool)in;
}

// "x  y  z" with the stream's default 6 significant digits.
inline void write_sites(std::ostream& out, bool soa, const std::vector<float>& s, int n) {
  out << n << std::endl;
  for (int i = 0; i < n; ++i) {
    if (soa)
      out << s[i] << "  " << s[i + (size_t)n] << "  " << s[i + 2 * (size_t)n] << std::endl;
    else
      out << s[3 * i] << "  " << s[3 * i + 1] << "  " << s[3 * i + 2] << std::endl;
  }
}

}  // namespace vdio

#endif
