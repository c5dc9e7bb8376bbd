// CPU oracle for the clipped-Voronoi Lloyd path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++17, the algorithm
// of the reference CUDA program (xh-liu-tech/3D-Voronoi-GPU) so that the HIP
// path can be checked against it. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never links it.
//
// Arithmetic convention (shared with the HIP kernels so that FP32 results can
// be compared bit for bit): every expression is written with explicit fma()
// where a contracting compiler would fuse, and the file is compiled with
// -ffp-contract=off so that no other fusion happens.  The reference's own
// contraction choices (nvcc) cannot be observed without nvcc; this is the
// build's canonical choice.
//
// Functions and the reference lines they follow:
//   det2/det3/det4, tri2plane, project ..... voronoi.cu:14-123
//   Poly (ConvexCell) ctors ................ voronoi.cu:125-217
//   Poly::security_reached ................. voronoi.cu:219-232
//   Poly::vertex ........................... voronoi.cu:238-249
//   Poly::conflict ......................... voronoi.cu:274-306
//   Poly::add_plane / add_triangle ......... voronoi.cu:356-389
//   Poly::cavity_boundary .................. voronoi.cu:391-441
//   Poly::clip ............................. voronoi.cu:443-530
//   pair_moments ........................... voronoi.cu:532-537,577-618
//   clip_pair_tet (kernel A) ............... voronoi.cu:662-701
//   site_cell / clip_pair_cell (B, C) ...... voronoi.cu:540-575,620-660,703-726
//   lloyd_update ........................... voronoi.cu:728-749
//   tet_centroid ........................... voronoi.cu:751-779
//   knn (exact, (d2,id) order) ............. knncuda.cu:20-91,143-189; knearests.cu:90-145
//   tet_k_heuristic ........................ voronoi.cu:936
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <limits>
#include <vector>

namespace vdo {

// Per-pair status codes; values are those of voronoi_defs.h:4-12.
enum Status : int32_t {
  triangle_overflow = 0,
  vertex_overflow = 1,
  inconsistent_boundary = 2,
  security_radius_not_reached = 3,
  success = 4,
  needs_exact_predicates = 5,
  no_intersection = 6
};

constexpr uint8_t kNone = 255;  // end-of-list marker for plane ids (voronoi.h:30)

template <class R>
struct V4 {
  R x, y, z, w;
};

// a*b - c*d with the product a*b fused.
template <class R>
inline R msub(R a, R b, R c, R d) {
  return std::fma(a, b, -(c * d));
}
template <class R>
inline R dot3(const V4<R>& a, const V4<R>& b) {
  return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x));
}
template <class R>
inline V4<R> sub3(const V4<R>& a, const V4<R>& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z, R(0)};
}
template <class R>
inline V4<R> cross3(const V4<R>& a, const V4<R>& b) {
  return {msub(a.y, b.z, a.z, b.y), msub(a.z, b.x, a.x, b.z), msub(a.x, b.y, a.y, b.x), R(0)};
}
// voronoi.cu:47-52: expansion along the first column.
template <class R>
inline R det2(R a11, R a12, R a21, R a22) {
  return msub(a11, a22, a12, a21);
}
template <class R>
inline R det3(R a11, R a12, R a13, R a21, R a22, R a23, R a31, R a32, R a33) {
  const R c1 = det2(a22, a23, a32, a33);
  const R c2 = det2(a12, a13, a32, a33);
  const R c3 = det2(a12, a13, a22, a23);
  return std::fma(a31, c3, msub(a11, c1, a21, c2));
}

// The four 3x3 minors of the planes (p1,p2,p3) that the reference's det4x4
// (voronoi.cu:54-73) builds before its last row; they depend on the vertex only,
// so a caller may keep them and finish the 4x4 determinant with conflict_det().
template <class R>
struct Minors {
  R m123, m124, m134, m234;
};
template <class R>
inline Minors<R> vertex_minors(const V4<R>& p1, const V4<R>& p2, const V4<R>& p3) {
  // columns (p1,p2,p3,eqn), rows (x,y,z,w) -> a11=p1.x a12=p2.x a13=p3.x ...
  const R m12 = msub(p1.y, p2.x, p1.x, p2.y);
  const R m13 = msub(p1.z, p2.x, p1.x, p2.z);
  const R m14 = msub(p1.w, p2.x, p1.x, p2.w);
  const R m23 = msub(p1.z, p2.y, p1.y, p2.z);
  const R m24 = msub(p1.w, p2.y, p1.y, p2.w);
  const R m34 = msub(p1.w, p2.z, p1.z, p2.w);
  Minors<R> m;
  m.m123 = std::fma(m12, p3.z, msub(m23, p3.x, m13, p3.y));
  m.m124 = std::fma(m12, p3.w, msub(m24, p3.x, m14, p3.y));
  m.m134 = std::fma(m13, p3.w, msub(m34, p3.x, m14, p3.z));
  m.m234 = std::fma(m23, p3.w, msub(m34, p3.y, m24, p3.z));
  return m;
}
template <class R>
inline R conflict_det(const Minors<R>& m, const V4<R>& e) {
  return std::fma(-m.m123, e.w, std::fma(m.m124, e.z, msub(m.m234, e.x, m.m134, e.y)));
}

// Plane through v0,v1,v2 with normal (v1-v0)x(v2-v0) (voronoi.cu:119-123, 41-46).
template <class R>
inline V4<R> tri2plane(const V4<R>& v0, const V4<R>& v1, const V4<R>& v2) {
  V4<R> n = cross3(sub3(v1, v0), sub3(v2, v0));
  n.w = -dot3(v0, n);
  return n;
}

// Orthogonal projection of P on a plane; no-op for near-degenerate normals
// (voronoi.cu:111-116).  The 1e-2 literal is a double in the reference, so the
// comparison is done in double.
template <class R>
inline V4<R> project(const V4<R>& P, const V4<R>& pl) {
  const R n2 = dot3(pl, pl);
  R lambda = R(0);
  if ((double)n2 > 1e-2) lambda = (dot3(pl, P) + pl.w) / n2;
  return {std::fma(-lambda, pl.x, P.x), std::fma(-lambda, pl.y, P.y), std::fma(-lambda, pl.z, P.z), R(1)};
}

// Signed volume and barycenter of tet (A,B,C,D) (voronoi.cu:104-110).  The
// reference divides by the double literal 6.; for float inputs the result
// equals the float quotient (double rounding of a quotient is innocuous when
// the wide format has >= 2p+2 bits).
template <class R>
inline void tet_vol_bary(const V4<R>& A, const V4<R>& B, const V4<R>& C, const V4<R>& D, R& vol, V4<R>& bary) {
  const V4<R> a = sub3(A, D), b = sub3(B, D), c = sub3(C, D);
  vol = -det3(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z) / R(6);
  bary.x = R(0.25) * (((A.x + B.x) + C.x) + D.x);
  bary.y = R(0.25) * (((A.y + B.y) + C.y) + D.y);
  bary.z = R(0.25) * (((A.z + B.z) + C.z) + D.z);
  bary.w = R(1);
}

struct Caps {
  int pcap = 64;  // _MAX_P_ (params.h:11)
  int tcap = 96;  // _MAX_T_ (params.h:12)
  // arithmetic filter of the conflict test (USE_ARITHMETIC_FILTER, params.h:32,
  // off in the reference): 0 off; 1 the reference's (flag needs_exact_predicates
  // when |det| is below its bound, voronoi.cu:285-301); 2 re-evaluate such a det
  // in FP64 from the same planes and flag only below the FP64 bound
  // (voronoi.cu:325-341), taking the FP64 sign
  int filter = 0;
};

// Error bounds of the filter: C * max|x| * max|y| * max|z| * max(max|.|)^2 over
// the four planes' x, y, z (voronoi.cu:285-297, 325-337).
inline float filter_bound32(float mx, float my, float mz) {
  float eps = (float)(6.6876506e-05 * (double)mx * (double)my * (double)mz);
  const float mm = std::max(std::max(mx, my), mz);
  eps *= (mm * mm);
  return eps;
}
inline double filter_bound64(double mx, double my, double mz) {
  double eps = 1.2466136531027298e-13 * mx * my * mz;
  const double mm = std::max(std::max(mx, my), mz);
  eps *= (mm * mm);
  return eps;
}

// Work counters for the FLOP model of the clipping kernel.
struct Events {
  int64_t pairs = 0;          // pairs started
  int64_t clips = 0;          // clip_by_plane calls
  int64_t conflict_tests = 0; // triangle_is_in_conflict evaluations
  int64_t secure_evals = 0;   // vertex evaluations inside security checks
  int64_t secure_checks = 0;  // security checks
  int64_t new_tris = 0;       // triangles created by clips
  int64_t acc_tris = 0;       // vertices integrated by the moment routine
  int64_t acc_pairs = 0;      // pairs that reached the moment routine
  int64_t max_nv = 0, max_nt = 0;  // largest plane / vertex count seen
  void add(const Events& o) {
    pairs += o.pairs; clips += o.clips; conflict_tests += o.conflict_tests;
    secure_evals += o.secure_evals; secure_checks += o.secure_checks; new_tris += o.new_tris;
    acc_tris += o.acc_tris; acc_pairs += o.acc_pairs;
    max_nv = std::max(max_nv, o.max_nv); max_nt = std::max(max_nt, o.max_nt);
  }
};

// Convex polytope in dual form: planes pl[0..nv) and vertices as oriented
// plane triples tr[0..nt) (voronoi.h:32-76).
template <class R>
struct Poly {
  static constexpr int PMAX = 255;
  static constexpr int TMAX = 1024;
  V4<R> pl[PMAX];
  uint8_t tr[TMAX][3];
  uint8_t nxt[PMAX];
  int nv = 0, nt = 0, nr = 0;
  int peak_nv = 0, peak_nt = 0;  // largest counts reached by the current pair
  uint8_t first = kNone;
  int32_t status = success;
  V4<R> seed{};
  Caps caps;
  Events* ev = nullptr;

  void set_tri(int t, int a, int b, int c) {
    tr[t][0] = (uint8_t)a; tr[t][1] = (uint8_t)b; tr[t][2] = (uint8_t)c;
  }
  void reset_links() {
    for (int i = 0; i < caps.pcap; ++i) nxt[i] = kNone;
    first = kNone;
  }

  // Tet start (voronoi.cu:163-201): planes of faces cbd, acd, bad, abc.
  void init_tet(const V4<R>& s, const V4<R> v[4]) {
    static const int face[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
    reset_links();
    seed = s;
    status = success;
    for (int i = 0; i < 4; ++i) pl[i] = tri2plane(v[face[i][0]], v[face[i][1]], v[face[i][2]]);
    nv = 4;
    set_tri(0, 1, 3, 2); set_tri(1, 0, 1, 2); set_tri(2, 0, 3, 1); set_tri(3, 0, 2, 3);
    nt = 4;
    peak_nv = nv; peak_nt = nt;
  }

  // Box start [-0.1, 1000.1]^3 (voronoi.cu:125-161).
  void init_box(const V4<R>& s) {
    const R lo = R(-0.1f), hi = R(1000.f + 0.1f);
    reset_links();
    seed = s;
    status = success;
    pl[0] = {R(1), R(0), R(0), -lo};
    pl[1] = {R(-1), R(0), R(0), hi};
    pl[2] = {R(0), R(1), R(0), -lo};
    pl[3] = {R(0), R(-1), R(0), hi};
    pl[4] = {R(0), R(0), R(1), -lo};
    pl[5] = {R(0), R(0), R(-1), hi};
    nv = 6;
    set_tri(0, 2, 5, 0); set_tri(1, 5, 3, 0); set_tri(2, 1, 5, 2); set_tri(3, 5, 1, 3);
    set_tri(4, 4, 2, 0); set_tri(5, 4, 0, 3); set_tri(6, 2, 4, 1); set_tri(7, 4, 3, 1);
    nt = 8;
    peak_nv = nv; peak_nt = nt;
  }

  // Vertex of triangle t: intersection of its three planes (voronoi.cu:238-249).
  V4<R> vertex(int t) const {
    const V4<R>& p1 = pl[tr[t][0]];
    const V4<R>& p2 = pl[tr[t][1]];
    const V4<R>& p3 = pl[tr[t][2]];
    const R X = -det3(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z);
    const R Y = -det3(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z);
    const R Z = -det3(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w);
    const R W = det3(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
    return {X / W, Y / W, Z / W, R(1)};
  }

  // Squared distance from the seed to vertex t (one term of voronoi.cu:222-227).
  R vertex_d2(int t) const {
    const V4<R> d = sub3(vertex(t), seed);
    return dot3(d, d);
  }

  // voronoi.cu:219-232: every vertex closer than half the distance to z.
  bool security_reached(const V4<R>& z) {
    R vmax = R(0);
    for (int t = 0; t < nt; ++t) vmax = std::fmax(vertex_d2(t), vmax);
    if (ev) { ev->secure_evals += nt; ev->secure_checks += 1; }
    const V4<R> d = sub3(z, seed);
    return dot3(d, d) > R(4) * vmax;
  }

  // voronoi.cu:274-306: vertex t lies strictly beyond plane e.
  bool conflict(int t, const V4<R>& e) {
    const V4<R>& p1 = pl[tr[t][0]];
    const V4<R>& p2 = pl[tr[t][1]];
    const V4<R>& p3 = pl[tr[t][2]];
    const Minors<R> m = vertex_minors(p1, p2, p3);
    const R det = conflict_det(m, e);
    if (caps.filter) {
      auto mx4 = [](R a, R b, R c, R d) { return std::max(std::max(std::abs(a), std::abs(b)), std::max(std::abs(c), std::abs(d))); };
      const R mx = mx4(p1.x, p2.x, p3.x, e.x), my = mx4(p1.y, p2.y, p3.y, e.y), mz = mx4(p1.z, p2.z, p3.z, e.z);
      const bool fp32 = sizeof(R) == sizeof(float);
      const double eps = fp32 ? (double)filter_bound32((float)mx, (float)my, (float)mz)
                              : filter_bound64((double)mx, (double)my, (double)mz);
      if (std::abs((double)det) < eps) {
        if (caps.filter == 1 || !fp32) {
          status = needs_exact_predicates;
        } else {
          auto up = [](const V4<R>& v) { return V4<double>{(double)v.x, (double)v.y, (double)v.z, (double)v.w}; };
          const double d64 = conflict_det(vertex_minors(up(p1), up(p2), up(p3)), up(e));
          if (std::abs(d64) < filter_bound64((double)mx, (double)my, (double)mz)) status = needs_exact_predicates;
          return d64 > 0.0;
        }
      }
    }
    return det > R(0);
  }

  // Bisector of the seed and site z (voronoi.cu:365-378).
  static V4<R> bisector(const V4<R>& s, const V4<R>& z) {
    const V4<R> dir = sub3(s, z);
    const V4<R> mid2 = {s.x + z.x, s.y + z.y, s.z + z.z, R(0)};
    V4<R> p = dir;
    p.w = -dot3(mid2, dir) / R(2);
    return p;
  }

  int add_plane(const V4<R>& p) {
    if (nv >= caps.pcap) {
      status = vertex_overflow;
      return -1;
    }
    pl[nv] = p;
    return nv++;
  }
  void add_triangle(int a, int b, int c) {
    if (nt + 1 >= caps.tcap) {
      status = triangle_overflow;
      return;
    }
    set_tri(nt, a, b, c);
    ++nt;
    if (ev) ev->new_tris += 1;
  }

  // Boundary loop of the removed triangles tr[nt..nt+nr), built by merging
  // them one at a time into a circular list of planes nxt[] (voronoi.cu:391-441).
  void cavity_boundary() {
    reset_links();
    int guard = 0;
    int cur = nt;
    while (nr > 0) {
      if (guard++ > 65535) {
        status = inconsistent_boundary;
        return;
      }
      const uint8_t e[3] = {tr[cur][0], tr[cur][1], tr[cur][2]};
      bool on[3], opp[3];
      for (int k = 0; k < 3; ++k) on[k] = nxt[e[k]] != kNone;
      for (int k = 0; k < 3; ++k) opp[k] = nxt[e[(k + 1) % 3]] == e[k];
      bool simple = true;
      for (int k = 0; k < 3; ++k)
        if (!opp[k] && !opp[(k + 1) % 3] && on[(k + 1) % 3]) simple = false;
      if (!opp[0] && !opp[1] && !opp[2]) {
        if (first == kNone) {
          for (int k = 0; k < 3; ++k) nxt[e[k]] = e[(k + 1) % 3];
          first = e[0];
        } else {
          simple = false;
        }
      }
      if (!simple) {
        ++cur;
        if (cur == nt + nr) cur = nt;
        continue;
      }
      for (int k = 0; k < 3; ++k)
        if (!opp[k]) nxt[e[k]] = e[(k + 1) % 3];
      for (int k = 0; k < 3; ++k) {
        if (opp[k] && opp[(k + 1) % 3]) {
          const uint8_t v = e[(k + 1) % 3];
          if (first == v) first = nxt[v];
          nxt[v] = kNone;
        }
      }
      std::swap(tr[cur][0], tr[nt + nr - 1][0]);
      std::swap(tr[cur][1], tr[nt + nr - 1][1]);
      std::swap(tr[cur][2], tr[nt + nr - 1][2]);
      cur = nt;
      --nr;
    }
  }

  // Clip by plane p (voronoi.cu:443-486 and 488-530).  `keep_srnr` selects the
  // variant used by kernels B/C, which goes on when the status already carries
  // security_radius_not_reached.
  void clip(const V4<R>& p, bool keep_srnr) {
    if (ev) ev->clips += 1;
    const int cur = add_plane(p);
    if (status == vertex_overflow) return;
    const V4<R> eqn = pl[cur];
    nr = 0;
    if (ev) ev->conflict_tests += nt;
    int i = 0;
    while (i < nt) {
      if (conflict(i, eqn)) {
        --nt;
        std::swap(tr[i][0], tr[nt][0]);
        std::swap(tr[i][1], tr[nt][1]);
        std::swap(tr[i][2], tr[nt][2]);
        ++nr;
      } else {
        ++i;
      }
    }
    if (status == needs_exact_predicates) return;  // voronoi.cu:459-461
    if (nr == 0) {
      --nv;
      return;
    }
    if (nt == 0) {
      status = no_intersection;
      return;
    }
    auto ok = [&]() { return status == success || (keep_srnr && status == security_radius_not_reached); };
    cavity_boundary();
    if (!ok()) return;
    if (first == kNone) return;
    uint8_t c = first;
    int guard = 0;
    do {
      if (c == kNone || guard++ > caps.pcap) {
        status = inconsistent_boundary;
        return;
      }
      if (nxt[c] == kNone) {  // open loop: the boundary walk failed
        status = inconsistent_boundary;
        return;
      }
      add_triangle(cur, c, nxt[c]);
      if (!ok()) return;
      c = nxt[c];
    } while (c != first);
    peak_nv = std::max(peak_nv, nv);
    peak_nt = std::max(peak_nt, nt);
    if (ev) {
      ev->max_nv = std::max<int64_t>(ev->max_nv, nv);
      ev->max_nt = std::max<int64_t>(ev->max_nt, nt);
    }
  }
};

// Volume and first moment of a polytope, decomposed around the seed into six
// tets per vertex (voronoi.cu:532-537, 577-618).  A is the accumulation type.
// Returns false when the piece is dropped (|vol| < 0.1 in the [0,1000] frame).
template <class R, class A>
inline bool pair_moments(const Poly<R>& c, A& vol, A& mx, A& my, A& mz) {
  vol = A(0); mx = A(0); my = A(0); mz = A(0);
  const V4<R> C = c.seed;
  V4<R> P[6];
  for (int t = 0; t < c.nt; ++t) {
    const V4<R> X = c.vertex(t);
    for (int i = 0; i < 3; ++i) P[2 * i] = project(C, c.pl[c.tr[t][i]]);
    for (int i = 0; i < 3; ++i) {
      V4<R> n = cross3(sub3(P[2 * i], C), sub3(P[(2 * i + 2) % 6], C));
      n.w = -dot3(C, n);
      P[2 * i + 1] = project(X, n);
    }
    for (int i = 0; i < 6; ++i) {
      R v;
      V4<R> b;
      tet_vol_bary(P[i], P[(i + 1) % 6], C, X, v, b);
      mx = std::fma((A)v, (A)b.x, mx);
      my = std::fma((A)v, (A)b.y, my);
      mz = std::fma((A)v, (A)b.z, mz);
      vol += (A)v;
    }
  }
  if (c.ev) { c.ev->acc_tris += c.nt; c.ev->acc_pairs += 1; }
  return !(std::fabs((double)vol) < 0.1);
}

// Well-conditioned variant of pair_moments (this build's own formulation, not
// the reference's).  Same six-tet-per-vertex ring (face foot, edge foot, ...)
// but (1) the apex O is the mean of the polytope's vertices instead of the seed,
// and (2) the edge foot is the projection of O on the edge line, whose
// direction is the cross product of the two plane normals.  The reference builds
// the edge foot by projecting on a plane whose normal is the cross product of
// the two face-foot offsets; that normal vanishes when the seed lies on or near
// a face plane, the projection is then skipped and the pair's volume is wrong
// (voronoi.cu:111-116, 536).  Exact for any convex polytope in exact arithmetic.
template <class R>
inline V4<R> foot_on_plane(const V4<R>& P, const V4<R>& pl) {
  const R n2 = dot3(pl, pl);
  const R lambda = (dot3(pl, P) + pl.w) / n2;
  return {std::fma(-lambda, pl.x, P.x), std::fma(-lambda, pl.y, P.y), std::fma(-lambda, pl.z, P.z), R(1)};
}
template <class R>
inline V4<R> foot_on_line(const V4<R>& P, const V4<R>& X, const V4<R>& u) {
  const R lambda = dot3(sub3(P, X), u) / dot3(u, u);
  return {std::fma(lambda, u.x, X.x), std::fma(lambda, u.y, X.y), std::fma(lambda, u.z, X.z), R(1)};
}
template <class R, class A>
inline bool pair_moments_robust(const Poly<R>& c, A& vol, A& mx, A& my, A& mz) {
  vol = A(0); mx = A(0); my = A(0); mz = A(0);
  V4<R> O = {R(0), R(0), R(0), R(1)};
  for (int t = 0; t < c.nt; ++t) {
    const V4<R> X = c.vertex(t);
    O.x += X.x; O.y += X.y; O.z += X.z;
  }
  const R inv = R(1) / R(c.nt > 0 ? c.nt : 1);
  O.x *= inv; O.y *= inv; O.z *= inv;
  V4<R> P[6];
  for (int t = 0; t < c.nt; ++t) {
    const V4<R> X = c.vertex(t);
    for (int i = 0; i < 3; ++i) P[2 * i] = foot_on_plane(O, c.pl[c.tr[t][i]]);
    for (int i = 0; i < 3; ++i)
      P[2 * i + 1] = foot_on_line(O, X, cross3(c.pl[c.tr[t][i]], c.pl[c.tr[t][(i + 1) % 3]]));
    for (int i = 0; i < 6; ++i) {
      R v;
      V4<R> b;
      tet_vol_bary(P[i], P[(i + 1) % 6], O, X, v, b);
      mx = std::fma((A)v, (A)b.x, mx);
      my = std::fma((A)v, (A)b.y, my);
      mz = std::fma((A)v, (A)b.z, mz);
      vol += (A)v;
    }
  }
  if (c.ev) { c.ev->acc_tris += c.nt; c.ev->acc_pairs += 1; }
  return !(std::fabs((double)vol) < 0.1);
}

// The robust formulation of an FP32 polytope, integrated in FP64 (this
// build's own formulation, as the HIP kernels compute it): the topology (planes,
// vertex triples) is the FP32 clipping's; each plane is rebuilt in FP64 from
// the FP32 inputs -- a tet face from the tet's corners, a bisector from its
// FP32 normal n (= s - z) and the seed s as the plane through s - n/2
// (w = -s.n + n.n/2), a box plane as it is -- so that the vertices, volumes
// and moments carry no FP32 rounding of the [0,1000] frame.  cell_nv < 0: a
// tet -> cell polytope (ids 0-3 the faces, later ids bisectors); cell_nv >= 0:
// a cell clipped by a tet (ids 0-5 the box, ids below cell_nv bisectors, later
// ids the tet's faces, recognised by their FP32 planes).
template <class R>
inline V4<double> up64(const V4<R>& q) {
  return {(double)q.x, (double)q.y, (double)q.z, (double)q.w};
}
template <class R>
inline V4<double> plane64(const Poly<R>& c, int id, const V4<R> tv[4], int cell_nv) {
  static const int face[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  const V4<R>& p = c.pl[id];
  auto bis = [&]() {
    V4<double> n = {(double)p.x, (double)p.y, (double)p.z, 0.0};
    n.w = std::fma(0.5, dot3(n, n), -dot3(up64(c.seed), n));
    return n;
  };
  int f = id;
  if (cell_nv < 0) {
    if (id >= 4) return bis();
  } else {
    if (id < 6) return up64(p);
    if (id < cell_nv) return bis();
    f = -1;
    for (int k = 0; k < 4 && f < 0; ++k) {
      const V4<R> q = tri2plane(tv[face[k][0]], tv[face[k][1]], tv[face[k][2]]);
      if (std::memcmp(&q, &p, sizeof q) == 0) f = k;
    }
    if (f < 0) return up64(p);
  }
  return tri2plane(up64(tv[face[f][0]]), up64(tv[face[f][1]]), up64(tv[face[f][2]]));
}
inline V4<double> vertex3d(const V4<double>& p1, const V4<double>& p2, const V4<double>& p3) {
  const double X = -det3(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z);
  const double Y = -det3(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z);
  const double Z = -det3(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w);
  const double r = 1.0 / det3(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
  return {X * r, Y * r, Z * r, 1.0};
}
template <class R, class A>
inline bool pair_moments_robust64(const Poly<R>& c, const V4<R> tv[4], int cell_nv, A& vol, A& mx, A& my, A& mz) {
  // apex: the seed (the signed ring is exact for any apex, and in FP64 a
  // seed outside the polytope costs nothing measurable)
  const V4<double> O = up64(c.seed);
  double sv = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;  // 6 x volume, 24 x moments
  V4<double> P[6], a[6];
  for (int t = 0; t < c.nt; ++t) {
    const V4<double> q[3] = {plane64(c, c.tr[t][0], tv, cell_nv), plane64(c, c.tr[t][1], tv, cell_nv),
                             plane64(c, c.tr[t][2], tv, cell_nv)};
    const V4<double> X = vertex3d(q[0], q[1], q[2]);
    for (int i = 0; i < 3; ++i) P[2 * i] = foot_on_plane(O, q[i]);
    for (int i = 0; i < 3; ++i) P[2 * i + 1] = foot_on_line(O, X, cross3(q[i], q[(i + 1) % 3]));
    // the six tets (P_i, P_i+1, O, X) share the edge cc = O - X:
    // 6 vol = -det(P_i - X, P_i+1 - X, cc) = (P_i - X) . (cc x (P_i+1 - X)), and
    // 24 moment = 6 vol (P_i + P_i+1 + O + X), whose O + X term is summed once
    const V4<double> cc = sub3(O, X);
    for (int i = 0; i < 6; ++i) a[i] = sub3(P[i], X);
    double tvol = 0.0, mx6 = 0.0, my6 = 0.0, mz6 = 0.0;
    for (int i = 0; i < 6; ++i) {
      const int j = (i + 1) % 6;
      const double vol6 = dot3(a[i], cross3(cc, a[j]));
      mx6 = std::fma(vol6, P[i].x + P[j].x, mx6);
      my6 = std::fma(vol6, P[i].y + P[j].y, my6);
      mz6 = std::fma(vol6, P[i].z + P[j].z, mz6);
      tvol += vol6;
    }
    sx = std::fma(O.x + X.x, tvol, sx + mx6);
    sy = std::fma(O.y + X.y, tvol, sy + my6);
    sz = std::fma(O.z + X.z, tvol, sz + mz6);
    sv += tvol;
  }
  vol = (A)(sv / 6.0); mx = (A)(sx / 24.0); my = (A)(sy / 24.0); mz = (A)(sz / 24.0);
  if (c.ev) { c.ev->acc_tris += c.nt; c.ev->acc_pairs += 1; }
  return !(std::fabs((double)vol) < 0.1);
}

// Moment formulation selector: 0 = the reference's, 1 = robust (an FP32
// polytope's integrated in FP64 as above; an FP64 polytope's in its own
// arithmetic).
template <class R, class A>
inline bool moments(const Poly<R>& c, int formulation, const V4<R> tv[4], int cell_nv, A& vol, A& mx, A& my,
                    A& mz) {
  if (formulation != 1) return pair_moments<R, A>(c, vol, mx, my, mz);
  if constexpr (std::is_same<R, float>::value) return pair_moments_robust64<R, A>(c, tv, cell_nv, vol, mx, my, mz);
  return pair_moments_robust<R, A>(c, vol, mx, my, mz);
}

// ---------------------------------------------------------------- kNN ------

// Squared distance, canonical form shared with the HIP kernels.
inline float dist2(const float* a, const float* b) {
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

inline uint64_t knn_key(float d2, uint32_t id) {
  uint32_t u;
  std::memcpy(&u, &d2, 4);
  return ((uint64_t)u << 32) | id;
}

// Exact k nearest reference points of every query by (d2, index) order,
// brute force (knncuda.cu:143-189 keeps lower indices first on ties).  With
// `self` set, query i is reference point i and is left out (knearests.cu:120).
void knn_brute(const float* ref, int n_ref, const float* qry, int n_q, int k, bool self, int32_t* out,
               int nthreads);

// Same result through a uniform grid with an exact stopping bound.
void knn_grid(const float* ref, int n_ref, const float* qry, int n_q, int k, bool self, int32_t* out,
              int nthreads);

// ---------------------------------------------------------- pair kernels ----

struct PairOut {
  int32_t status;
  double vol, mx, my, mz;  // zero when not accumulated
  int32_t accumulated;
};

// Moments of a finished pair: dropped below 0.1 (no_intersection,
// voronoi.cu:599-603); a non-finite result (vertices formed by parallel planes
// in a degenerate configuration) is not accumulated and reported as
// needs_exact_predicates -- the reference would add the NaN.
template <class R, class A>
inline void finish_pair(Poly<R>& c, int mform, const V4<R> tv[4], int cell_nv, PairOut& o) {
  A vol, mx, my, mz;
  const bool keep = moments<R, A>(c, mform, tv, cell_nv, vol, mx, my, mz);
  if (!(std::isfinite((double)vol) && std::isfinite((double)mx) && std::isfinite((double)my) &&
        std::isfinite((double)mz))) {
    c.status = needs_exact_predicates;
    return;
  }
  if (!keep) {
    c.status = no_intersection;
    return;
  }
  o.vol = (double)vol; o.mx = (double)mx; o.my = (double)my; o.mz = (double)mz;
  o.accumulated = 1;
}

// Kernel A for one (tet, seed) pair (voronoi.cu:662-701).  `nbr` holds the
// seed's k nearest other sites, nearest first.
template <class R, class A>
inline PairOut clip_pair_tet(Poly<R>& c, const float* sites, const int32_t* nbr, int k, int seed,
                             const V4<R> tv[4], int mform) {
  PairOut o{};
  const V4<R> s = {(R)sites[3 * seed], (R)sites[3 * seed + 1], (R)sites[3 * seed + 2], R(1)};
  c.init_tet(s, tv);
  if (c.ev) c.ev->pairs += 1;
  auto site = [&](int z) { return V4<R>{(R)sites[3 * z], (R)sites[3 * z + 1], (R)sites[3 * z + 2], R(1)}; };
  for (int v = 0; v < k; ++v) {
    const int z = nbr[v];
    c.clip(Poly<R>::bisector(s, site(z)), false);
    if (c.security_reached(site(z))) break;
    if (c.status != success) {
      o.status = c.status;
      return o;
    }
  }
  if (k == 0 || !c.security_reached(site(nbr[k - 1]))) c.status = security_radius_not_reached;
  if (c.status != no_intersection) finish_pair<R, A>(c, mform, tv, -1, o);
  o.status = c.status;
  return o;
}

// Kernel B for one site: its Voronoi cell from the box (voronoi.cu:540-575).
// Returns the status; on a failing status the polytope is left as it was.
template <class R>
inline int32_t site_cell(Poly<R>& c, const float* sites, const int32_t* nbr, int k, int seed) {
  const V4<R> s = {(R)sites[3 * seed], (R)sites[3 * seed + 1], (R)sites[3 * seed + 2], R(1)};
  c.init_box(s);
  auto site = [&](int z) { return V4<R>{(R)sites[3 * z], (R)sites[3 * z + 1], (R)sites[3 * z + 2], R(1)}; };
  for (int v = 0; v < k; ++v) {
    const int z = nbr[v];
    c.clip(Poly<R>::bisector(s, site(z)), false);
    if (c.security_reached(site(z))) break;
    if (c.status != success) return c.status;
  }
  if (k == 0 || !c.security_reached(site(nbr[k - 1]))) c.status = security_radius_not_reached;
  return c.status;
}

// Kernel C for one (tet, site) pair: the stored cell clipped by the tet's
// four face planes (voronoi.cu:620-645, 703-726).
template <class R, class A>
inline PairOut clip_pair_cell(Poly<R>& c, const V4<R> tv[4], int mform) {
  static const int face[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
  PairOut o{};
  const int cell_nv = c.nv;
  for (int i = 0; i < 4; ++i) {
    c.clip(tri2plane(tv[face[i][0]], tv[face[i][1]], tv[face[i][2]]), true);
    if (c.status == no_intersection) break;
  }
  if (c.status != no_intersection) finish_pair<R, A>(c, mform, tv, cell_nv, o);
  o.status = c.status;
  return o;
}

// voronoi.cu:936
inline int tet_k_heuristic(int n_site, int n_tet, int preferred) {
  if (preferred) return preferred;
  const int h = 10 * (int)std::ceil(8.0 * std::log10(7.0 * n_site / n_tet + 1.0));
  return std::max(h, 20);
}

// Fixed-point scale of the order-independent per-site reduction: per-pair
// FP64 sums are rounded to multiples of 2^-20 and added as 64-bit integers.
constexpr double kFixScale = 1048576.0;  // 2^20

}  // namespace vdo
