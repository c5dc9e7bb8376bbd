// Clipping kernels: per (tet, candidate site) pair, the tet clipped by the
// bisectors of the site with its nearest neighbours (the reference's kernel A,
// voronoi.cu:662-701), the volume and first moment of the result, and the
// per-site reduction and Lloyd update (voronoi.cu:577-618, 728-749).
//
// Two launch forms share the polytope code:
//   * tet-major ("plain"): lane i of the grid takes pair i = tet*TK + rank, the
//     reference's pair order; every lane keeps its own planes.
//   * site-major ("fast"): a wavefront takes the pairs of one site; the site's
//     bisector planes are built once per wave in LDS and shared by all lanes,
//     and lanes that finish pick up the site's next pair.
// Both produce the same per-pair results bit for bit.
//
// Polytope state lives in LDS, slot-major ([slot][lane]) so that lanes working
// on the same slot hit distinct banks.
#include <algorithm>
#include <type_traits>

#include "vd_clip.hpp"
#include "vd_geom.hpp"
#include "vd_knn.hpp"

namespace vd {

namespace {

// Emptiness test of the site-major form: required depth of every tet vertex
// beyond a bisector (distance, frame units) and the relative rounding bound of
// the FP32 plane evaluation fma(nz,z,fma(ny,y,nx*x)) + w.
constexpr float kCullMargin = 1e-2f;
constexpr float kCullRound = 1.2e-6f;

__device__ __forceinline__ uint32_t pack3(int a, int b, int c) { return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16); }
__device__ __forceinline__ int tri_a(uint32_t t) { return (int)(t & 255u); }
__device__ __forceinline__ int tri_b(uint32_t t) { return (int)((t >> 8) & 255u); }
__device__ __forceinline__ int tri_c(uint32_t t) { return (int)((t >> 16) & 255u); }
__device__ __forceinline__ int tri_at(uint32_t t, int k) { return (int)((t >> (8 * k)) & 255u); }

// Slots per trip of the cached clip's distance loop and determinant loop.
#ifndef VD_FAR_UNROLL
#define VD_FAR_UNROLL 4
#endif
#ifndef VD_CAND_UNROLL
#define VD_CAND_UNROLL 1
#endif
constexpr int kFarUnroll = VD_FAR_UNROLL, kCandUnroll = VD_CAND_UNROLL;

// Bisectors a lane of the cached clipping loops may test per trip before the
// trip's cuts run.
#ifndef VD_TESTS_PER_TRIP
#define VD_TESTS_PER_TRIP 3
#endif
constexpr int kTestsPerTrip = VD_TESTS_PER_TRIP;
// Bisectors the site-major kernel tests per trip as one batch (test_batch).
#ifndef VD_BATCH_TESTS
#define VD_BATCH_TESTS 6
#endif
constexpr int kBatchTests = VD_BATCH_TESTS;

// Storage policies --------------------------------------------------------
//
// PlaneStoreLocal: planes [pcap][64] per lane, ids are positions (reference).
// PlaneStoreShared: tet faces of the lane (ids 0..3) in LDS, bisector v at id
//   4+v in a table shared by the wave; the id picks the address, so a fetch is
//   one ds_read_b128 (a select among registers would cost more VALU work than
//   the LDS it saves).  A separate counter nv keeps the reference's
//   plane count so that vertex_overflow happens exactly where it would.

struct PlaneStoreLocal {
  float4* pl;  // [pcap][64]
  int lane;
  __device__ __forceinline__ P4 get(int i) const { return ld4(pl, i * 64 + lane); }
  __device__ __forceinline__ void put(int i, const P4& p) { pl[i * 64 + lane] = make_float4(p.x, p.y, p.z, p.w); }
};

// Face planes of the lane (ids 0..3) in registers, the wave's bisector table
// (id 4+v) in LDS: a fetch reads the table entry and selects.
struct PlaneStoreShared {
  float4* face;        // [4][64] per lane
  const float4* bis;   // [K] shared
  int lane;
  __device__ __forceinline__ P4 get(int i) const {
    const float4* p = i < 4 ? face + (i * 64 + lane) : bis + (i - 4);
    const float4 v = *p;
    return {v.x, v.y, v.z, v.w};
  }
  __device__ __forceinline__ void put(int i, const P4& p) { face[i * 64 + lane] = make_float4(p.x, p.y, p.z, p.w); }
  __device__ __forceinline__ void bind(float4* f, int l) {
    face = f;
    lane = l;
  }
};

// Kernel C: the stored cell's planes (ids 0..nvc-1) in a table shared by the
// wave, the tet's face planes appended after them (ids nvc..nvc+3) in a
// per-lane LDS table; the id picks the address, so a fetch is one LDS read
// (in registers, the selects among the four faces became an indexed array in
// scratch memory).
struct PlaneStoreCell {
  const float4* cpl;  // [nvc] shared
  int nvc;
  float4* fl;         // [4][64] the lane's tet faces
  int lane;
  __device__ __forceinline__ P4 get(int i) const {
    const float4* p = i < nvc ? cpl + i : fl + ((i - nvc) * 64 + lane);
    const float4 v = *p;
    return {v.x, v.y, v.z, v.w};
  }
  __device__ __forceinline__ void put(int i, const P4& p) { fl[(i - nvc) * 64 + lane] = make_float4(p.x, p.y, p.z, p.w); }
};

// Volume pass of the site-major form: face planes of the lane (ids 0..3) in
// LDS, bisector v (id 4+v) rebuilt from the seed and the neighbour's position
// with the operations that built the clipping's table, so the same bits.
struct PlaneStoreRebuild {
  float4* face;          // [4][64] per lane
  const float4* sites;
  const int32_t* nbr;    // the seed's neighbour row
  int n_sites;
  P4 seed;
  int lane;
  __device__ __forceinline__ P4 get(int i) const {
    if (i < 4) return ld4(face, i * 64 + lane);
    const int z = nbr[i - 4];
    if ((unsigned)z >= (unsigned)n_sites) return {0.f, 0.f, 0.f, 0.f};  // as the table marks it
    return bisector(seed, ld4(sites, z));
  }
  __device__ __forceinline__ void put(int i, const P4& p) { face[i * 64 + lane] = make_float4(p.x, p.y, p.z, p.w); }
};

// PlaneStoreTable: as PlaneStoreRebuild, but the seed's first `depth`
// bisectors come from a table in LDS that the wave built once for the seed
// (tbl null: none), with the same function and inputs, hence the same bits.
struct PlaneStoreTable {
  float4* face;          // [4][64] per lane
  const float4* tbl;     // bisectors 0 .. depth-1 of the lane's seed, or null
  int depth;
  const float4* sites;
  const int32_t* nbr;    // the seed's neighbour row
  int n_sites;
  P4 seed;
  int lane;
  __device__ __forceinline__ P4 get(int i) const {
    if (i < 4) return ld4(face, i * 64 + lane);
    if (tbl && i - 4 < depth) return ld4(tbl, i - 4);
    const int z = nbr[i - 4];
    if ((unsigned)z >= (unsigned)n_sites) return {0.f, 0.f, 0.f, 0.f};  // as the table marks it
    return bisector(seed, ld4(sites, z));
  }
  __device__ __forceinline__ void put(int i, const P4& p) { face[i * 64 + lane] = make_float4(p.x, p.y, p.z, p.w); }
};

// Triangle (= polytope vertex) storage policies.
// TriPlain: the vertex's plane triple only (the reference's layout); the
//   security-radius test recomputes every vertex.
// TriCached: plane triple + the vertex's squared distance to the seed in one
//   8-byte word.  A vertex is a function of its three planes only, so the
//   cached value is the one the recomputation would give, bit for bit.
struct TriPlain {
  static constexpr bool kCached = false;
  static constexpr bool kHasD2 = false;
  uint32_t* tr;  // [tcap][64]
  int lane;
  __device__ __forceinline__ uint32_t get(int t) const { return tr[t * 64 + lane]; }
  __device__ __forceinline__ float d2(int) const { return 0.f; }
  __device__ __forceinline__ void put(int t, uint32_t v, float) { tr[t * 64 + lane] = v; }
  __device__ __forceinline__ void swap(int a, int b) {
    const uint32_t x = tr[a * 64 + lane], y = tr[b * 64 + lane];
    tr[a * 64 + lane] = y;
    tr[b * 64 + lane] = x;
  }
  __device__ __forceinline__ void copy(int dst, int src) { tr[dst * 64 + lane] = tr[src * 64 + lane]; }
};
struct TriCached {
  static constexpr bool kCached = true;
  static constexpr bool kHasD2 = true;
  uint2* tr;  // [tcap][64]: (plane triple, d2 bits)
  int lane;
  __device__ __forceinline__ uint32_t get(int t) const { return tr[t * 64 + lane].x; }
  __device__ __forceinline__ float d2(int t) const { return __uint_as_float(tr[t * 64 + lane].y); }
  __device__ __forceinline__ void put(int t, uint32_t v, float d2) { tr[t * 64 + lane] = make_uint2(v, __float_as_uint(d2)); }
  __device__ __forceinline__ void swap(int a, int b) {
    const uint2 x = tr[a * 64 + lane], y = tr[b * 64 + lane];
    tr[a * 64 + lane] = y;
    tr[b * 64 + lane] = x;
  }
  // swap, returning the triangle that moves to slot b
  __device__ __forceinline__ uint32_t swap_get(int a, int b) {
    const uint2 x = tr[a * 64 + lane], y = tr[b * 64 + lane];
    tr[a * 64 + lane] = y;
    tr[b * 64 + lane] = x;
    return x.x;
  }
  __device__ __forceinline__ void copy(int dst, int src) { tr[dst * 64 + lane] = tr[src * 64 + lane]; }
};

// TriPlainD2: the layout of TriCached (plane triple + the vertex's squared
// distance to the seed) for the plain clip, whose vertex count is not bound
// by the cached clip's 32-bit vertex masks: the distances serve the
// security-radius test only.
struct TriPlainD2 : TriCached {
  static constexpr bool kCached = false;
  static constexpr bool kHasD2 = true;
};

// TriRecord: the triangles of a volume-pass record (read only).
struct TriRecord {
  static constexpr bool kCached = false;
  static constexpr bool kHasD2 = false;
  const uint32_t* tr;
  __device__ __forceinline__ uint32_t get(int t) const { return tr[t]; }
};

// Corners of a tet from its record (x0 y0 z0 x1 | y1 z1 x2 y2 | z2 x3 y3 z3).
__device__ __forceinline__ void corners_of(const float4& q0, const float4& q1, const float4& q2, P4 v[4]) {
  v[0] = {q0.x, q0.y, q0.z, 0.f};
  v[1] = {q0.w, q1.x, q1.y, 0.f};
  v[2] = {q1.z, q1.w, q2.x, 0.f};
  v[3] = {q2.y, q2.z, q2.w, 0.f};
}

template <class PS, class TS>
struct Cell {
  PS ps;
  TS ts;
  uint8_t* nx;   // [nxcap][64]
  int lane;
  int pcap, tcap;
  int nv, nt, nr;
  int first;
  int status;
  bool keep_srnr;  // kernel C: a cell flagged security_radius_not_reached keeps clipping
  int filter;      // arithmetic filter of the conflict test: 0 off, 1 the reference's, 2 FP64 re-check
  P4 seed;
  // cached form: slots whose vertex is not yet known to lie within the near
  // ball of the current bisector (see clip), and the d2z of the last clip
  uint32_t far;
  float last_d2z;
  // cached form: the largest squared vertex distance to the seed (fmaxf from
  // 0, as the security test takes it), and the tet's bounding box relative to
  // the seed (every vertex of the polytope lies in it)
  float vmax;
  float blo[3], bhi[3];

  __device__ __forceinline__ bool ok() const {
    return status == kSuccess || (keep_srnr && status == kSecurityRadiusNotReached);
  }
  __device__ __forceinline__ uint32_t tri(int t) const { return ts.get(t); }

  // Conflict of vertex tt with plane e: det4x4 > 0 (voronoi.cu:274-306).  With
  // the filter on, a det below the reference's FP32 error bound flags
  // needs_exact_predicates (mode 1), or is re-evaluated in FP64 from the same
  // planes, whose sign decides and whose own bound flags (mode 2).
  __device__ __forceinline__ bool conflict(uint32_t tt, const P4& e) {
    const P4 p1 = ps.get(tri_a(tt)), p2 = ps.get(tri_b(tt)), p3 = ps.get(tri_c(tt));
    const float det = conflict_det(vertex_minors(p1, p2, p3), e);
    if (filter) {
      const float mx = fmaxf(fmaxf(fabsf(p1.x), fabsf(p2.x)), fmaxf(fabsf(p3.x), fabsf(e.x)));
      const float my = fmaxf(fmaxf(fabsf(p1.y), fabsf(p2.y)), fmaxf(fabsf(p3.y), fabsf(e.y)));
      const float mz = fmaxf(fmaxf(fabsf(p1.z), fabsf(p2.z)), fmaxf(fabsf(p3.z), fabsf(e.z)));
      if (fabsf(det) < filter_bound32(mx, my, mz)) {
        if (filter == 1) {
          status = kNeedsExactPredicates;
        } else {
          const double d64 = conflict_det64(p1, p2, p3, e);
          if (fabs(d64) < filter_bound64(mx, my, mz)) status = kNeedsExactPredicates;
          return d64 > 0.0;
        }
      }
    }
    return det > 0.f;
  }
  __device__ __forceinline__ int nxt(int p) const { return nx[p * 64 + lane]; }
  __device__ __forceinline__ void set_nxt(int p, int v) { nx[p * 64 + lane] = (uint8_t)v; }

  __device__ __forceinline__ P4 vertex(uint32_t t) const { return vertex3(ps.get(tri_a(t)), ps.get(tri_b(t)), ps.get(tri_c(t))); }
  __device__ __forceinline__ float seed_d2(const P4& X) const {
    const P4 d = sub3(X, seed);
    return dot3(d, d);
  }
  // Stores triangle t; the cached form also computes its vertex distance,
  // which it returns.
  __device__ __forceinline__ float put_tri(int t, uint32_t v) {
    if constexpr (TS::kHasD2) {
      const float d2 = seed_d2(vertex(v));
      ts.put(t, v, d2);
      return d2;
    } else {
      ts.put(t, v, 0.f);
      return 0.f;
    }
  }

  // Planes of faces cbd, acd, bad, abc as ids 0..3.
  __device__ __forceinline__ void put_tet_faces(const P4 v[4]) {
    ps.put(0, tri2plane(v[2], v[1], v[3]));
    ps.put(1, tri2plane(v[0], v[2], v[3]));
    ps.put(2, tri2plane(v[1], v[0], v[3]));
    ps.put(3, tri2plane(v[0], v[1], v[2]));
  }

  // Tet start (voronoi.cu:163-201): the faces cbd, acd, bad, abc as plane
  // ids 0..3 and the tet's 4 vertices as triangles (1,3,2), (0,1,2), (0,3,1),
  // (0,2,3), from the tet's precomputed record (tet_prep: the four face planes,
  // then the vertices of triangles 0..3 as x0 y0 z0 x1 | y1 z1 x2 y2 |
  // z2 x3 y3 z3), which holds what put_tet_faces and put_tri would compute.
  __device__ void init_tet_prep(const P4& s, const float4* rec) {
    seed = s;
    status = kSuccess;
    keep_srnr = false;
    const float4 f0 = rec[0], f1 = rec[1], f2 = rec[2], f3 = rec[3];
    ps.put(0, {f0.x, f0.y, f0.z, f0.w});
    ps.put(1, {f1.x, f1.y, f1.z, f1.w});
    ps.put(2, {f2.x, f2.y, f2.z, f2.w});
    ps.put(3, {f3.x, f3.y, f3.z, f3.w});
    nv = 4;
    nt = 4;
    first = kNoPlane;
    far = 0xFu;
    last_d2z = 0.f;
    if constexpr (TS::kCached) {
      const float4 q0 = rec[4], q1 = rec[5], q2 = rec[6];
      P4 X[4];
      corners_of(q0, q1, q2, X);
      const uint32_t tri[4] = {pack3(1, 3, 2), pack3(0, 1, 2), pack3(0, 3, 1), pack3(0, 2, 3)};
      float m = 0.f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float d2 = seed_d2(X[k]);
        ts.put(k, tri[k], d2);
        m = fmaxf(d2, m);
      }
      vmax = m;
      blo[0] = fminf(fminf(X[0].x, X[1].x), fminf(X[2].x, X[3].x)) - s.x;
      blo[1] = fminf(fminf(X[0].y, X[1].y), fminf(X[2].y, X[3].y)) - s.y;
      blo[2] = fminf(fminf(X[0].z, X[1].z), fminf(X[2].z, X[3].z)) - s.z;
      bhi[0] = fmaxf(fmaxf(X[0].x, X[1].x), fmaxf(X[2].x, X[3].x)) - s.x;
      bhi[1] = fmaxf(fmaxf(X[0].y, X[1].y), fmaxf(X[2].y, X[3].y)) - s.y;
      bhi[2] = fmaxf(fmaxf(X[0].z, X[1].z), fmaxf(X[2].z, X[3].z)) - s.z;
    } else {
      ts.put(0, pack3(1, 3, 2), 0.f);
      ts.put(1, pack3(0, 1, 2), 0.f);
      ts.put(2, pack3(0, 3, 1), 0.f);
      ts.put(3, pack3(0, 2, 3), 0.f);
      vmax = 0.f;
    }
  }

  // Box start of kernel B: planes x >= -0.1, x <= 1000.1, y.., z.. and the
  // 8 corner triangles (voronoi.cu:125-160).
  __device__ void init_box(const P4& s) {
    seed = s;
    status = kSuccess;
    keep_srnr = false;
    const float lo = -0.1f, hi = 1000.f + 0.1f;
    ps.put(0, {1.f, 0.f, 0.f, -lo});
    ps.put(1, {-1.f, 0.f, 0.f, hi});
    ps.put(2, {0.f, 1.f, 0.f, -lo});
    ps.put(3, {0.f, -1.f, 0.f, hi});
    ps.put(4, {0.f, 0.f, 1.f, -lo});
    ps.put(5, {0.f, 0.f, -1.f, hi});
    nv = 6;
    float m = put_tri(0, pack3(2, 5, 0));
    m = fmaxf(put_tri(1, pack3(5, 3, 0)), m);
    m = fmaxf(put_tri(2, pack3(1, 5, 2)), m);
    m = fmaxf(put_tri(3, pack3(5, 1, 3)), m);
    m = fmaxf(put_tri(4, pack3(4, 2, 0)), m);
    m = fmaxf(put_tri(5, pack3(4, 0, 3)), m);
    m = fmaxf(put_tri(6, pack3(2, 4, 1)), m);
    m = fmaxf(put_tri(7, pack3(4, 3, 1)), m);
    nt = 8;
    first = kNoPlane;
    far = 0xFFu;
    last_d2z = 0.f;
    // the cached form's largest vertex distance and, for its bisector box
    // test, a box holding every vertex (relative to the seed): the domain
    vmax = m;
    blo[0] = lo - s.x;
    blo[1] = lo - s.y;
    blo[2] = lo - s.z;
    bhi[0] = hi - s.x;
    bhi[1] = hi - s.y;
    bhi[2] = hi - s.z;
  }

  // Every vertex closer to the seed than half the distance to z
  // (voronoi.cu:219-232); d2z = |z - seed|^2.
  __device__ bool security_reached(float d2z) const {
    float vmax = 0.f;
    for (int t = 0; t < nt; ++t) {
      if constexpr (TS::kHasD2)
        vmax = fmaxf(ts.d2(t), vmax);
      else
        vmax = fmaxf(seed_d2(vertex(tri(t))), vmax);
    }
    return d2z > 4.f * vmax;
  }

  // Boundary loop of the removed triangles tr[nt..nt+nr) (voronoi.cu:391-441):
  // removed triangles are merged one at a time into a circular list nx[] of
  // planes; a triangle whose merge would pinch the loop is deferred.  The
  // links of the removed triangles' planes start cleared (`cleared`: the
  // caller cleared them as it removed the triangles).
  __device__ void cavity_boundary(bool cleared = false) {
    for (int r = nt; r < nt + nr && !cleared; ++r) {
      const uint32_t t = tri(r);
      set_nxt(tri_a(t), kNoPlane);
      set_nxt(tri_b(t), kNoPlane);
      set_nxt(tri_c(t), kNoPlane);
    }
    first = kNoPlane;
    int guard = 0;
    int cur = nt;
    while (nr > 0) {
      if (guard++ > 65535) {
        status = kInconsistentBoundary;
        return;
      }
      const uint32_t t = tri(cur);
      const int e0 = tri_a(t), e1 = tri_b(t), e2 = tri_c(t);
      const int n0 = nxt(e0), n1 = nxt(e1), n2 = nxt(e2);
      const bool on0 = n0 != kNoPlane, on1 = n1 != kNoPlane, on2 = n2 != kNoPlane;
      // opp[k]: the loop already runs e[k+1] -> e[k]
      const bool op0 = n1 == e0, op1 = n2 == e1, op2 = n0 == e2;
      bool simple = true;
      if (!op0 && !op1 && on1) simple = false;
      if (!op1 && !op2 && on2) simple = false;
      if (!op2 && !op0 && on0) simple = false;
      const bool none = !op0 && !op1 && !op2;
      if (none && first != kNoPlane) simple = false;
      if (!simple) {
        ++cur;
        if (cur == nt + nr) cur = nt;
        continue;
      }
      // The merge as selects: a link e -> e' is added where the loop does
      // not already run e' -> e, a plane between two such edges leaves the
      // loop, and `first` steps past a plane that leaves (each plane's link
      // is still its n value where it is read: its own edge is in the loop).
      // The first triangle of an empty loop is the loop.
      if (none) first = e0;
      if (op0 && op1 && first == e1) first = n1;
      if (op1 && op2 && first == e2) first = n2;
      if (op2 && op0 && first == e0) first = n0;
      set_nxt(e0, op2 && op0 ? kNoPlane : (op0 ? n0 : e1));
      set_nxt(e1, op0 && op1 ? kNoPlane : (op1 ? n1 : e2));
      set_nxt(e2, op1 && op2 ? kNoPlane : (op2 ? n2 : e0));
      // the merged triangle is not needed again: the last unmerged one takes
      // its slot
      ts.copy(cur, nt + nr - 1);
      cur = nt;
      --nr;
    }
  }

  // Cached form of the clip in two steps, so that a caller can test a lane's
  // next bisectors one after another and run the expensive cut for many lanes
  // at once.  clip_test: the plane-cap check, the conflict flags and the
  // security outcome; returns as clip() when the plane removes no vertex (or
  // flags needs_exact_predicates, or overflows), else 2 with the flags in
  // conf_out.
  __device__ __forceinline__ int clip_test(const P4& eqn, float d2z, uint32_t& conf_out) {
    if (nv >= pcap) {
      status = kVertexOverflow;
      return -1;
    }
    ++nv;
    nr = 0;
    // A vertex within 0.9 * |z - s| / 2 of the seed (d2 < 0.2025 d2z) is
    // closer to s than to z by at least 0.05 |z - s|, far beyond the FP32
    // error of the determinant, so the test's outcome (no conflict) is known
    // without computing it.  d2z does not decrease along the neighbour list,
    // so a vertex found inside that ball stays inside for the later
    // bisectors: `far` keeps the slots not yet found inside.
    // The same margin decides a whole bisector at once: the plane lies at
    // |z - s| / 2 from the seed along u = (z - s) / |z - s|, so when every
    // point of the tet's box (which holds every vertex) projects below
    // 0.45 |z - s| on u, no vertex can be in conflict.  Then only the security
    // outcome is needed, and the largest vertex distance gives it.
    if (!filter) {
      // (z - s) = -e.xyz; max over the box of (x - s).(z - s)
      const float px = -fminf(blo[0] * eqn.x, bhi[0] * eqn.x);
      const float py = -fminf(blo[1] * eqn.y, bhi[1] * eqn.y);
      const float pz = -fminf(blo[2] * eqn.z, bhi[2] * eqn.z);
      if (px + py + pz < 0.45f * d2z) {
        --nv;
        return d2z > 4.f * vmax ? 1 : 0;
      }
    }
    if (!(d2z >= last_d2z)) far = (1u << nt) - 1u;
    last_d2z = d2z;
    // (with the filter on every determinant is evaluated, as in the reference)
    const float near2 = filter ? -1.f : 0.2025f * d2z;
    uint32_t cand = 0;
    // kFarUnroll slots per loop trip, so that their LDS reads are in flight
    // together
    for (uint32_t m = far; m;) {
      int t[kFarUnroll];
      bool has[kFarUnroll];
      float d[kFarUnroll];
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) {
        has[j] = m != 0;
        t[j] = has[j] ? __builtin_ctz(m) : 0;
        m &= m - 1;
      }
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) d[j] = ts.d2(t[j]);
      // as selects rather than branches: per slot, found inside the near
      // ball (cleared from far) or a candidate (possibly blocking)
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) {
        const uint32_t bit = has[j] ? 1u << t[j] : 0u;
        const uint32_t keep = d[j] < near2 ? 0u : bit;
        far &= ~(bit & ~keep);
        cand |= keep;
      }
    }
    // the determinant only for candidates: a wave pays for the largest
    // candidate count rather than the largest vertex count (kCandUnroll per
    // trip)
    uint32_t conf = 0;
    for (uint32_t m = cand; m;) {
      int t[kCandUnroll];
      bool has[kCandUnroll];
      uint32_t tt[kCandUnroll];
#pragma unroll
      for (int j = 0; j < kCandUnroll; ++j) {
        has[j] = m != 0;
        t[j] = has[j] ? __builtin_ctz(m) : 0;
        m &= m - 1;
      }
#pragma unroll
      for (int j = 0; j < kCandUnroll; ++j) tt[j] = tri(t[j]);
      if (filter) {
#pragma unroll
        for (int j = 0; j < kCandUnroll; ++j)
          if (has[j] && conflict(tt[j], eqn)) conf |= 1u << t[j];
      } else {
        // without the filter the test has no side effect: every slot's
        // determinant (slot 0 stands in for a missing one) and a select
#pragma unroll
        for (int j = 0; j < kCandUnroll; ++j) {
          const float det = conflict_det(vertex_minors(ps.get(tri_a(tt[j])), ps.get(tri_b(tt[j])), ps.get(tri_c(tt[j]))), eqn);
          conf |= has[j] && det > 0.f ? 1u << t[j] : 0u;
        }
      }
    }
    if (status == kNeedsExactPredicates) {
      // the reference stops right after its conflict loop (voronoi.cu:459-461):
      // the conflicting vertices are gone, no plane is dropped, no fan
      while (conf) {
        const int i = __builtin_ctz(conf);
        --nt;
        ++nr;
        const uint32_t bi = 1u << i, bn = 1u << nt;
        if (i != nt) {
          ts.swap(i, nt);
          conf = (conf & ~(bi | bn)) | ((conf & bn) ? bi : 0u);
          far = (far & ~(bi | bn)) | ((far & bn) ? bi : 0u);
        } else {
          conf &= ~bi;
          far &= ~bi;
        }
      }
      return -1;
    }
    if (conf == 0u) {
      --nv;
      return d2z > 4.f * vmax ? 1 : 0;
    }
    conf_out = conf;
    return 2;
  }

  // The next bisectors e[0..n) of the neighbour list tested against the same
  // polytope, as clip_test tests them one at a time (filter off): no vertex is
  // removed before the first bisector that conflicts, so the tests share one
  // pass over the far vertices (with the smallest near ball of the batch) and
  // one pass over the candidates, in which a vertex's minors serve every
  // bisector's determinant.  conf[k]: the vertices in conflict with bisector
  // k, 0 where its box test holds (k >= n: unused).  Each determinant is
  // clip_test's, bit for bit, and a candidate of the batch that is not one of
  // bisector k's own lies inside k's near ball, where its determinant is
  // negative, so the first k with conf[k] != 0 and conf[k] itself are what the
  // tests one at a time find.  The caller checks the vertex cap first.
  template <int N>
  __device__ __forceinline__ void test_batch(const P4 (&e)[N], const float (&d2z)[N], int n, uint32_t (&conf)[N]) {
    bool skip[N];
    float dmin = INFINITY;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const float px = -fminf(blo[0] * e[k].x, bhi[0] * e[k].x);
      const float py = -fminf(blo[1] * e[k].y, bhi[1] * e[k].y);
      const float pz = -fminf(blo[2] * e[k].z, bhi[2] * e[k].z);
      skip[k] = k >= n || !(d2z[k] == d2z[k]) || px + py + pz < 0.45f * d2z[k];
      dmin = skip[k] ? dmin : fminf(dmin, d2z[k]);
      conf[k] = 0u;
    }
    if (dmin == INFINITY) return;
    if (!(dmin >= last_d2z)) far = (1u << nt) - 1u;
    last_d2z = dmin;
    const float near2 = 0.2025f * dmin;
    uint32_t cand = 0;
    for (uint32_t m = far; m;) {
      int t[kFarUnroll];
      bool has[kFarUnroll];
      float d[kFarUnroll];
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) {
        has[j] = m != 0;
        t[j] = has[j] ? __builtin_ctz(m) : 0;
        m &= m - 1;
      }
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) d[j] = ts.d2(t[j]);
#pragma unroll
      for (int j = 0; j < kFarUnroll; ++j) {
        const uint32_t bit = has[j] ? 1u << t[j] : 0u;
        const uint32_t keep = d[j] < near2 ? 0u : bit;
        far &= ~(bit & ~keep);
        cand |= keep;
      }
    }
    for (uint32_t m = cand; m;) {
      int t[kCandUnroll];
      bool has[kCandUnroll];
      uint32_t tt[kCandUnroll];
#pragma unroll
      for (int j = 0; j < kCandUnroll; ++j) {
        has[j] = m != 0;
        t[j] = has[j] ? __builtin_ctz(m) : 0;
        m &= m - 1;
      }
#pragma unroll
      for (int j = 0; j < kCandUnroll; ++j) tt[j] = tri(t[j]);
#pragma unroll
      for (int j = 0; j < kCandUnroll; ++j) {
        const Minors mn = vertex_minors(ps.get(tri_a(tt[j])), ps.get(tri_b(tt[j])), ps.get(tri_c(tt[j])));
#pragma unroll
        for (int k = 0; k < N; ++k) conf[k] |= has[j] && conflict_det(mn, e[k]) > 0.f ? 1u << t[j] : 0u;
      }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) conf[k] = skip[k] ? 0u : conf[k];
  }

  // clip_cut: removal of the flagged vertices, the cavity walk and the fan of
  // new vertices around plane `pid`; returns as clip().
  __device__ __forceinline__ int clip_cut(int pid, uint32_t conf, float d2z) {
    const float near2 = filter ? -1.f : 0.2025f * d2z;
    // The reference's loop (ascending slot, a conflicting slot takes the
    // last live triangle and is tested again) replayed from the flags: a
    // flag depends on the triangle's planes only, so the resulting order is
    // the one the in-loop tests produce; each trip removes one triangle.
    while (conf) {
      const int i = __builtin_ctz(conf);
      --nt;
      ++nr;
      // slot i takes slot nt's triangle and flags (a no-op swap when i == nt);
      // the removed triangle's planes get their links cleared for the
      // cavity walk
      const uint32_t bi = 1u << i, bn = 1u << nt;
      const bool mv = i != nt;
      const uint32_t rt = ts.swap_get(i, nt);
      set_nxt(tri_a(rt), kNoPlane);
      set_nxt(tri_b(rt), kNoPlane);
      set_nxt(tri_c(rt), kNoPlane);
      conf = (conf & ~(bi | bn)) | (mv && (conf & bn) ? bi : 0u);
      far = (far & ~(bi | bn)) | (mv && (far & bn) ? bi : 0u);
    }
    if (nt == 0) {
      status = kNoIntersection;
      return -1;
    }
    // the largest distance of the kept vertices, then of the new ones (the
    // security test: every vertex closer than half of |z - s|)
    float m = 0.f;
    for (int t = 0; t < nt; ++t) m = fmaxf(ts.d2(t), m);
    cavity_boundary(true);
    if (status != kSuccess) return -1;
    if (first == kNoPlane) {
      vmax = m;
      return d2z > 4.f * m ? 1 : 0;
    }
    int c = first;
    int guard = 0;
    // the new vertex (pid, c, n): the plane of pid and of c are already in
    // registers (c is the previous vertex's n), one plane read per vertex
    const P4 pp = ps.get(pid);
    P4 pc = ps.get(first);
    // one exit for the three failures (an open or runaway loop, the vertex
    // cap), taken before anything is stored for the failing vertex
    // (err: the failure status, or -1; kTriangleOverflow is 0)
    int err = -1;
    do {
      err = (c == kNoPlane || guard++ > pcap) ? kInconsistentBoundary : (nt + 1 >= tcap ? kTriangleOverflow : -1);
      const int n = err >= 0 ? kNoPlane : nxt(c);
      if (err < 0 && n == kNoPlane) err = kInconsistentBoundary;
      if (err >= 0) break;
      const P4 pn = ps.get(n);
      const float d = seed_d2(vertex3(pp, pc, pn));
      ts.put(nt, pack3(pid, c, n), d);
      far |= d < near2 ? 0u : 1u << nt;
      m = fmaxf(d, m);
      ++nt;
      c = n;
      pc = pn;
    } while (c != first);
    if (err >= 0) {
      status = err;
      return -1;
    }
    vmax = m;
    return d2z > 4.f * m ? 1 : 0;
  }

  // Clip by plane eqn stored under id `pid` (voronoi.cu:443-486); the plane is
  // the bisector of the seed with a site at squared distance d2z.
  //
  // Returns the outcome of the security-radius test that follows the clip in
  // kernel A (d2z > 4 max_t |x_t - seed|^2) when the cached form knows it (1
  // or 0), or -1 when the caller has to evaluate it.
  __device__ int clip(int pid, const P4& eqn, float d2z) {
    if constexpr (TS::kCached) {
      uint32_t conf = 0;
      const int r = clip_test(eqn, d2z, conf);
      return r != 2 ? r : clip_cut(pid, conf, d2z);
    } else {
      if (nv >= pcap) {
        status = kVertexOverflow;
        return -1;
      }
      ++nv;
      nr = 0;
      int i = 0;
      while (i < nt) {
        if (conflict(tri(i), eqn)) {
          --nt;
          ts.swap(i, nt);
          ++nr;
        } else {
          ++i;
        }
      }
      if (status == kNeedsExactPredicates) return -1;  // voronoi.cu:459-461
      if (nr == 0) {
        --nv;
        return -1;
      }
      if (nt == 0) {
        status = kNoIntersection;
        return -1;
      }
      cavity_boundary();
      if (!ok()) return -1;
      if (first == kNoPlane) return -1;
      int c = first;
      int guard = 0;
      do {
        if (c == kNoPlane || guard++ > pcap) {
          status = kInconsistentBoundary;
          return -1;
        }
        if (nt + 1 >= tcap) {
          status = kTriangleOverflow;
          return -1;
        }
        const int n = nxt(c);
        if (n == kNoPlane) {
          status = kInconsistentBoundary;
          return -1;
        }
        put_tri(nt, pack3(pid, c, n));
        ++nt;
        c = n;
      } while (c != first);
      return -1;
    }
  }

  // Volume and first moment: the reference's six-tet-per-vertex decomposition
  // around the seed (voronoi.cu:532-537, 577-618).
  template <int ACC>
  __device__ void moments_ref(PairSums<ACC>& S) const {
    const P4 C = seed;
    for (int t = 0; t < nt; ++t) {
      const uint32_t tt = tri(t);
      const P4 q0 = ps.get(tri_a(tt)), q1 = ps.get(tri_b(tt)), q2 = ps.get(tri_c(tt));
      const P4 X = vertex3(q0, q1, q2);
      P4 P[6];
      P[0] = project_ref(C, q0);
      P[2] = project_ref(C, q1);
      P[4] = project_ref(C, q2);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        P4 n = cross3(sub3(P[2 * k], C), sub3(P[(2 * k + 2) % 6], C));
        n.w = -dot3(C, n);
        P[2 * k + 1] = project_ref(X, n);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        float v;
        P4 b;
        tet_vol_bary(P[k], P[(k + 1) % 6], C, X, v, b);
        S.add(v, b);
      }
    }
  }

  // FP64 plane of plane id `id` rebuilt from the FP32 inputs (vd_geom.hpp):
  // cell_nv < 0, a tet -> cell polytope: ids 0-3 the tet's faces, later ids
  // bisectors; cell_nv >= 0, a cell clipped by a tet (kernel C): ids 0-5 the
  // box (exact in FP32), ids below cell_nv bisectors, later ids the tet's
  // faces, each recognised by its FP32 plane.
  __device__ D4 plane64(int id, const P4 v[4], int cell_nv) const {
    const P4 p = ps.get(id);
    int f = id;
    if (cell_nv < 0) {
      if (id >= 4) return bisector64(p, seed);
    } else {
      if (id < 6) return up4(p);
      if (id < cell_nv) return bisector64(p, seed);
      f = -1;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const P4 q = face_plane(k, v);
        if (__float_as_uint(q.x) == __float_as_uint(p.x) && __float_as_uint(q.y) == __float_as_uint(p.y) &&
            __float_as_uint(q.z) == __float_as_uint(p.z) && __float_as_uint(q.w) == __float_as_uint(p.w) && f < 0)
          f = k;
      }
      if (f < 0) return up4(p);
    }
    if (f == 0) return face64(v[2], v[1], v[3]);
    if (f == 1) return face64(v[0], v[2], v[3]);
    if (f == 2) return face64(v[1], v[0], v[3]);
    return face64(v[0], v[1], v[2]);
  }
  // FP32 plane of face k of the tet, as put_tet_faces stores it (k constant)
  __device__ __forceinline__ static P4 face_plane(int k, const P4 v[4]) {
    return k == 0 ? tri2plane(v[2], v[1], v[3])
                  : k == 1 ? tri2plane(v[0], v[2], v[3]) : k == 2 ? tri2plane(v[1], v[0], v[3]) : tri2plane(v[0], v[1], v[2]);
  }

  // The same six-tet ring around an apex O, with edge feet taken on the edge
  // line (the library's well-conditioned formulation), integrated in FP64 over
  // planes rebuilt in FP64; v: the tet's corners.  The signed ring is exact for
  // any apex, so O is the seed (in FP64 a seed outside the polytope costs
  // nothing measurable).
  template <int ACC>
  __device__ void moments_robust(PairSums<ACC>& S, const P4 v[4], int cell_nv) const {
    const D4 O = up4(seed);
    double sv = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;  // 6 x volume, 24 x moments
    // the next vertex's triangle is read one vertex ahead (in the volume pass
    // it comes from the record in global memory)
    uint32_t tt_next = nt > 0 ? tri(0) : 0u;
    for (int t = 0; t < nt; ++t) {
      const uint32_t tt = tt_next;
      tt_next = tri(min(t + 1, nt - 1));
      const D4 q[3] = {plane64(tri_a(tt), v, cell_nv), plane64(tri_b(tt), v, cell_nv),
                       plane64(tri_c(tt), v, cell_nv)};
      const D4 X = vertex3d(q[0], q[1], q[2]);
      // ring point k: the apex's foot on plane k/2 (k even) or on the line of
      // planes k/2 and k/2+1 (k odd)
      auto ring = [&](int k) -> D4 {
        return (k & 1) ? foot_on_line_d(O, X, cross3d(q[k >> 1], q[((k >> 1) + 1) % 3]))
                       : foot_on_plane_d(O, q[k >> 1]);
      };
      // The six tets (P_k, P_k+1, O, X) share the edge c = O - X:
      // 6 vol = -det(P_k - X, P_k+1 - X, c) = (P_k - X) . (c x (P_k+1 - X)), and
      // 24 moment = 6 vol (P_k + P_k+1 + O + X), whose O + X term is summed once.
      const D4 c = sub3d(O, X);
      const D4 P0 = ring(0);
      const D4 a0 = sub3d(P0, X);
      D4 Pp = P0, ap = a0;
      double tv = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const D4 Pn = k < 5 ? ring(k + 1) : P0;
        const D4 an = k < 5 ? sub3d(Pn, X) : a0;
        const double v6 = dot3d(ap, cross3d(c, an));
        mx = fma(v6, Pp.x + Pn.x, mx);
        my = fma(v6, Pp.y + Pn.y, my);
        mz = fma(v6, Pp.z + Pn.z, mz);
        tv += v6;
        Pp = Pn;
        ap = an;
      }
      sx = fma(O.x + X.x, tv, sx + mx);
      sy = fma(O.y + X.y, tv, sy + my);
      sz = fma(O.z + X.z, tv, sz + mz);
      sv += tv;
    }
    S.set(sv / 6.0, sx / 24.0, sy / 24.0, sz / 24.0);
  }
};

// Result of one pair, written to the per-site accumulators.
template <int ACC>
__device__ __forceinline__ void accumulate(const ClipArgs& a, int site, const PairSums<ACC>& S);

template <>
__device__ __forceinline__ void accumulate<0>(const ClipArgs& a, int site, const PairSums<0>& S) {
  float* p = a.acc_f32 + 4 * (size_t)site;
  unsafeAtomicAdd(p + 0, S.v);
  unsafeAtomicAdd(p + 1, S.x);
  unsafeAtomicAdd(p + 2, S.y);
  unsafeAtomicAdd(p + 3, S.z);
}
template <>
__device__ __forceinline__ void accumulate<1>(const ClipArgs& a, int site, const PairSums<1>& S) {
  unsigned long long* p = a.acc_fix + 4 * (size_t)site;
  atomicAdd(p + 0, (unsigned long long)__double2ll_rn(S.v * kFixScale));
  atomicAdd(p + 1, (unsigned long long)__double2ll_rn(S.x * kFixScale));
  atomicAdd(p + 2, (unsigned long long)__double2ll_rn(S.y * kFixScale));
  atomicAdd(p + 3, (unsigned long long)__double2ll_rn(S.z * kFixScale));
}

__device__ __forceinline__ void count_status(const ClipArgs& a, int status, bool live) {
  if (!a.status_hist) return;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    const unsigned long long m = __ballot(live && status == s);
    if (lane == 0 && m) atomicAdd(a.status_hist + s, (unsigned long long)__popcll(m));
  }
}

__device__ __forceinline__ P4 site_at(const ClipArgs& a, int i) { return ld4(a.sites, i); }

__device__ __forceinline__ void tet_corners(const ClipArgs& a, int t, P4 v[4]) {
  const float4* r = a.tet_xyz + 3 * (size_t)t;
  corners_of(r[0], r[1], r[2], v);
}


// Clips one pair whose cell has been set up; returns through S and status.
// kernel A's loop (voronoi.cu:679-700) with the bisector of the v-th neighbour
// under plane id pid(v).
template <int MF, int ACC, class PS, class TS, class PlaneOf>
__device__ __forceinline__ bool run_pair(const ClipArgs& a, Cell<PS, TS>& c, const int32_t* nbr, PlaneOf plane_of,
                                         int tet, PairSums<ACC>& S) {
  const int K = a.K;
  bool stopped = false;
  if constexpr (TS::kCached) {
    // the cached form: up to kTestsPerTrip bisectors tested per loop trip,
    // until one removes a vertex, then its cut, so that the lanes of a wave
    // run their cuts together (the same operations in the same order per lane)
    int v = 0;
    bool go = K > 0;
    // Without the arithmetic filter the trip's bisectors are tested as one
    // batch (Cell::test_batch) and their outcomes taken in list order, as
    // the tests one at a time below take them; only the cutting bisector is
    // stored, at the next free plane position.
    while (go && !c.filter) {
      P4 e[kBatchTests];
      float d2z[kBatchTests];
      bool ok[kBatchTests];
      uint32_t cf[kBatchTests];
      const int n = min(kBatchTests, K - v);
      const P4 s = c.seed;
#pragma unroll
      for (int k = 0; k < kBatchTests; ++k) {
        const int z = nbr[min(v + k, K - 1)];
        ok[k] = (unsigned)z < (unsigned)a.n_sites;
        const P4 zp = site_at(a, ok[k] ? z : 0);
        e[k] = bisector(s, zp);
        const P4 d = sub3(zp, s);
        d2z[k] = ok[k] ? dot3(d, d) : __int_as_float(0x7fc00000);
        cf[k] = 0u;
      }
      if (c.nv >= c.pcap) {
        // the vertex cap: the first test reports it (after its neighbour check)
        if (!ok[0]) {  // defensive: never read outside the site array
          c.status = kInconsistentBoundary;
          return false;
        }
        c.status = kVertexOverflow;
        if (!c.security_reached(d2z[0])) return false;
        stopped = true;
        break;
      }
      c.test_batch(e, d2z, n, cf);
      int cut_k = -1;
      bool bad = false;
      P4 eq = e[0];
      float dz = d2z[0];
      uint32_t conf = 0u;
#pragma unroll
      for (int k = 0; k < kBatchTests; ++k) {
        if (!go || cut_k >= 0) continue;
        if (!ok[k]) {
          bad = true;
          go = false;
        } else if (cf[k] != 0u) {
          cut_k = k;
          eq = e[k];
          dz = d2z[k];
          conf = cf[k];
        } else if (d2z[k] > 4.f * c.vmax) {
          stopped = true;
          go = false;
        } else if (++v == K) {
          go = false;
        }
      }
      if (bad) {  // defensive: never read outside the site array
        c.status = kInconsistentBoundary;
        return false;
      }
      if (cut_k >= 0) {
        const int pid = c.nv;  // below pcap: not capped
        c.ps.put(pid, eq);
        ++c.nv;
        c.nr = 0;
        const int r = c.clip_cut(pid, conf, dz);
        if (r >= 0 ? r != 0 : c.security_reached(dz)) {
          stopped = true;
          go = false;
        } else if (c.status != kSuccess) {
          return false;
        } else if (++v == K) {
          go = false;
        }
      }
    }
    while (go) {
      bool cut = false;
      uint32_t conf = 0;
      float cut_d2z = 0.f;
      int cut_pid = 0;
      for (int it = 0; it < kTestsPerTrip && go; ++it) {
        const int z = nbr[v];
        if ((unsigned)z >= (unsigned)a.n_sites) {  // defensive: never read outside the site array
          c.status = kInconsistentBoundary;
          return false;
        }
        int pid;
        P4 eq;
        float d2z;
        plane_of(v, z, pid, eq, d2z);
        const int r = c.clip_test(eq, d2z, conf);
        if (r == 2) {
          cut = true;
          cut_d2z = d2z;
          cut_pid = pid;
          break;
        }
        if (r >= 0 ? r != 0 : c.security_reached(d2z)) {
          stopped = true;
          go = false;
        } else if (c.status != kSuccess) {
          return false;
        } else if (++v == K) {
          go = false;
        }
      }
      if (cut) {
        const int r = c.clip_cut(cut_pid, conf, cut_d2z);
        if (r >= 0 ? r != 0 : c.security_reached(cut_d2z)) {
          stopped = true;
          go = false;
        } else if (c.status != kSuccess) {
          return false;
        } else if (++v == K) {
          go = false;
        }
      }
    }
  } else {
    for (int v = 0; v < K; ++v) {
      const int z = nbr[v];
      if ((unsigned)z >= (unsigned)a.n_sites) {  // defensive: never read outside the site array
        c.status = kInconsistentBoundary;
        return false;
      }
      int pid;
      P4 eq;
      float d2z;
      plane_of(v, z, pid, eq, d2z);
      const int sec = c.clip(pid, eq, d2z);
      if (sec >= 0 ? sec != 0 : c.security_reached(d2z)) {
        stopped = true;
        break;
      }
      if (c.status != kSuccess) return false;
    }
  }
  if (!stopped || K == 0) {
    // the last neighbour's test, as the reference makes it after the loop
    bool ok = false;
    if (K > 0 && (unsigned)nbr[K - 1] < (unsigned)a.n_sites) {
      int pid;
      P4 eq;
      float d2z;
      plane_of(K - 1, nbr[K - 1], pid, eq, d2z);
      ok = c.security_reached(d2z);
    }
    if (!ok) c.status = kSecurityRadiusNotReached;
  }
  if (c.status == kNoIntersection) return false;
  if (MF == 0) {
    c.template moments_ref<ACC>(S);
  } else {
    P4 v[4];
    tet_corners(a, tet, v);
    c.template moments_robust<ACC>(S, v, -1);
  }
  // vertices formed by parallel planes (degenerate input) give non-finite sums;
  // such a pair is not accumulated
  if (!(isfinite((double)S.v) && isfinite((double)S.x) && isfinite((double)S.y) && isfinite((double)S.z))) {
    c.status = kNeedsExactPredicates;
    return false;
  }
  if (fabs((double)S.v) < 0.1) {
    c.status = kNoIntersection;
    return false;
  }
  return true;
}

// Per tet: its face planes and the vertices of its four start triangles, as
// put_tet_faces and put_tri compute them (the same functions, so the same
// bits), for init_tet_prep.
__global__ __launch_bounds__(256) void k_tet_prep(const float4* tet_xyz, long long n, float4* out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  P4 v[4];
  corners_of(tet_xyz[3 * t], tet_xyz[3 * t + 1], tet_xyz[3 * t + 2], v);
  const P4 pl[4] = {tri2plane(v[2], v[1], v[3]), tri2plane(v[0], v[2], v[3]), tri2plane(v[1], v[0], v[3]),
                    tri2plane(v[0], v[1], v[2])};
  const P4 X[4] = {vertex3(pl[1], pl[3], pl[2]), vertex3(pl[0], pl[1], pl[2]), vertex3(pl[0], pl[3], pl[1]),
                   vertex3(pl[0], pl[2], pl[3])};
  float4* o = out + kTetPrepWords * t;
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = make_float4(pl[k].x, pl[k].y, pl[k].z, pl[k].w);
  o[4] = make_float4(X[0].x, X[0].y, X[0].z, X[1].x);
  o[5] = make_float4(X[1].y, X[1].z, X[2].x, X[2].y);
  o[6] = make_float4(X[2].z, X[3].x, X[3].y, X[3].z);
}

// ------------------------------------------------------------ tet-major ----

// CACHED: vertices stored with their distance to the seed (at most 32 per
// polytope), as in the site-major form; used to re-clip the pairs that
// overflow the site-major launch's smaller vertex cap.
template <int MF, int ACC, int CACHED>
__global__ __launch_bounds__(64) void k_clip_plain(ClipArgs a) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int lane = threadIdx.x;
  using TS = typename std::conditional<CACHED != 0, TriCached, TriPlain>::type;
  Cell<PlaneStoreLocal, TS> c;
  c.ps.pl = (float4*)smem;
  c.ps.lane = lane;
  c.ts.tr = (decltype(c.ts.tr))(c.ps.pl + (size_t)a.pcap * 64);
  c.ts.lane = lane;
  c.nx = (uint8_t*)(c.ts.tr + (size_t)a.tcap * 64);
  c.lane = lane;
  c.pcap = a.pcap;
  c.tcap = a.tcap;
  c.filter = a.filter;
  // Either all pairs (one per lane) or the pairs of a retry list (grid-strided).
  const long long n_items = a.pair_list ? (long long)min(*a.pair_list_count, (unsigned)a.n_pairs) : a.n_pairs;
  for (long long base = (long long)blockIdx.x * 64; base < n_items; base += (long long)gridDim.x * 64) {
  const long long item = base + lane;
  const bool live = item < n_items;
  const long long pair = !live ? 0 : (a.pair_list ? (long long)a.pair_list[item] : item);
  int status = kNoIntersection;
  PairSums<ACC> S;
  bool acc = false;
  int seed_id = 0;
  if (live && (unsigned)a.tet_knn[pair] >= (unsigned)a.n_sites) {
    status = kInconsistentBoundary;  // defensive: candidate id out of range
    if (a.pair_status) a.pair_status[pair] = (uint8_t)status;
  } else if (live) {
    const int t = (int)(pair / a.TK);
    seed_id = a.tet_knn[pair];
    const P4 s = site_at(a, seed_id);
    c.init_tet_prep(s, a.tet_prep + kTetPrepWords * (size_t)t);
    const int32_t* nbr = a.site_knn + (size_t)seed_id * a.K;
    ClipArgs ak = a;
    if (a.knn2) {
      nbr = a.knn2 + (size_t)a.knn2_row[seed_id] * a.K2;
      ak.K = a.K2;
    }
    auto plane_of = [&](int v, int z, int& pid, P4& eq, float& d2z) {
      const P4 zp = site_at(a, z);
      eq = bisector(s, zp);
      const P4 d = sub3(zp, s);
      d2z = dot3(d, d);
      pid = c.nv;  // next free position
      if (pid < a.pcap) c.ps.put(pid, eq);
    };
    acc = run_pair<MF, ACC>(ak, c, nbr, plane_of, t, S);
    status = c.status;
    if (a.retry_list && ((status == kTriangleOverflow && a.tcap < a.tcap_full) ||
                         (status == kVertexOverflow && a.pcap < a.pcap_full))) {
      // over one of this launch's caps: recomputed by a launch at wider ones
      const unsigned slot = atomicAdd(a.retry_count, 1u);
      if (slot < (unsigned)a.n_pairs) a.retry_list[slot] = (unsigned)pair;
      status = -1;
    } else if (status == kSecurityRadiusNotReached && a.srnr_list) {
      // clipped again with deeper neighbour lists; counted there
      const unsigned slot = atomicAdd(a.srnr_count, 1u);
      if (slot < (unsigned)a.n_pairs) a.srnr_list[slot] = (unsigned)pair;
      status = -1;
    } else {
      if (acc) accumulate<ACC>(a, seed_id, S);
      if (a.pair_status) a.pair_status[pair] = (uint8_t)status;
      if (a.pair_vals) {
        double* pv = a.pair_vals + 4 * pair;
        pv[0] = acc ? (double)S.v : 0.0;
        pv[1] = acc ? (double)S.x : 0.0;
        pv[2] = acc ? (double)S.y : 0.0;
        pv[3] = acc ? (double)S.z : 0.0;
      }
    }
  }
  count_status(a, status, live);
  }
}

// ----------------------------------------------------------- site-major ----
//
// Pairs grouped by seed: site s owns pair_ids[pair_start[s] .. pair_start[s+1]),
// or the first pair_count[s] of them when the emptiness test (k_cull_site)
// has run.  A wavefront takes sites one at a time from a global counter.  For
// each site it builds the bisector table once (shared by the wave) and queues
// the site's pairs in an LDS ring.  Every lane clips its own pair a few
// bisectors per loop trip and takes the next queued pair as soon as it is
// done, so lanes do not wait for the slowest pair of a group.  Two table slots let lanes start on the next site
// while others finish the previous one.  Tet corners come from a per-tet
// record (3 float4 per tet), one dependent load per pair.  A pair sees the
// same planes in the same order as in kernel A, so its result is the same bit
// for bit.
//
// The volume and moments of a finished polytope are not computed here: the
// lane stores its triangles as a record and takes its next pair at once, and
// k_site_moments runs the volume pass over all records with every lane busy
// (in place, a lane would either stall the wave's clipping or wait idle for
// others to finish).

constexpr int kQueue = 128;           // ring of queued pair ids
constexpr float kInvalidNbr = -1.f;   // table mark of an out-of-range neighbour id
constexpr int kIdle = 0, kClipping = 1;  // lane states
constexpr int kRecChunk = 128;           // record slots a wave reserves at a time
// Records of polytopes with more than kRecSplit vertices go to chunks of their
// own, so that a wave of the volume pass (64 records of one chunk) runs over
// polytopes of similar size.
constexpr int kRecSplit = 9;

__host__ __device__ inline size_t site_lds_layout(int K, int tcap, size_t* off) {
  const size_t K4 = (size_t)((K + 3) & ~3);
  size_t o = 0;
  off[0] = o; o += 16 * 2 * K4;        // bis   float4 [2][K4]
  off[1] = o;                          // (unused)
  off[2] = o;                          // (unused)
  off[3] = o; o += 16 * 4 * 64;        // face float4 [4][64]
  off[4] = o; o += 8 * (size_t)tcap * 64;  // tr uint2 [tcap][64]
  off[5] = o; o += 4 * kQueue;         // queue u32 [kQueue]
  off[6] = o; o += 4 * 8;              // hist  unsigned [8]
  off[7] = o; o += (size_t)(4 + K) * 64;  // nx u8 [4+K][64]
  return (o + 15) & ~(size_t)15;
}

template <int MF, int ACC>
__global__ __launch_bounds__(64) void k_clip_site(ClipArgs a) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int lane = threadIdx.x;
  // bisectors in the site tables: a pair that needs more is recomputed from
  // scratch through the retry list (the same result, at the full depth)
  const int K = a.site_depth > 0 ? min(a.site_depth, a.K) : a.K;
  const bool deeper = K < a.K;
  const int K4 = (K + 3) & ~3;
  size_t off[8];
  site_lds_layout(K, a.tcap, off);
  float4* bis = (float4*)(smem + off[0]);
  unsigned* queue = (unsigned*)(smem + off[5]);
  unsigned* hist = (unsigned*)(smem + off[6]);
  Cell<PlaneStoreShared, TriCached> c;
  c.ps.bind((float4*)(smem + off[3]), lane);
  c.ps.bis = bis;
  c.ts.tr = (uint2*)(smem + off[4]);
  c.ts.lane = lane;
  c.nx = smem + off[7];
  c.lane = lane;
  c.pcap = a.pcap;
  c.tcap = a.tcap;
  c.filter = a.filter;
  if (lane < 8) hist[lane] = 0;

  // wave-uniform scheduling state
  int site0 = -1, site1 = -1;         // site of each table slot
  int feed_slot = 1;                  // slot of the site being fed into the queue
  unsigned cursor = 0, feed_end = 0;  // its pairs not yet through the emptiness test
  unsigned qhead = 0, qtail = 0;      // queue ring positions
  unsigned qsplit = 0;                // entries before qsplit belong to the other slot
  bool no_more_sites = false;

  // lane state
  int lst = kIdle;
  int pv = 0, pslot = 0;
  unsigned ppair = 0, ptet = 0;

  // record slots reserved by this wave, per size class: [rec_next, rec_end)
  unsigned rec_next0 = 0, rec_end0 = 0, rec_next1 = 0, rec_end1 = 0;  // wave-uniform

  for (;;) {
    // (1) refill the queue up to 64 entries
    while (qtail - qhead < 64) {
      if (cursor < feed_end) {
        __syncthreads();  // queue entries read by the previous trip
        const unsigned i = cursor + lane;
        const bool live = i < feed_end;
        const unsigned pair = live ? a.pair_ids[i] : 0u;
        const unsigned long long mask = __ballot(live);
        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
        if (live) queue[(qtail + (unsigned)below) % kQueue] = pair;
        qtail += (unsigned)__popcll(mask);
        cursor += 64;
        continue;
      }
      if (no_more_sites) break;
      // the other slot is free once no lane and no queued entry refers to it
      const int other = feed_slot ^ 1;
      if (__ballot(lst != kIdle && pslot == other) != 0 || qhead < qsplit) break;
      int s = 0;
      if (lane == 0) s = (int)atomicAdd(a.site_counter, 1u);
      s = __builtin_amdgcn_readfirstlane(s);
      if (s >= a.n_sites) {
        no_more_sites = true;
        break;
      }
      if (a.site_order) s = a.site_order[s];
      const unsigned pb = a.pair_start[s], pe = a.pair_count ? pb + a.pair_count[s] : a.pair_start[s + 1];
      if (pb == pe) continue;
      const P4 sp = site_at(a, s);
      const int32_t* nbr = a.site_knn + (size_t)s * a.K;
      __syncthreads();  // the slot's last readers are done
      float4* eb = bis + other * K4;
      for (int v = lane; v < K; v += 64) {
        const int z = nbr[v];
        if ((unsigned)z >= (unsigned)a.n_sites) {  // clipping stops at such a neighbour
          const float nan = __int_as_float(0x7fc00000);
          eb[v] = make_float4(nan, nan, nan, nan);
          continue;
        }
        const P4 eq = bisector(sp, site_at(a, z));
        eb[v] = make_float4(eq.x, eq.y, eq.z, eq.w);
      }
      __syncthreads();
      if (other) site1 = s; else site0 = s;
      feed_slot = other;
      cursor = pb;
      feed_end = pe;
      qsplit = qtail;
    }
    __syncthreads();

    // (2) idle lanes take queued pairs in queue order
    const unsigned long long idle = __ballot(lst == kIdle);
    const unsigned avail = qtail - qhead;
    const unsigned rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0));
    if (lst == kIdle && rank < avail) {
      // the entries before qsplit were queued for the other slot's site
      const unsigned qp = qhead + rank;
      ppair = queue[qp % kQueue];
      pslot = qp < qsplit ? feed_slot ^ 1 : feed_slot;
      pv = 0;
      lst = kClipping;
      ptet = ppair / (unsigned)a.TK;
      c.ps.bis = bis + pslot * K4;
      c.init_tet_prep(site_at(a, pslot ? site1 : site0), a.tet_prep + kTetPrepWords * (size_t)ptet);
    }
    qhead += min((unsigned)__popcll(idle), avail);
    const int n_clip = __popcll(__ballot(lst == kClipping));
    if (n_clip == 0) {
      if (no_more_sites && cursor >= feed_end && qhead == qtail) break;
      continue;
    }
    // Result of a finished pair: retry entry, or accumulation and records.
    auto finish = [&](const PairSums<ACC>& S, bool acc) {
      const int status = c.status;
      if (status == kTriangleOverflow && a.retry_list && a.tcap < a.tcap_full) {
        // this launch's vertex cap is below the configured one: the pair is
        // recomputed from scratch at the configured caps by the retry pass
        const unsigned slot = atomicAdd(a.retry_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.retry_list[slot] = ppair;
      } else if (status == kSecurityRadiusNotReached && a.srnr_list) {
        // clipped again with deeper neighbour lists
        const unsigned slot = atomicAdd(a.srnr_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.srnr_list[slot] = ppair;
      } else {
        if (acc) accumulate<ACC>(a, pslot ? site1 : site0, S);
        if (a.pair_status) a.pair_status[ppair] = (uint8_t)status;
        if (a.pair_vals) {
          double* pvals = a.pair_vals + 4 * (size_t)ppair;
          pvals[0] = acc ? (double)S.v : 0.0;
          pvals[1] = acc ? (double)S.x : 0.0;
          pvals[2] = acc ? (double)S.y : 0.0;
          pvals[3] = acc ? (double)S.z : 0.0;
        }
        atomicAdd(&hist[status], 1u);
      }
      lst = kIdle;
    };

    // Volume and moments of this lane's polytope, in place, in the reference's
    // FP32 decomposition.  The robust decomposition's FP64 moments stay out of
    // this kernel (their registers would cost it resident waves): such a pair
    // goes to the retry list instead, whose pass computes the same result.
    auto moments_now = [&]() {
      if constexpr (MF == 0) {
        PairSums<ACC> S;
        c.template moments_ref<ACC>(S);
        bool acc = false;
        // vertices formed by parallel planes (degenerate input) give
        // non-finite sums; such a pair is not accumulated
        if (!(isfinite((double)S.v) && isfinite((double)S.x) && isfinite((double)S.y) && isfinite((double)S.z)))
          c.status = kNeedsExactPredicates;
        else if (fabs((double)S.v) < 0.1)
          c.status = kNoIntersection;
        else
          acc = true;
        finish(S, acc);
      } else {
        const unsigned slot = atomicAdd(a.retry_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.retry_list[slot] = ppair;
        lst = kIdle;
      }
    };

    bool defer = false;
    if (lst == kClipping) {
      // (3) kernel A's loop (voronoi.cu:679-700), a few bisectors per trip: a
      // lane tests its next bisectors until one removes a vertex of its
      // polytope (at most kTestsPerTrip of them), then the cuts of all lanes
      // run together, so that the costly cut is not run for one lane while
      // the others wait
      bool cut = false;
      uint32_t conf = 0;
      float cut_d2z = 0.f;
      // The loop's exits after the clip of bisector pv: 0 go on, 1 finished
      // (the volume pass follows), 2 ended without one, 3 recomputed deeper.
      auto step_end = [&](int sec, float d2z) -> int {
        if (sec >= 0 ? sec != 0 : c.security_reached(d2z)) return 1;
        if (c.status != kSuccess) return 2;
        if (++pv == K) {
          if (deeper) return 3;
          c.status = kSecurityRadiusNotReached;
          return 1;
        }
        return 0;
      };
      int out = 0;
      if (K == 0) {
        c.status = kSecurityRadiusNotReached;
        out = 1;
      } else if (!a.filter) {
        // the trip's bisectors tested as one batch (test_batch), then their
        // outcomes taken in list order as the tests one at a time take them
        P4 e[kBatchTests];
        float d2z[kBatchTests];
        uint32_t cf[kBatchTests];
#pragma unroll
        for (int k = 0; k < kBatchTests; ++k) {
          // |z - s|^2 from the bisector's normal s - z, the same bits as
          // from z - s; NaN marks an out-of-range neighbour
          e[k] = ld4(c.ps.bis, min(pv + k, K - 1));
          d2z[k] = dot3(e[k], e[k]);
          cf[k] = 0u;
        }
        const bool capped = c.nv >= c.pcap;
        if (!capped) c.test_batch(e, d2z, min(kBatchTests, K - pv), cf);
#pragma unroll
        for (int k = 0; k < kBatchTests; ++k) {
          if (out != 0 || cut) continue;
          const float d = d2z[k];
          if (d != d) {
            c.status = kInconsistentBoundary;
            out = 2;
          } else if (capped) {
            c.status = kVertexOverflow;
            out = step_end(-1, d);
          } else if (cf[k] != 0u) {
            cut = true;
            cut_d2z = d;
            conf = cf[k];
            ++c.nv;
            c.nr = 0;
          } else {
            out = step_end(d > 4.f * c.vmax ? 1 : 0, d);
          }
        }
      } else {
        // (a wave-uniform loop with the lanes that stop predicated off: the
        // same tests as a per-lane loop with breaks, and cheaper control flow)
        for (int it = 0; it < kTestsPerTrip; ++it) {
          const bool testing = out == 0 && !cut;
          if (__ballot(testing) == 0) break;
          if (testing) {
            // |z - s|^2 from the bisector's normal s - z, the same bits as
            // from z - s; NaN marks an out-of-range neighbour
            const P4 e = ld4(c.ps.bis, pv);
            const float d2z = dot3(e, e);
            if (d2z != d2z) {
              c.status = kInconsistentBoundary;
              out = 2;
            } else {
              const int r = c.clip_test(e, d2z, conf);
              if (r == 2) {
                cut = true;
                cut_d2z = d2z;
              } else {
                out = step_end(r, d2z);
              }
            }
          }
        }
      }
      if (cut) out = step_end(c.clip_cut(4 + pv, conf, cut_d2z), cut_d2z);
      if (out == 3) {
        const unsigned slot = atomicAdd(a.retry_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.retry_list[slot] = ppair;
        lst = kIdle;
      } else if (out == 1 && c.status != kNoIntersection) {
        // A polytope whose cut overflowed this launch's vertex cap may go to
        // the retry list after its volume is known: that is decided here, in
        // place, so that the list is final when this kernel ends (the re-clip
        // that reads it runs beside the volume pass, not after it).
        if (c.status == kTriangleOverflow && a.retry_list && a.tcap < a.tcap_full)
          moments_now();  // the robust form: retried without the volume test
        else
          defer = true;
      } else if (out != 0) {
        PairSums<ACC> S;
        finish(S, false);
      }
    }
    // (4) records of the polytopes for the volume pass, one slot per
    // deferring lane, from slots the wave reserves kRecChunk at a time (one
    // global counter shared by every wave would serialise them)
    const unsigned long long dm = __ballot(defer);
    if (dm != 0) {
      unsigned slot = ~0u;
      if (a.rec_meta) {
        const unsigned long long dm1 = __ballot(defer && c.nt > kRecSplit);
        const unsigned long long dm0 = dm & ~dm1;
        // slot of this lane among the lanes of mask m (m wave-uniform)
        auto reserve = [&](unsigned long long m, unsigned& next, unsigned& end) {
          if (m == 0) return;
          const unsigned need = (unsigned)__popcll(m);
          const unsigned rk =
              (unsigned)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
          const unsigned have = end - next;
          const bool mine = (m >> lane) & 1ull;
          if (have >= need) {
            if (mine) slot = next + rk;
            next += need;
          } else {
            const int leader = __builtin_ctzll(m);
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(a.rec_count, (unsigned)kRecChunk);
            base = __shfl(base, leader);
            if (mine) slot = rk < have ? next + rk : base + (rk - have);
            next = base + (need - have);
            end = base + (unsigned)kRecChunk;
          }
        };
        reserve(dm0, rec_next0, rec_end0);
        reserve(dm1, rec_next1, rec_end1);
      }
      if (defer) {
        if (a.rec_meta && slot < a.rec_cap) {
          a.rec_meta[slot] = make_uint4(ppair, (unsigned)(pslot ? site1 : site0), ptet,
                                        (unsigned)c.nt | ((unsigned)c.status << 8));
          uint4* dst = (uint4*)(a.rec_tris + (size_t)slot * rec_stride(a.tcap));
          // all LDS reads of a group of 16 first, then the stores (words past
          // nt are not read by the volume pass)
          for (int t0 = 0; t0 < c.nt; t0 += 16) {
            uint32_t w[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) w[j] = c.tri(min(t0 + j, c.tcap - 1));
#pragma unroll
            for (int j = 0; j < 16; j += 4)
              if (t0 + j < c.nt) dst[(t0 + j) >> 2] = make_uint4(w[j], w[j + 1], w[j + 2], w[j + 3]);
          }
          lst = kIdle;
        } else {
          moments_now();  // past the record capacity
        }
      }
    }
  }
  // reserved record slots left unused are marked for the volume pass
  if (a.rec_meta) {
    for (unsigned r = rec_next0 + lane; r < rec_end0 && r < a.rec_cap; r += 64) a.rec_meta[r] = make_uint4(~0u, 0u, 0u, 0u);
    for (unsigned r = rec_next1 + lane; r < rec_end1 && r < a.rec_cap; r += 64) a.rec_meta[r] = make_uint4(~0u, 0u, 0u, 0u);
  }
  __syncthreads();
  if (a.status_hist && lane < 7 && hist[lane]) atomicAdd(a.status_hist + lane, (unsigned long long)hist[lane]);
}

// Emptiness test of the site-major form as a pass of its own, before the
// clipping: one wavefront per site builds the site's first `depth` bisectors
// as the clipping's tables hold them, tests the site's pairs 64 at a time
// (all four tet corners beyond one of them by more than the FP32 rounding
// bound plus kCullMargin, with the early exit once the tet's nearest corner
// is within |z - s| / 2 of the seed) and compacts the survivors, in their
// order, to the front of the site's pair range; pair_count[s] gets their
// number.  Small and without polytope state, it runs at full occupancy, where
// inside the clipping kernel the same test ran at under two waves per SIMD.
constexpr int kCullMaxDepth = 256;
constexpr int kXcds = 8;  // MI355X: accelerator complex dies, one L2 each

// Pairs per lane per trip of the emptiness pass: one (a lane's loop runs
// until every pair it holds is decided, so two per lane left more lanes idle:
// 0.05 ms slower per iteration at 1M tets, four 0.12 ms).
constexpr int kCullPairs = 1;
__device__ __forceinline__ bool any_open(const bool* o) {
  bool r = false;
#pragma unroll
  for (int h = 0; h < kCullPairs; ++h) r = r || o[h];
  return r;
}

__global__ __launch_bounds__(64) void k_cull_site(ClipArgs a, int depth, unsigned* pair_ids, unsigned* pair_count) {
  extern __shared__ __align__(16) uint8_t smem[];
  float4* eb = (float4*)smem;   // [depth] bisector (n, w)
  float4* ec = eb + depth;      // [depth] |z - s|^2, |n|_1, |w|, kCullMargin * |n|
  const int lane = threadIdx.x;
  unsigned n_culled = 0;  // over the block's sites: one histogram atomic per block
  // Blocks go to the 8 XCDs in turn, each with an L2 of its own: the blocks of
  // one XCD take one contiguous stretch of the site order, so that the tets
  // that nearby sites share are fetched into the same L2.
  const int xcd = blockIdx.x % kXcds, per = gridDim.x / kXcds;
  const int span = (a.n_sites + kXcds - 1) / kXcds;
  const int s_begin = min(a.n_sites, xcd * span), s_end = min(a.n_sites, s_begin + span);
  for (int si = s_begin + (int)blockIdx.x / kXcds; si < s_end; si += per) {
    const int s = a.site_order ? a.site_order[si] : si;
    const unsigned pb = a.pair_start[s], pe = a.pair_start[s + 1];
    if (pb == pe) {
      if (lane == 0) pair_count[s] = 0;
      continue;
    }
    // Cell form: no cell (every pair of the site takes the cell's status), or
    // a cell whose clip by a skipped pair's tet could overflow the caps: a
    // culled pair's cell is cut by at most three of the tet's faces before
    // one empties it, which adds at most 3 planes and 6 triangles (twice the
    // planes less 4), and where that could reach a cap the reference reports
    // the overflow.  Every pair is kept.
    if (a.cell_status && ((a.cell_status[s] != kSuccess && a.cell_status[s] != kSecurityRadiusNotReached) ||
                          a.cell_nv[s] + 4 > a.pcap || a.cell_nt[s] + 6 >= a.tcap)) {
      if (lane == 0) pair_count[s] = pe - pb;
      continue;
    }
    const P4 sp = site_at(a, s);
    const int32_t* nbr = a.site_knn + (size_t)s * a.K;
    __syncthreads();  // the previous site's readers are done
    for (int v = lane; v < depth; v += 64) {
      const int z = nbr[v];
      if ((unsigned)z >= (unsigned)a.n_sites) {
        eb[v] = make_float4(0.f, 0.f, 0.f, 0.f);
        ec[v] = make_float4(kInvalidNbr, 0.f, 0.f, 0.f);
        continue;
      }
      const P4 zp = site_at(a, z);
      const P4 eq = bisector(sp, zp);
      const P4 d = sub3(zp, sp);
      const float dd = dot3(d, d);
      eb[v] = make_float4(eq.x, eq.y, eq.z, eq.w);
      ec[v] = make_float4(dd, fabsf(eq.x) + fabsf(eq.y) + fabsf(eq.z), fabsf(eq.w), kCullMargin * __builtin_amdgcn_sqrtf(dd));
    }
    __syncthreads();
    unsigned w = pb;
    // kCullPairs pairs per lane per trip (c0 + lane, c0 + 64 + lane, ...)
    for (unsigned c0 = pb; c0 < pe; c0 += 64 * kCullPairs) {
      unsigned pair[kCullPairs];
      bool live[kCullPairs], culled[kCullPairs] = {}, open[kCullPairs];
      P4 v[kCullPairs][4] = {};
      float pmax[kCullPairs] = {}, vmin2[kCullPairs] = {};
#pragma unroll
      for (int h = 0; h < kCullPairs; ++h) {
        const unsigned i = c0 + 64 * h + lane;
        live[h] = i < pe;
        pair[h] = live[h] ? pair_ids[i] : 0u;
      }
#pragma unroll
      for (int h = 0; h < kCullPairs; ++h)
        if (live[h]) tet_corners(a, (int)(pair[h] / (unsigned)a.TK), v[h]);
#pragma unroll
      for (int h = 0; h < kCullPairs; ++h) {
        open[h] = live[h];
        if (live[h]) {
          vmin2[h] = INFINITY;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            pmax[h] = fmaxf(pmax[h], fmaxf(fabsf(v[h][j].x), fmaxf(fabsf(v[h][j].y), fabsf(v[h][j].z))));
            const P4 d = sub3(v[h][j], sp);
            vmin2[h] = fminf(vmin2[h], dot3(d, d));
          }
        }
      }
      for (int q = 0; q < depth && any_open(open); ++q) {
        const float4 f = ec[q];
        const float4 e = eb[q];
        // rounding bound of the plane evaluation (relative kCullRound of
        // |n|_1 * max|x| + |w|: three roundings of partial sums bounded by
        // that) plus the required depth beyond the plane (kCullMargin in
        // distance, times |n|)
        // (as selects: a pair's loads are in flight together and every lane
        // runs the same instructions)
#pragma unroll
        for (int h = 0; h < kCullPairs; ++h) {
          // an invalid neighbour ends the clipping there; and once the tet's
          // nearest corner is within |z - s| / 2 of the seed it is on the
          // seed's side of this bisector and of every later one (the
          // neighbour distances do not decrease)
          const bool stop = f.x == kInvalidNbr || 4.f * vmin2[h] <= f.x;
          const float lim = -fmaf(kCullRound, fmaf(f.y, pmax[h], f.z), f.w);
          float hi = -INFINITY;
#pragma unroll
          for (int k = 0; k < 4; ++k)
            hi = fmaxf(hi, fmaf(e.z, v[h][k].z, fmaf(e.y, v[h][k].y, fmaf(e.x, v[h][k].x, e.w))));
          const bool cut = open[h] && !stop && hi < lim;
          culled[h] = culled[h] || cut;
          open[h] = open[h] && !stop && !cut;
        }
      }
#pragma unroll
      for (int h = 0; h < kCullPairs; ++h) {
        if (culled[h] && a.pair_status) a.pair_status[pair[h]] = (uint8_t)kNoIntersection;
        if (culled[h] && a.pair_vals) {
          double* pvals = a.pair_vals + 4 * (size_t)pair[h];
          pvals[0] = pvals[1] = pvals[2] = pvals[3] = 0.0;
        }
        n_culled += (unsigned)__popcll(__ballot(culled[h]));
        // survivors to the front, in order: every position written here has
        // been read (w never passes the pairs read so far)
        const bool keep = live[h] && !culled[h];
        const unsigned long long mask = __ballot(keep);
        const unsigned below =
            (unsigned)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
        if (keep) pair_ids[w + below] = pair[h];
        w += (unsigned)__popcll(mask);
      }
    }
    if (lane == 0) pair_count[s] = w - pb;
  }
  if (lane == 0 && a.status_hist && n_culled) atomicAdd(a.status_hist + kNoIntersection, (unsigned long long)n_culled);
}

// Volume pass of the site-major form: one lane per record.  The planes are
// rebuilt from their ids (faces from the tet's corners, bisectors from the
// seed's neighbour row), the sums run over the triangles in the order they
// were stored, and the pair's outcome follows as in the in-place pass:
// non-finite sums flag needs_exact_predicates, a volume below 0.1
// no_intersection; then a vertex-cap overflow goes to the retry list and
// security_radius_not_reached to the deeper-list re-clip when they are on,
// anything else is accumulated.
#ifndef VD_MOMENTS_BLOCK
#define VD_MOMENTS_BLOCK 128
#endif
constexpr int kMomentsBlock = VD_MOMENTS_BLOCK;

// Sum over the lanes of mask m (wave-uniform) of a per-lane value, in every
// lane; 64-bit integers add exactly, so the order does not matter.
__device__ __forceinline__ long long wave_sum_ll(long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Accumulation of a wave's pairs: lanes with acc set add S to site s.  Lanes
// of one site are summed across the wave first (a wave's records mostly
// belong to one or two sites), so each site gets one set of atomics per wave
// instead of one per lane.  The FP64 mode's per-pair values are rounded to
// the fixed point first, as accumulate<1> does, and integer sums are exact,
// so the site sums are the same bit for bit; the FP32 mode's order of
// additions varies anyway.
template <int ACC>
__device__ __forceinline__ void accumulate_wave(const ClipArgs& a, bool acc, int s, const PairSums<ACC>& S) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(acc);
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const int ks = __shfl(s, leader);
    const bool mine = acc && s == ks && ((todo >> lane) & 1ull);
    todo &= ~__ballot(mine);
    if constexpr (ACC == 1) {
      const long long v = wave_sum_ll(mine ? __double2ll_rn(S.v * kFixScale) : 0ll);
      const long long x = wave_sum_ll(mine ? __double2ll_rn(S.x * kFixScale) : 0ll);
      const long long y = wave_sum_ll(mine ? __double2ll_rn(S.y * kFixScale) : 0ll);
      const long long z = wave_sum_ll(mine ? __double2ll_rn(S.z * kFixScale) : 0ll);
      if (lane == leader) {
        unsigned long long* p = a.acc_fix + 4 * (size_t)ks;
        atomicAdd(p + 0, (unsigned long long)v);
        atomicAdd(p + 1, (unsigned long long)x);
        atomicAdd(p + 2, (unsigned long long)y);
        atomicAdd(p + 3, (unsigned long long)z);
      }
    } else {
      const float v = wave_sum_f(mine ? S.v : 0.f);
      const float x = wave_sum_f(mine ? S.x : 0.f);
      const float y = wave_sum_f(mine ? S.y : 0.f);
      const float z = wave_sum_f(mine ? S.z : 0.f);
      if (lane == leader) {
        float* p = a.acc_f32 + 4 * (size_t)ks;
        unsafeAtomicAdd(p + 0, v);
        unsafeAtomicAdd(p + 1, x);
        unsafeAtomicAdd(p + 2, y);
        unsafeAtomicAdd(p + 3, z);
      }
    }
  }
}

// Bisector tables of the volume pass: per wave, up to kMomentsTables seeds
// of its 64 records (records come in runs of one seed) with their first
// kMomentsDepth bisectors; the records of further seeds rebuild theirs.
constexpr int kMomentsTables = 4;
constexpr int kMomentsDepth = 48;

template <int MF, int ACC>
__global__ __launch_bounds__(kMomentsBlock) void k_site_moments(ClipArgs a) {
  __shared__ float4 face_all[kMomentsBlock / 64][4][64];
  __shared__ float4 tbl_all[kMomentsBlock / 64][kMomentsTables][kMomentsDepth];
  __shared__ unsigned hist[8];
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < 8) hist[threadIdx.x] = 0;
  __syncthreads();
  Cell<PlaneStoreTable, TriRecord> c;
  c.ps.face = &face_all[threadIdx.x >> 6][0][0];
  c.ps.sites = a.sites;
  c.ps.n_sites = a.n_sites;
  c.ps.lane = lane;
  c.keep_srnr = false;
  float4(*tbl)[kMomentsDepth] = tbl_all[threadIdx.x >> 6];
  // records hold bisectors below the launch's table depth only
  const int depth = min(min(a.site_depth > 0 ? a.site_depth : a.K, a.K), kMomentsDepth);
  c.ps.depth = depth;
  const unsigned n = min(*a.rec_count, a.rec_cap);
  const int stride = rec_stride(a.tcap);
  // wave-uniform loop: every lane takes part in the accumulation's shuffles
  for (unsigned base = blockIdx.x * kMomentsBlock + (threadIdx.x & ~63u); base < n;
       base += gridDim.x * kMomentsBlock) {
    const unsigned i = base + lane;
    uint4 m = make_uint4(~0u, 0u, 0u, 0u);
    if (i < n) m = a.rec_meta[i];
    const bool rec = m.x != ~0u;  // a reserved slot left unused, or past the end
    const unsigned pair = m.x;
    const int s = rec ? (int)m.y : 0;
    // the bisector tables of the wave's first kMomentsTables seeds: the
    // seeds first, then every table entry's loads at once
    int slot = -1;
    {
      int kst[kMomentsTables];
      int ntab = 0;
      unsigned long long todo = __ballot(rec);
#pragma unroll
      for (int t = 0; t < kMomentsTables; ++t) {
        kst[t] = 0;
        if (!todo) continue;
        const int leader = __builtin_ctzll(todo);
        const int ks = __shfl(s, leader);
        const bool mine = rec && s == ks;
        todo &= ~__ballot(mine);
        if (mine) slot = t;
        kst[t] = ks;
        ntab = t + 1;
      }
      constexpr int kRounds = (kMomentsTables * kMomentsDepth + 63) / 64;
      int z[kRounds], ks[kRounds], tv[kRounds];
#pragma unroll
      for (int r = 0; r < kRounds; ++r) {
        const int ent = r * 64 + lane, t = depth > 0 ? ent / depth : kMomentsTables;
        tv[r] = t < ntab ? ent : -1;
        int k = kst[0];
#pragma unroll
        for (int j = 1; j < kMomentsTables; ++j) k = t == j ? kst[j] : k;
        ks[r] = k;
        z[r] = tv[r] >= 0 ? a.site_knn[(size_t)k * a.K + (ent - t * depth)] : -1;
      }
#pragma unroll
      for (int r = 0; r < kRounds; ++r) {
        if (tv[r] < 0) continue;
        const int t = tv[r] / depth;
        const P4 p = (unsigned)z[r] >= (unsigned)a.n_sites ? P4{0.f, 0.f, 0.f, 0.f}
                                                           : bisector(site_at(a, ks[r]), site_at(a, z[r]));
        tbl[t][tv[r] - t * depth] = make_float4(p.x, p.y, p.z, p.w);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    PairSums<ACC> S;
    int status = kNoIntersection;
    bool acc = false;
    if (rec) {
      P4 v[4];
      tet_corners(a, (int)m.z, v);
      c.seed = site_at(a, s);
      c.ps.seed = c.seed;
      c.ps.tbl = slot >= 0 ? tbl[slot] : nullptr;
      c.ps.nbr = a.site_knn + (size_t)s * a.K;
      c.put_tet_faces(v);
      c.ts.tr = a.rec_tris + (size_t)i * stride;
      c.nt = (int)(m.w & 255u);
      status = (int)(m.w >> 8);
      if (MF == 0)
        c.template moments_ref<ACC>(S);
      else
        c.template moments_robust<ACC>(S, v, -1);
      if (!(isfinite((double)S.v) && isfinite((double)S.x) && isfinite((double)S.y) && isfinite((double)S.z)))
        status = kNeedsExactPredicates;
      else if (fabs((double)S.v) < 0.1)
        status = kNoIntersection;
      else
        acc = true;
      // then as the site-major kernel's finish step (which keeps the records
      // that could go to the retry list to itself)
      if (status == kTriangleOverflow && a.retry_list && a.tcap < a.tcap_full) {
        const unsigned slot = atomicAdd(a.retry_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.retry_list[slot] = pair;
        acc = false;
      } else if (status == kSecurityRadiusNotReached && a.srnr_list) {
        const unsigned slot = atomicAdd(a.srnr_count, 1u);
        if (slot < (unsigned)a.n_pairs) a.srnr_list[slot] = pair;
        acc = false;
      } else {
        if (a.pair_status) a.pair_status[pair] = (uint8_t)status;
        if (a.pair_vals) {
          double* pvals = a.pair_vals + 4 * (size_t)pair;
          pvals[0] = acc ? (double)S.v : 0.0;
          pvals[1] = acc ? (double)S.x : 0.0;
          pvals[2] = acc ? (double)S.y : 0.0;
          pvals[3] = acc ? (double)S.z : 0.0;
        }
        atomicAdd(&hist[status], 1u);
      }
    }
    accumulate_wave<ACC>(a, acc, s, S);
    // every lane is done with the tables before the next records' are built
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  if (a.status_hist && threadIdx.x < 7 && hist[threadIdx.x])
    atomicAdd(a.status_hist + threadIdx.x, (unsigned long long)hist[threadIdx.x]);
}

// ----------------------------------------------------- cell -> tet form ----
//
// Kernel B (voronoi.cu:540-575, 648-660): every site's Voronoi cell, the box
// [-0.1, 1000.1]^3 clipped by the bisectors of the site's neighbours with the
// security-radius exit; one lane per site; the cell (planes, triangles) is
// stored for kernel C with the squared radius of its farthest vertex.
// Kernel C (voronoi.cu:620-645, 703-726): for every (tet, site) pair, the
// stored cell clipped by the tet's four face planes, then volume and moments
// as in kernel A.  Site-major: a wavefront takes its sites one after another,
// loads each site's cell into one of two LDS slots and queues the site's
// pairs; every 64 queued pairs (of one or two sites) are clipped together.

// Kernel B's launches: every site at caps of at most kCellsFastPcap planes
// and kCellsFastTcap vertices (76 KB of LDS per wave, two waves per CU
// instead of one at 64/96), then the sites whose cell failed there (the
// redo list, about 4 in 100k seeded cells at 40/64) at the configured caps.
// A cell that succeeds at the smaller caps ran the same steps the configured
// caps run (a cap only ends a build), so it is the same cell; every other
// outcome is built again, since a runaway boundary walk's guard depends on
// the plane cap.
constexpr int kCellsFastPcap = 40;
constexpr int kCellsFastTcap = 64;

// pstride / tstride: the cell store's planes and triangles per site (the
// configured caps); redo_out: the sites whose cell failed are listed there
// ([0] count, then ids); redo_in: build only the listed sites.
template <int U>
__global__ __launch_bounds__(64) void k_cells(ClipArgs a, CellStore cs, int pstride, int tstride, unsigned* redo_out,
                                              const unsigned* redo_in) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int lane = threadIdx.x;
  // the plain clip with the vertices' distances stored: the security test
  // after every bisector reads them instead of solving every vertex again
  // (the cached clip's vertex masks hold 32 vertices; a cell has more)
  Cell<PlaneStoreLocal, TriPlainD2> c;
  c.ps.pl = (float4*)smem;
  c.ps.lane = lane;
  c.ts.tr = (uint2*)(c.ps.pl + (size_t)a.pcap * 64);
  c.ts.lane = lane;
  c.nx = (uint8_t*)(c.ts.tr + (size_t)a.tcap * 64);
  c.lane = lane;
  c.pcap = a.pcap;
  c.tcap = a.tcap;
  c.filter = a.filter;
  const int K = a.K;
  const int n_items = redo_in ? (int)min(redo_in[0], (unsigned)a.n_sites) : a.n_sites;
  for (int base = blockIdx.x * 64; base < n_items; base += gridDim.x * 64) {
    if (base + lane >= n_items) continue;
    const int s = redo_in ? (int)redo_in[1 + base + lane] : base + lane;
    const P4 sp = site_at(a, s);
    c.init_box(sp);
    const int32_t* nbr = a.site_knn + (size_t)s * K;
    bool stopped = false;
    for (int v = 0; v < K; ++v) {
      const int z = nbr[v];
      if ((unsigned)z >= (unsigned)a.n_sites) {  // defensive: never read outside the site array
        c.status = kInconsistentBoundary;
        break;
      }
      const P4 zp = site_at(a, z);
      const P4 eq = bisector(sp, zp);
      const P4 d = sub3(zp, sp);
      const float d2z = dot3(d, d);
      const int pid = c.nv;
      if (pid < a.pcap) c.ps.put(pid, eq);
      const int sec = c.clip(pid, eq, d2z);
      if (sec >= 0 ? sec != 0 : c.security_reached(d2z)) {
        stopped = true;
        break;
      }
      if (c.status != kSuccess) break;
    }
    if (!stopped && c.status == kSuccess) c.status = kSecurityRadiusNotReached;
    // store: header, planes, triangles, farthest vertex
    float r2 = 0.f;
    for (int t = 0; t < c.nt; ++t) r2 = fmaxf(r2, c.ts.d2(t));
    cs.status[s] = c.status;
    cs.nv[s] = min(c.nv, a.pcap);
    cs.nt[s] = c.nt;
    cs.r2[s] = r2;
    float4* pl = cs.planes + (size_t)s * pstride;
    for (int i = 0; i < min(c.nv, a.pcap); ++i) pl[i] = c.ps.pl[i * 64 + lane];
    uint32_t* tr = cs.tris + (size_t)s * tstride;
    for (int t = 0; t < c.nt; ++t) tr[t] = c.tri(t);
    if (redo_out && c.status != kSuccess && c.status != kSecurityRadiusNotReached)
      redo_out[1 + atomicAdd(redo_out, 1u)] = (unsigned)s;  // each site once: below n_sites entries
  }
}

// pstride / tstride: the cell store's planes and triangles per site; part 0
// takes the sites whose cell has fewer than split - 8 triangles and at most
// psplit - 4 planes (and the cells that were not built), part 1 the other
// built cells.  Four cuts add at most 4 planes and 8 triangles to a cell (a
// closed polytope of this form has twice its planes less 4 triangles), so
// part 0's pairs cannot reach caps of psplit / split, and part 0 runs at
// those smaller caps (less LDS per wave, more waves per CU) with the results
// of the configured ones.
template <int MF, int ACC>
__global__ __launch_bounds__(64) void k_cells_clip(ClipArgs a, CellStore cs, int pstride, int tstride, int split,
                                                   int psplit, int part) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int lane = threadIdx.x;
  // two cell slots, so that the pairs of consecutive sites share batches
  float4* cpl = (float4*)smem;                              // [2][pcap]
  uint32_t* ctr = (uint32_t*)(cpl + 2 * a.pcap);            // [2][tcap]
  uint32_t* tr = ctr + 2 * a.tcap;                          // [tcap][64]
  uint8_t* nx = (uint8_t*)(tr + (size_t)a.tcap * 64);       // [pcap][64]
  int* surv = (int*)(nx + (size_t)a.pcap * 64);             // [128] queued pairs
  uint8_t* sslot = (uint8_t*)(surv + 128);                  // [128] their slots
  float4* fl = (float4*)(((uintptr_t)(sslot + 128) + 15) & ~(uintptr_t)15);  // [4][64]
  Cell<PlaneStoreCell, TriPlain> c;
  c.ps.fl = fl;
  c.ps.lane = lane;
  c.ts.tr = tr;
  c.ts.lane = lane;
  c.nx = nx;
  c.lane = lane;
  c.pcap = a.pcap;
  c.tcap = a.tcap;
  c.filter = a.filter;
  // wave-uniform: the site of each slot, its cell status and counts, and the
  // queued pairs of each slot
  // (scalars, not arrays: a slot is picked at run time, and an indexed array
  // would go to scratch memory)
  int site0 = 0, site1 = 0, cst0 = kSuccess, cst1 = kSuccess, cnv0 = 0, cnv1 = 0, cnt0 = 0, cnt1 = 0;
  int queued = 0, in_q0 = 0, in_q1 = 0;
  int slot = 1;  // slot of the current site

  // Clips the queued pairs [0, n) (n <= 64), one per lane, each with its
  // slot's cell; the rest of the queue moves to the front.
  auto clip_batch = [&](int n) {
    const bool live = lane < n;
    const unsigned pair = live ? (unsigned)surv[lane] : 0u;
    const int sl = live ? (int)sslot[lane] : 0;
    const int rest = lane + 64 < queued ? surv[lane + 64] : 0;
    const uint8_t rest_sl = lane + 64 < queued ? sslot[lane + 64] : (uint8_t)0;
    __syncthreads();
    if (lane + 64 < queued) {
      surv[lane] = rest;
      sslot[lane] = rest_sl;
    }
    __syncthreads();
    in_q0 -= __popcll(__ballot(live && sl == 0));
    in_q1 -= __popcll(__ballot(live && sl == 1));
    queued -= n;
    const int s = sl ? site1 : site0;
    const int cst = sl ? cst1 : cst0;
    const int cnv = sl ? cnv1 : cnv0;
    const int cnt = sl ? cnt1 : cnt0;
    const uint32_t* ct = ctr + sl * a.tcap;
    int status = kNoIntersection;
    if (live) {
      P4 v[4];
      tet_corners(a, (int)(pair / (unsigned)a.TK), v);
      c.ps.cpl = cpl + sl * a.pcap;
      c.ps.nvc = cnv;
      c.seed = site_at(a, s);
      c.status = cst;
      c.keep_srnr = true;
      c.nv = cnv;
      c.nt = cnt;
      c.first = kNoPlane;
      for (int t = 0; t < cnt; ++t) c.ts.put(t, ct[t], 0.f);
      static constexpr int face[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const P4 eq = tri2plane(v[face[f][0]], v[face[f][1]], v[face[f][2]]);
        const int pid = c.nv;
        if (pid < a.pcap) c.ps.put(pid, eq);
        (void)c.clip(pid, eq, 0.f);
        if (c.status == kNoIntersection) break;
      }
      PairSums<ACC> S;
      bool acc = false;
      if (c.status != kNoIntersection) {
        if (MF == 0)
          c.template moments_ref<ACC>(S);
        else
          c.template moments_robust<ACC>(S, v, cnv);  // planes below cnv: the cell's
        if (!(isfinite((double)S.v) && isfinite((double)S.x) && isfinite((double)S.y) && isfinite((double)S.z)))
          c.status = kNeedsExactPredicates;
        else if (fabs((double)S.v) < 0.1)
          c.status = kNoIntersection;
        else
          acc = true;
      }
      status = c.status;
      if (acc) accumulate<ACC>(a, s, S);
      if (a.pair_status) a.pair_status[pair] = (uint8_t)status;
      if (a.pair_vals) {
        double* pv = a.pair_vals + 4 * (size_t)pair;
        pv[0] = acc ? (double)S.v : 0.0;
        pv[1] = acc ? (double)S.x : 0.0;
        pv[2] = acc ? (double)S.y : 0.0;
        pv[3] = acc ? (double)S.z : 0.0;
      }
    }
    count_status(a, status, live);
  };

  for (int s = blockIdx.x; s < a.n_sites; s += gridDim.x) {
    const unsigned pb = a.pair_start[s], pe = a.pair_count ? pb + a.pair_count[s] : a.pair_start[s + 1];
    if (pb == pe) continue;
    const int cst = cs.status[s];
    const int cnv = cs.nv[s], cnt = cs.nt[s];
    const float cr2 = cs.r2[s];
    const P4 sp = site_at(a, s);
    const bool built = cst == kSuccess || cst == kSecurityRadiusNotReached;
    if ((part == 0) != (!built || (cnt + 8 < split && cnv + 4 <= psplit))) continue;  // the other launch's site
    if (!built) {
      // the cell could not be built: every pair of the site takes its status
      for (unsigned i = pb + lane; i < pe; i += 64) {
        const unsigned pair = a.pair_ids[i];
        if (a.pair_status) a.pair_status[pair] = (uint8_t)cst;
        if (a.pair_vals) {
          double* pv = a.pair_vals + 4 * (size_t)pair;
          pv[0] = pv[1] = pv[2] = pv[3] = 0.0;
        }
      }
      count_status(a, cst, false);
      if (a.status_hist && lane == 0) atomicAdd(a.status_hist + cst, (unsigned long long)(pe - pb));
      continue;
    }
    // the other slot takes this site's cell once no queued pair refers to it
    slot ^= 1;
    if ((slot ? in_q1 : in_q0) > 0) clip_batch(queued);  // (queued < 64 here)
    __syncthreads();
    for (int i = lane; i < cnv; i += 64) cpl[slot * a.pcap + i] = cs.planes[(size_t)s * pstride + i];
    for (int t = lane; t < cnt; t += 64) ctr[slot * a.tcap + t] = cs.tris[(size_t)s * tstride + t];
    __syncthreads();
    if (slot) {
      site1 = s;
      cst1 = cst;
      cnv1 = cnv;
      cnt1 = cnt;
    } else {
      site0 = s;
      cst0 = cst;
      cnv0 = cnv;
      cnt0 = cnt;
    }
    // the cell lies in the ball of radius sqrt(cr2) around the site; a tet
    // whose bounding ball around its centroid misses that ball by more than
    // the margin has an empty intersection with the cell (skipped only where
    // the clip could not overflow the caps first, as in cull_sites)
    const float cr = sqrtf(cr2) * (1.f + 1e-5f) + kCullMargin;
    const bool may_skip = cnv + 4 <= a.pcap && cnt + 6 < a.tcap;
    for (unsigned base = pb; base < pe; base += 64) {
      const unsigned i = base + lane;
      const bool live = i < pe;
      const unsigned pair = live ? a.pair_ids[i] : 0u;
      bool culled = false;
      if (live && a.cull_depth > 0 && may_skip) {
        P4 v[4];
        tet_corners(a, (int)(pair / (unsigned)a.TK), v);
        P4 g = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          g.x += 0.25f * v[j].x;
          g.y += 0.25f * v[j].y;
          g.z += 0.25f * v[j].z;
        }
        float tr2 = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const P4 d = sub3(v[j], g);
          tr2 = fmaxf(tr2, dot3(d, d));
        }
        const P4 d = sub3(g, sp);
        const float gap = sqrtf(dot3(d, d)) - sqrtf(tr2) * (1.f + 1e-5f);
        culled = gap > cr;
        if (culled && a.pair_status) a.pair_status[pair] = (uint8_t)kNoIntersection;
        if (culled && a.pair_vals) {
          double* pv = a.pair_vals + 4 * (size_t)pair;
          pv[0] = pv[1] = pv[2] = pv[3] = 0.0;
        }
      }
      count_status(a, kNoIntersection, culled);
      const bool keep = live && !culled;
      const unsigned long long mask = __ballot(keep);
      const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
      if (keep) {
        surv[queued + below] = (int)pair;
        sslot[queued + below] = (uint8_t)slot;
      }
      queued += __popcll(mask);
      if (slot) in_q1 += __popcll(mask);
      else in_q0 += __popcll(mask);
      __syncthreads();
      if (queued >= 64) clip_batch(64);
    }
  }
  if (queued > 0) clip_batch(queued);
}

// Pairs grouped by seed: a counting sort over sites.  A block takes a run of
// kGroupPairs consecutive pairs (neighbouring tets share most candidates) and
// counts them per site in an LDS hash table, then adds each site's count to
// the global counter once; the scatter pass reserves one range per (block,
// site) the same way.  Sites that do not fit the table go through global
// atomics directly.  The order inside a site's list is unspecified.
#ifndef VD_GROUP_PAIRS
#define VD_GROUP_PAIRS 8192
#endif
#ifndef VD_GROUP_HASH_BITS
#define VD_GROUP_HASH_BITS 11
#endif
constexpr int kGroupPairs = VD_GROUP_PAIRS;
constexpr int kHashBits = VD_GROUP_HASH_BITS;
constexpr int kHashSlots = 1 << kHashBits;
constexpr int kEmptySlot = -1;

__device__ __forceinline__ unsigned hash_slot(int site) { return ((unsigned)site * 2654435761u) >> (32 - kHashBits); }

__device__ int hash_insert(int* keys, int site) {
  unsigned h = hash_slot(site);
  for (int probe = 0; probe < kHashSlots; ++probe) {
    const int k = keys[h];
    if (k == site) return (int)h;
    if (k == kEmptySlot) {
      const int prev = atomicCAS(&keys[h], kEmptySlot, site);
      if (prev == kEmptySlot || prev == site) return (int)h;
    }
    h = (h + 1) & (kHashSlots - 1);
  }
  return -1;
}

__global__ __launch_bounds__(256) void k_pairs_count(const int32_t* tet_knn, long long n_pairs, int n_sites,
                                                      int* count) {
  constexpr int kPerThread = kGroupPairs / 256;
  static_assert(kGroupPairs % 256 == 0, "count group");
  __shared__ int keys[kHashSlots];
  __shared__ int vals[kHashSlots];
  // this thread's pairs' sites first, all their loads in flight at once (a
  // load per trip behind the previous trip's table atomics waited out the
  // full memory latency every trip)
  const long long base = (long long)blockIdx.x * kGroupPairs;
  int sv[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const long long p = base + k * 256 + threadIdx.x;
    sv[k] = p < n_pairs ? tet_knn[p] : -1;
  }
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
    keys[i] = kEmptySlot;
    vals[i] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int s = sv[k];
    if ((unsigned)s >= (unsigned)n_sites) continue;
    const int h = hash_insert(keys, s);
    if (h >= 0)
      atomicAdd(&vals[h], 1);
    else
      atomicAdd(count + s, 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x)
    if (keys[i] != kEmptySlot) atomicAdd(count + keys[i], vals[i]);
}

// The scatter pass takes kScatterPairs pairs per block: it ranks each pair
// within its site's slot of the block's hash table, reserves one global range
// per slot, places the pair ids in LDS in slot order and writes them out from
// there, so that consecutive lanes write consecutive words of a site's range
// (a store of one tet's 64 candidates would touch 64 cache lines).
#ifndef VD_SCATTER_PAIRS
#define VD_SCATTER_PAIRS 4096
#endif
constexpr int kScatterPairs = VD_SCATTER_PAIRS;
constexpr int kScatterThreads = 256;
constexpr int kScatterPerThread = kScatterPairs / kScatterThreads;
static_assert(kScatterPairs % kScatterThreads == 0 && kScatterPairs <= 65536, "scatter group");
static_assert(kHashSlots % kScatterThreads == 0 && kHashSlots <= 65536, "scatter hash table");

__global__ __launch_bounds__(kScatterThreads) void k_pairs_scatter(const int32_t* tet_knn, long long n_pairs,
                                                                   int n_sites, const unsigned* start, int* cursor,
                                                                   unsigned* pair_ids) {
  constexpr int kSlotsPerThread = kHashSlots / kScatterThreads;
  __shared__ int keys[kHashSlots];         // site of the slot; after the reservation, its global
                                           // offset minus its local base (mod 2^32)
  __shared__ unsigned cnt[kHashSlots];     // pairs of the slot; after the scan, its local base
  __shared__ unsigned stage[kScatterPairs];       // pair ids in slot order
  __shared__ uint16_t stage_slot[kScatterPairs];  // their slots
  unsigned* wsum = stage;  // the waves' count sums while the slots are scanned
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long base = (long long)blockIdx.x * kScatterPairs;
  // this thread's pairs' sites first, all their loads in flight at once
  int sv[kScatterPerThread];
#pragma unroll
  for (int k = 0; k < kScatterPerThread; ++k) {
    const long long p = base + k * kScatterThreads + tid;
    sv[k] = p < n_pairs ? tet_knn[p] : -1;
  }
  for (int i = tid; i < kHashSlots; i += kScatterThreads) {
    keys[i] = kEmptySlot;
    cnt[i] = 0;
  }
  __syncthreads();
  // slot (high 16 bits) and rank within it (low 16) of this thread's pairs
  uint32_t code[kScatterPerThread];
#pragma unroll
  for (int k = 0; k < kScatterPerThread; ++k) {
    const long long p = base + k * kScatterThreads + tid;
    code[k] = ~0u;
    const int s = sv[k];
    if ((unsigned)s >= (unsigned)n_sites) continue;
    const int h = hash_insert(keys, s);
    if (h >= 0) {
      code[k] = ((uint32_t)h << 16) | (uint32_t)atomicAdd(&cnt[h], 1u);
    } else {
      pair_ids[start[s] + (unsigned)atomicAdd(cursor + s, 1)] = (unsigned)p;
    }
  }
  __syncthreads();
  // exclusive scan of the slot counts (kSlotsPerThread consecutive slots per
  // thread), then one global reservation per occupied slot
  unsigned c[kSlotsPerThread], sum = 0;
#pragma unroll
  for (int k = 0; k < kSlotsPerThread; ++k) {
    c[k] = cnt[tid * kSlotsPerThread + k];
    sum += c[k];
  }
  unsigned incl = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  unsigned run = incl - sum, total = 0;
#pragma unroll
  for (int w = 0; w < kScatterThreads / 64; ++w) {
    if (w < wv) run += wsum[w];
    total += wsum[w];
  }
#pragma unroll
  for (int k = 0; k < kSlotsPerThread; ++k) {
    const int j = tid * kSlotsPerThread + k;
    const int key = keys[j];
    cnt[j] = run;
    if (key != kEmptySlot && c[k] != 0u) {
      const unsigned off = start[key] + (unsigned)atomicAdd(cursor + key, (int)c[k]);
      keys[j] = (int)(off - run);
    }
    run += c[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScatterPerThread; ++k) {
    if (code[k] == ~0u) continue;
    const unsigned h = code[k] >> 16;
    const unsigned j = cnt[h] + (code[k] & 0xffffu);
    stage[j] = (unsigned)(base + k * kScatterThreads + tid);
    stage_slot[j] = (uint16_t)h;
  }
  __syncthreads();
  for (unsigned j = tid; j < total; j += kScatterThreads) pair_ids[(unsigned)keys[stage_slot[j]] + j] = stage[j];
}

__global__ __launch_bounds__(256) void k_srnr_rows(const unsigned* list, const unsigned* n_list, long long cap,
                                                    const int32_t* tet_knn, int n_sites, int* row_of, int* row_site,
                                                    unsigned* n_rows) {
  const long long n = min((long long)*n_list, cap);
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    const int s = tet_knn[list[j]];
    if ((unsigned)s >= (unsigned)n_sites) continue;
    if (atomicCAS(row_of + s, -1, -2) == -1) {
      const unsigned r = atomicAdd(n_rows, 1u);
      row_site[r] = s;
      row_of[s] = (int)r;
    }
  }
}

// Lloyd update (voronoi.cu:728-749): site = moment / volume where volume != 0.
__global__ void k_update_f32(float4* sites, int n, const float* acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = acc[4 * i];
  if (v != 0.f) {
    float4 p = sites[i];
    p.x = acc[4 * i + 1] / v;
    p.y = acc[4 * i + 2] / v;
    p.z = acc[4 * i + 3] / v;
    sites[i] = p;
  }
}
__global__ void k_update_fix(float4* sites, int n, const unsigned long long* acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = (double)(long long)acc[4 * i] / kFixScale;
  if (v != 0.0) {
    float4 p = sites[i];
    p.x = (float)(((double)(long long)acc[4 * i + 1] / kFixScale) / v);
    p.y = (float)(((double)(long long)acc[4 * i + 2] / kFixScale) / v);
    p.z = (float)(((double)(long long)acc[4 * i + 3] / kFixScale) / v);
    sites[i] = p;
  }
}
// Per-site sums as FP64 (vol, mx, my, mz), for readback.
__global__ void k_moments_out(int n, int acc_mode, const float* acc_f32, const unsigned long long* acc_fix,
                              double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * n) return;
  out[i] = acc_mode == 0 ? (double)acc_f32[i] : (double)(long long)acc_fix[i] / kFixScale;
}

}  // namespace

size_t clip_plain_lds(int pcap, int tcap) { return (size_t)64 * (16 * pcap + 4 * tcap + pcap); }
size_t clip_plain_cached_lds(int pcap, int tcap) { return (size_t)64 * (16 * pcap + 8 * tcap + pcap); }
size_t clip_site_lds(int K, int tcap) {
  size_t off[8];
  return site_lds_layout(K, tcap, off);
}

template <class Kern>
static hipError_t launch_lds(Kern k, dim3 grid, dim3 block, size_t lds, hipStream_t st, const ClipArgs& a) {
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, grid, block, lds, st, a);
  return hipGetLastError();
}

// The site-major kernel is persistent (waves take sites from a counter): by
// default one workgroup per slot the device can keep resident.
template <class Kern>
static hipError_t launch_site(Kern k, int n_blocks, size_t lds, hipStream_t st, const ClipArgs& a) {
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (n_blocks <= 0) {
    int per_cu = 0, dev = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, lds)) != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    n_blocks = std::max(1, per_cu * cus);
  }
  n_blocks = std::max(1, std::min(n_blocks, a.n_sites));
  hipLaunchKernelGGL(k, dim3((unsigned)n_blocks), dim3(64), lds, st, a);
  return hipGetLastError();
}

hipError_t clip_launch(const ClipArgs& a, int form, int moments, int acc_mode, int n_blocks, hipStream_t st,
                       bool volume_pass) {
  if (a.n_pairs <= 0) return hipSuccess;
  // the tet-major and site-major kernels start every pair from its tet's record
  if ((form == 0 || form == kFormTetCached || form == 1) && !a.tet_prep) return hipErrorInvalidValue;
  if (form == 0) {
    const size_t lds = clip_plain_lds(a.pcap, a.tcap);
    const dim3 grid(a.pair_list ? (unsigned)n_blocks : (unsigned)((a.n_pairs + 63) / 64)), block(64);
    if (moments == 0 && acc_mode == 0) return launch_lds(k_clip_plain<0, 0, 0>, grid, block, lds, st, a);
    if (moments == 0 && acc_mode == 1) return launch_lds(k_clip_plain<0, 1, 0>, grid, block, lds, st, a);
    if (moments == 1 && acc_mode == 0) return launch_lds(k_clip_plain<1, 0, 0>, grid, block, lds, st, a);
    return launch_lds(k_clip_plain<1, 1, 0>, grid, block, lds, st, a);
  }
  if (form == kFormTetCached) {
    if (a.tcap > 32) return hipErrorInvalidValue;
    const size_t lds = clip_plain_cached_lds(a.pcap, a.tcap);
    const dim3 grid(a.pair_list ? (unsigned)n_blocks : (unsigned)((a.n_pairs + 63) / 64)), block(64);
    if (moments == 0 && acc_mode == 0) return launch_lds(k_clip_plain<0, 0, 1>, grid, block, lds, st, a);
    if (moments == 0 && acc_mode == 1) return launch_lds(k_clip_plain<0, 1, 1>, grid, block, lds, st, a);
    if (moments == 1 && acc_mode == 0) return launch_lds(k_clip_plain<1, 0, 1>, grid, block, lds, st, a);
    return launch_lds(k_clip_plain<1, 1, 1>, grid, block, lds, st, a);
  }
  if (!a.site_counter) return hipErrorInvalidValue;
  if (a.rec_meta && (!a.rec_tris || !a.rec_count)) return hipErrorInvalidValue;
  // the robust moments are never computed in the site-major kernel: the
  // polytopes it cannot hand to the volume pass go to the retry list
  if (moments == 1 && (!a.retry_list || !a.retry_count)) return hipErrorInvalidValue;
  if (a.site_depth > 0 && a.site_depth < a.K && (!a.retry_list || !a.retry_count)) return hipErrorInvalidValue;
  const size_t lds = clip_site_lds(a.site_depth > 0 ? std::min(a.site_depth, a.K) : a.K, a.tcap);
  hipError_t e;
  if (moments == 0 && acc_mode == 0)
    e = launch_site(k_clip_site<0, 0>, n_blocks, lds, st, a);
  else if (moments == 0 && acc_mode == 1)
    e = launch_site(k_clip_site<0, 1>, n_blocks, lds, st, a);
  else if (moments == 1 && acc_mode == 0)
    e = launch_site(k_clip_site<1, 0>, n_blocks, lds, st, a);
  else
    e = launch_site(k_clip_site<1, 1>, n_blocks, lds, st, a);
  if (e != hipSuccess || !volume_pass) return e;
  return site_volume_pass(a, moments, acc_mode, st);
}

#ifndef VD_MOMENTS_GRID
#define VD_MOMENTS_GRID 8192
#endif
hipError_t site_volume_pass(const ClipArgs& a, int moments, int acc_mode, hipStream_t st) {
  if (a.n_pairs <= 0 || !a.rec_meta || a.rec_cap == 0) return hipSuccess;
  // the record count is read on the device
  const dim3 grid((unsigned)std::min<long long>(((long long)a.rec_cap + kMomentsBlock - 1) / kMomentsBlock, VD_MOMENTS_GRID)),
      block(kMomentsBlock);
  if (moments == 0 && acc_mode == 0)
    hipLaunchKernelGGL((k_site_moments<0, 0>), grid, block, 0, st, a);
  else if (moments == 0 && acc_mode == 1)
    hipLaunchKernelGGL((k_site_moments<0, 1>), grid, block, 0, st, a);
  else if (moments == 1 && acc_mode == 0)
    hipLaunchKernelGGL((k_site_moments<1, 0>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_site_moments<1, 1>), grid, block, 0, st, a);
  return hipGetLastError();
}

hipError_t tet_prep(const float4* tet_xyz, long long n, float4* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tet_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tet_xyz, n, out);
  return hipGetLastError();
}

hipError_t cull_sites(const ClipArgs& a, unsigned* pair_ids, unsigned* pair_count, hipStream_t st) {
  if (a.n_sites <= 0) return hipSuccess;
  const int K = a.site_depth > 0 ? std::min(a.site_depth, a.K) : a.K;
  const int depth = std::min(a.cull_depth, K);
  if (depth <= 0 || depth > kCullMaxDepth) return hipErrorInvalidValue;
  // a multiple of the XCD count, so that every XCD gets the same number of blocks
#ifndef VD_CULL_BLOCKS
#define VD_CULL_BLOCKS 32768
#endif
  const unsigned blocks =
      (unsigned)(kXcds * std::max(1, std::min((a.n_sites + kXcds - 1) / kXcds, VD_CULL_BLOCKS / kXcds)));
  hipLaunchKernelGGL(k_cull_site, dim3(blocks), dim3(64), 32 * (size_t)depth, st, a, depth, pair_ids, pair_count);
  return hipGetLastError();
}

hipError_t srnr_rows(const unsigned* list, const unsigned* n_list, long long cap, const int32_t* tet_knn, int n_sites,
                     int* row_of, int* row_site, unsigned* n_rows, hipStream_t st) {
  hipLaunchKernelGGL(k_srnr_rows, dim3(256), dim3(256), 0, st, list, n_list, cap, tet_knn, n_sites, row_of, row_site,
                     n_rows);
  return hipGetLastError();
}

size_t cells_lds(int pcap, int tcap) { return (size_t)64 * (16 * pcap + 8 * tcap + pcap); }
size_t cells_clip_lds(int pcap, int tcap) {
  return (size_t)32 * pcap + 8 * (size_t)tcap + (size_t)64 * (4 * tcap + pcap) + 5 * 128 + 16 + 16 * 4 * 64;
}

// Workgroups (one wave each) of kernel C: 8,192 at 1M tets / 100k sites
// (mean clipping ms over iterations 1-3: 1,024 → 28.3, 4,096 → 28.1, 8,192
// → 27.4-27.6, 16,384 → 27.5; contiguous stretches of the grid order per
// wave instead of every 8,192nd site: 27.4-27.6).
#ifndef VD_CELLS_CLIP_BLOCKS
#define VD_CELLS_CLIP_BLOCKS 8192
#endif
// Caps of kernel C's first launch, which takes the cells of fewer than 38
// triangles and at most 32 planes: 36 planes and 46 vertices, 20,352 bytes of
// LDS per wave, so that 8 waves share a CU (mean clipping ms of the cell form
// over iterations 1-3 at 1M tets, the rest at 64/96: vertex cap 40 → 22.0,
// 44 → 19.8, 45 → 19.2, 46 → 19.1, 48 → 20.7 at 64 planes (7 waves); 36/46 →
// 17.8, 40/44 → 19.6, 40/46 → 19.5; 27.6 in one launch at 64/96).
#ifndef VD_CELLS_CLIP_FAST_TCAP
#define VD_CELLS_CLIP_FAST_TCAP 46
#endif
constexpr int kCellsClipFastTcap = VD_CELLS_CLIP_FAST_TCAP;
#ifndef VD_CELLS_CLIP_FAST_PCAP
#define VD_CELLS_CLIP_FAST_PCAP 36
#endif
constexpr int kCellsClipFastPcap = VD_CELLS_CLIP_FAST_PCAP;
constexpr int kCellsClipBlocks = VD_CELLS_CLIP_BLOCKS;

hipError_t cells_launch(const ClipArgs& a0, const CellStore& cs, int moments, int acc_mode, hipStream_t st,
                        unsigned* pair_ids, unsigned* cull_count, int fast_pcap, int fast_tcap) {
  if (a0.n_sites <= 0) return hipSuccess;
  size_t lds = cells_lds(a0.pcap, a0.tcap);
  hipError_t e = hipFuncSetAttribute((const void*)k_cells<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  ClipArgs ab = a0;
  ab.pcap = std::min(a0.pcap, fast_pcap > 0 ? fast_pcap : kCellsFastPcap);
  ab.tcap = std::min(a0.tcap, fast_tcap > 0 ? fast_tcap : kCellsFastTcap);
  const bool redo = cs.redo && (ab.pcap < a0.pcap || ab.tcap < a0.tcap);
  if (!redo) ab = a0;
  if (redo && (e = fill_u32(cs.redo, 1, 0u, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_cells<0>, dim3((unsigned)((a0.n_sites + 63) / 64)), dim3(64), cells_lds(ab.pcap, ab.tcap), st, ab,
                     cs, a0.pcap, a0.tcap, redo ? cs.redo : nullptr, (const unsigned*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (redo) {
    // the listed sites at the configured caps (64 workgroups, looping over
    // the list when it is longer)
    hipLaunchKernelGGL(k_cells<0>, dim3(64), dim3(64), lds, st, a0, cs, a0.pcap, a0.tcap, (unsigned*)nullptr,
                       (const unsigned*)cs.redo);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (a0.n_pairs <= 0) return hipSuccess;
  ClipArgs a = a0;
  if (a.cull_depth > 0 && pair_ids && cull_count) {
    // A tet beyond one of the site's first bisectors (by the emptiness test's
    // margin) misses the cell: kernel B clipped the cell by that bisector, or
    // stopped earlier with every vertex on the seed's side of it, so clipping
    // the cell by the tet leaves nothing either.  Kernel C takes the survivors.
    ClipArgs ac = a;
    ac.cell_status = cs.status;
    ac.cell_nv = cs.nv;
    ac.cell_nt = cs.nt;
    if ((e = cull_sites(ac, pair_ids, cull_count, st)) != hipSuccess) return e;
    a.pair_count = cull_count;
  }
  lds = cells_clip_lds(a.pcap, a.tcap);
  // the cells with fewer than kCellsClipFastTcap - 8 triangles and at most
  // kCellsClipFastPcap - 4 planes (most) at those caps, then the others at
  // the configured ones
  const int split = std::min(a.tcap, kCellsClipFastTcap);
  const int psplit = std::min(a.pcap, kCellsClipFastPcap);
  auto go = [&](auto k) -> hipError_t {
    hipError_t e2 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e2 != hipSuccess) return e2;
    // each wave takes every kCellsClipBlocks-th site
    const dim3 grid((unsigned)std::min(a.n_sites, kCellsClipBlocks));
    if (split < a.tcap || psplit < a.pcap) {
      ClipArgs af = a;
      af.tcap = split;
      af.pcap = psplit;
      hipLaunchKernelGGL(k, grid, dim3(64), cells_clip_lds(af.pcap, af.tcap), st, af, cs, a.pcap, a.tcap, split,
                         psplit, 0);
      if ((e2 = hipGetLastError()) != hipSuccess) return e2;
      hipLaunchKernelGGL(k, grid, dim3(64), lds, st, a, cs, a.pcap, a.tcap, split, psplit, 1);
    } else {
      hipLaunchKernelGGL(k, grid, dim3(64), lds, st, a, cs, a.pcap, a.tcap, INT_MAX, INT_MAX, 0);
    }
    return hipGetLastError();
  };
  if (moments == 0 && acc_mode == 0) return go(k_cells_clip<0, 0>);
  if (moments == 0 && acc_mode == 1) return go(k_cells_clip<0, 1>);
  if (moments == 1 && acc_mode == 0) return go(k_cells_clip<1, 0>);
  return go(k_cells_clip<1, 1>);
}

hipError_t pairs_by_site(const int32_t* tet_knn, long long n_pairs, int n_sites, int* count, unsigned* start,
                         int* cursor, unsigned* pair_ids, unsigned* scan_tmp, hipStream_t st) {
  const unsigned blocks = (unsigned)((n_pairs + kGroupPairs - 1) / kGroupPairs);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pairs_count, dim3(blocks), dim3(256), 0, st, tet_knn, n_pairs, n_sites, count);
  hipError_t e = scan_excl(count, nullptr, n_sites, start, cursor, scan_tmp, st);
  if (e != hipSuccess) return e;
  const unsigned sblocks = (unsigned)((n_pairs + kScatterPairs - 1) / kScatterPairs);
  hipLaunchKernelGGL(k_pairs_scatter, dim3(sblocks), dim3(kScatterThreads), 0, st, tet_knn, n_pairs, n_sites, start,
                     cursor, pair_ids);
  return hipGetLastError();
}

hipError_t lloyd_update(float4* sites, int n, int acc_mode, const float* acc_f32, const unsigned long long* acc_fix,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (acc_mode == 0)
    hipLaunchKernelGGL(k_update_f32, dim3((n + 255) / 256), dim3(256), 0, st, sites, n, acc_f32);
  else
    hipLaunchKernelGGL(k_update_fix, dim3((n + 255) / 256), dim3(256), 0, st, sites, n, acc_fix);
  return hipGetLastError();
}

hipError_t moments_out(int n, int acc_mode, const float* acc_f32, const unsigned long long* acc_fix, double* out,
                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_moments_out, dim3((4 * n + 255) / 256), dim3(256), 0, st, n, acc_mode, acc_f32, acc_fix,
                     out);
  return hipGetLastError();
}

}  // namespace vd
