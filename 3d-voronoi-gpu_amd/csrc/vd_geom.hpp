// Device geometry for the clipping kernels (gfx950).
//
// Operation order and fusion are fixed: every fused multiply-add is an explicit
// fmaf() and the library is compiled with -ffp-contract=off, so an FP32 result
// is a deterministic function of its inputs and can be compared bit for bit with
// a host computation that follows the same expressions.  The formulas follow
// the reference's geometry (voronoi.cu:14-123, 219-249, 274-306, 532-537,
// 577-618); the contraction pattern is this library's own choice.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vd {

// Per-pair status values (voronoi_defs.h:4-12).
enum PairStatus : int32_t {
  kTriangleOverflow = 0,
  kVertexOverflow = 1,
  kInconsistentBoundary = 2,
  kSecurityRadiusNotReached = 3,
  kSuccess = 4,
  kNeedsExactPredicates = 5,
  kNoIntersection = 6
};

constexpr uint8_t kNoPlane = 255;
constexpr double kFixScale = 1048576.0;  // 2^20: per-site reduction in 64-bit fixed point

struct P4 {
  float x, y, z, w;
};

__device__ __forceinline__ float msub(float a, float b, float c, float d) { return fmaf(a, b, -(c * d)); }
__device__ __forceinline__ float dot3(const P4& a, const P4& b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ P4 sub3(const P4& a, const P4& b) { return {a.x - b.x, a.y - b.y, a.z - b.z, 0.f}; }
__device__ __forceinline__ P4 cross3(const P4& a, const P4& b) {
  return {msub(a.y, b.z, a.z, b.y), msub(a.z, b.x, a.x, b.z), msub(a.x, b.y, a.y, b.x), 0.f};
}
__device__ __forceinline__ float det2(float a11, float a12, float a21, float a22) { return msub(a11, a22, a12, a21); }
__device__ __forceinline__ float det3(float a11, float a12, float a13, float a21, float a22, float a23, float a31,
                                      float a32, float a33) {
  const float c1 = det2(a22, a23, a32, a33);
  const float c2 = det2(a12, a13, a32, a33);
  const float c3 = det2(a12, a13, a22, a23);
  return fmaf(a31, c3, msub(a11, c1, a21, c2));
}

// The four 3x3 minors of the reference's det4x4 (voronoi.cu:54-73) that do not
// involve the clipping plane; the conflict determinant of a vertex with plane e
// is then four more terms.
struct Minors {
  float m123, m124, m134, m234;
};
__device__ __forceinline__ Minors vertex_minors(const P4& p1, const P4& p2, const P4& p3) {
  const float m12 = msub(p1.y, p2.x, p1.x, p2.y);
  const float m13 = msub(p1.z, p2.x, p1.x, p2.z);
  const float m14 = msub(p1.w, p2.x, p1.x, p2.w);
  const float m23 = msub(p1.z, p2.y, p1.y, p2.z);
  const float m24 = msub(p1.w, p2.y, p1.y, p2.w);
  const float m34 = msub(p1.w, p2.z, p1.z, p2.w);
  Minors m;
  m.m123 = fmaf(m12, p3.z, msub(m23, p3.x, m13, p3.y));
  m.m124 = fmaf(m12, p3.w, msub(m24, p3.x, m14, p3.y));
  m.m134 = fmaf(m13, p3.w, msub(m34, p3.x, m14, p3.z));
  m.m234 = fmaf(m23, p3.w, msub(m34, p3.y, m24, p3.z));
  return m;
}
__device__ __forceinline__ float conflict_det(const Minors& m, const P4& e) {
  return fmaf(-m.m123, e.w, fmaf(m.m124, e.z, msub(m.m234, e.x, m.m134, e.y)));
}

// The same determinant in FP64 from the FP32 planes, and the error bounds of
// the reference's arithmetic filter (voronoi.cu:285-301 FP32, 325-341 FP64):
// C * max|x| * max|y| * max|z| * max(max|.|)^2 over the four planes.
__device__ __forceinline__ double msub64(double a, double b, double c, double d) { return fma(a, b, -(c * d)); }
__device__ __forceinline__ double conflict_det64(const P4& q1, const P4& q2, const P4& q3, const P4& qe) {
  const double p1x = q1.x, p1y = q1.y, p1z = q1.z, p1w = q1.w;
  const double p2x = q2.x, p2y = q2.y, p2z = q2.z, p2w = q2.w;
  const double p3x = q3.x, p3y = q3.y, p3z = q3.z, p3w = q3.w;
  const double m12 = msub64(p1y, p2x, p1x, p2y);
  const double m13 = msub64(p1z, p2x, p1x, p2z);
  const double m14 = msub64(p1w, p2x, p1x, p2w);
  const double m23 = msub64(p1z, p2y, p1y, p2z);
  const double m24 = msub64(p1w, p2y, p1y, p2w);
  const double m34 = msub64(p1w, p2z, p1z, p2w);
  const double m123 = fma(m12, p3z, msub64(m23, p3x, m13, p3y));
  const double m124 = fma(m12, p3w, msub64(m24, p3x, m14, p3y));
  const double m134 = fma(m13, p3w, msub64(m34, p3x, m14, p3z));
  const double m234 = fma(m23, p3w, msub64(m34, p3y, m24, p3z));
  return fma(-m123, (double)qe.w, fma(m124, (double)qe.z, msub64(m234, (double)qe.x, m134, (double)qe.y)));
}
__device__ __forceinline__ float filter_bound32(float mx, float my, float mz) {
  float eps = (float)(6.6876506e-05 * (double)mx * (double)my * (double)mz);
  const float mm = fmaxf(fmaxf(mx, my), mz);
  eps *= (mm * mm);
  return eps;
}
__device__ __forceinline__ double filter_bound64(double mx, double my, double mz) {
  double eps = 1.2466136531027298e-13 * mx * my * mz;
  const double mm = fmax(fmax(mx, my), mz);
  eps *= (mm * mm);
  return eps;
}

// Vertex where three planes meet (voronoi.cu:238-249).
__device__ __forceinline__ P4 vertex3(const P4& p1, const P4& p2, const P4& p3) {
  const float X = -det3(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z);
  const float Y = -det3(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z);
  const float Z = -det3(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w);
  const float W = det3(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
  return {X / W, Y / W, Z / W, 1.f};
}

__device__ __forceinline__ P4 tri2plane(const P4& v0, const P4& v1, const P4& v2) {
  P4 n = cross3(sub3(v1, v0), sub3(v2, v0));
  n.w = -dot3(v0, n);
  return n;
}

// Bisector plane of seed s against site z, positive on the seed's side
// (voronoi.cu:365-378).
__device__ __forceinline__ P4 bisector(const P4& s, const P4& z) {
  P4 p = sub3(s, z);
  const P4 mid2 = {s.x + z.x, s.y + z.y, s.z + z.z, 0.f};
  p.w = -dot3(mid2, p) / 2.f;
  return p;
}

// voronoi.cu:111-116 (1e-2 is a double literal in the reference).
__device__ __forceinline__ P4 project_ref(const P4& P, const P4& pl) {
  const float n2 = dot3(pl, pl);
  float lambda = 0.f;
  if ((double)n2 > 1e-2) lambda = (dot3(pl, P) + pl.w) / n2;
  return {fmaf(-lambda, pl.x, P.x), fmaf(-lambda, pl.y, P.y), fmaf(-lambda, pl.z, P.z), 1.f};
}
__device__ __forceinline__ P4 foot_on_plane(const P4& P, const P4& pl) {
  const float lambda = (dot3(pl, P) + pl.w) / dot3(pl, pl);
  return {fmaf(-lambda, pl.x, P.x), fmaf(-lambda, pl.y, P.y), fmaf(-lambda, pl.z, P.z), 1.f};
}
__device__ __forceinline__ P4 foot_on_line(const P4& P, const P4& X, const P4& u) {
  const float lambda = dot3(sub3(P, X), u) / dot3(u, u);
  return {fmaf(lambda, u.x, X.x), fmaf(lambda, u.y, X.y), fmaf(lambda, u.z, X.z), 1.f};
}

// Signed volume and barycenter of tet (A,B,C,D) (voronoi.cu:104-110).
__device__ __forceinline__ void tet_vol_bary(const P4& A, const P4& B, const P4& C, const P4& D, float& vol,
                                             P4& bary) {
  const P4 a = sub3(A, D), b = sub3(B, D), c = sub3(C, D);
  vol = -det3(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z) / 6.f;
  bary.x = 0.25f * (((A.x + B.x) + C.x) + D.x);
  bary.y = 0.25f * (((A.y + B.y) + C.y) + D.y);
  bary.z = 0.25f * (((A.z + B.z) + C.z) + D.z);
  bary.w = 1.f;
}

// ---- FP64 moments of the robust formulation ----
//
// The polytope's topology (which planes bound it, which triples meet at its
// vertices) comes from the FP32 clipping; its volume and first moment are then
// integrated in FP64 from planes rebuilt in FP64 from the FP32 inputs: a tet
// face from the tet's corners, a bisector from its FP32 normal n (= s - z, the
// subtraction of two nearby FP32 coordinates) and the seed s, as the plane
// through s - n/2 with normal n, so w = -s.n + n.n/2.  In the reference's
// [0,1000] frame FP32 offsets of planes through points ~500 units out lose
// about 1e-4 of a piece's volume; these planes do not.
struct D4 {
  double x, y, z, w;
};
__device__ __forceinline__ double dot3d(const D4& a, const D4& b) { return fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ D4 sub3d(const D4& a, const D4& b) { return {a.x - b.x, a.y - b.y, a.z - b.z, 0.0}; }
__device__ __forceinline__ D4 cross3d(const D4& a, const D4& b) {
  return {msub64(a.y, b.z, a.z, b.y), msub64(a.z, b.x, a.x, b.z), msub64(a.x, b.y, a.y, b.x), 0.0};
}
__device__ __forceinline__ D4 up4(const P4& p) { return {(double)p.x, (double)p.y, (double)p.z, (double)p.w}; }
__device__ __forceinline__ double det3d(double a11, double a12, double a13, double a21, double a22, double a23,
                                        double a31, double a32, double a33) {
  const double c1 = msub64(a22, a33, a23, a32);
  const double c2 = msub64(a12, a33, a13, a32);
  const double c3 = msub64(a12, a23, a13, a22);
  return fma(a31, c3, msub64(a11, c1, a21, c2));
}
__device__ __forceinline__ D4 face64(const P4& v0, const P4& v1, const P4& v2) {
  const D4 a = up4(v0);
  D4 n = cross3d(sub3d(up4(v1), a), sub3d(up4(v2), a));
  n.w = -dot3d(a, n);
  return n;
}
__device__ __forceinline__ D4 bisector64(const P4& pl, const P4& s) {
  D4 n = {(double)pl.x, (double)pl.y, (double)pl.z, 0.0};
  n.w = fma(0.5, dot3d(n, n), -dot3d(up4(s), n));
  return n;
}
__device__ __forceinline__ D4 vertex3d(const D4& p1, const D4& p2, const D4& p3) {
  const double X = -det3d(p1.w, p1.y, p1.z, p2.w, p2.y, p2.z, p3.w, p3.y, p3.z);
  const double Y = -det3d(p1.x, p1.w, p1.z, p2.x, p2.w, p2.z, p3.x, p3.w, p3.z);
  const double Z = -det3d(p1.x, p1.y, p1.w, p2.x, p2.y, p2.w, p3.x, p3.y, p3.w);
  const double r = 1.0 / det3d(p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z);
  return {X * r, Y * r, Z * r, 1.0};
}
__device__ __forceinline__ D4 foot_on_plane_d(const D4& P, const D4& pl) {
  const double lambda = (dot3d(pl, P) + pl.w) / dot3d(pl, pl);
  return {fma(-lambda, pl.x, P.x), fma(-lambda, pl.y, P.y), fma(-lambda, pl.z, P.z), 1.0};
}
__device__ __forceinline__ D4 foot_on_line_d(const D4& P, const D4& X, const D4& u) {
  const double lambda = dot3d(sub3d(P, X), u) / dot3d(u, u);
  return {fma(lambda, u.x, X.x), fma(lambda, u.y, X.y), fma(lambda, u.z, X.z), 1.0};
}
// Six times the signed volume of tet (A,B,C,D) and four times its barycenter:
// the 1/6 and 1/4 are applied once to a polytope's sums.
__device__ __forceinline__ void tet_vol6_bary4_d(const D4& A, const D4& B, const D4& C, const D4& D, double& vol6,
                                                 D4& bary4) {
  const D4 a = sub3d(A, D), b = sub3d(B, D), c = sub3d(C, D);
  vol6 = -det3d(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
  bary4.x = ((A.x + B.x) + C.x) + D.x;
  bary4.y = ((A.y + B.y) + C.y) + D.y;
  bary4.z = ((A.z + B.z) + C.z) + D.z;
  bary4.w = 1.0;
}

__device__ __forceinline__ P4 ld4(const float4* p, int i) {
  const float4 v = p[i];
  return {v.x, v.y, v.z, v.w};
}

// Per-pair accumulator: FP32 (the reference's) or FP64.
template <int ACC>
struct PairSums;
template <>
struct PairSums<0> {
  float v = 0.f, x = 0.f, y = 0.f, z = 0.f;
  __device__ __forceinline__ void add(float tv, const P4& b) {
    x = fmaf(tv, b.x, x);
    y = fmaf(tv, b.y, y);
    z = fmaf(tv, b.z, z);
    v += tv;
  }
  // a polytope's sums computed in FP64, rounded to FP32
  __device__ __forceinline__ void set(double tv, double tx, double ty, double tz) {
    v = (float)tv;
    x = (float)tx;
    y = (float)ty;
    z = (float)tz;
  }
};
template <>
struct PairSums<1> {
  double v = 0.0, x = 0.0, y = 0.0, z = 0.0;
  __device__ __forceinline__ void add(float tv, const P4& b) {
    x = fma((double)tv, (double)b.x, x);
    y = fma((double)tv, (double)b.y, y);
    z = fma((double)tv, (double)b.z, z);
    v += (double)tv;
  }
  __device__ __forceinline__ void set(double tv, double tx, double ty, double tz) {
    v = tv;
    x = tx;
    y = ty;
    z = tz;
  }
};

}  // namespace vd
