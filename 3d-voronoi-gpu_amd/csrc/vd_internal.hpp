// Engine functions shared with the job server (vd_serve.cpp); not part of the
// C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "vd_api.h"

namespace vdi {

// One problem resident in device memory: vertices, tets with vertex ids local
// to it, their corner records (3 float4 per tet), sites, and the mesh box
// (lo xyz, hi xyz).
struct JobDev {
  const float4* verts;
  int nv;
  const int4* tets;
  const float4* tet_xyz;
  int nt;
  float4* sites;
  int ns;
  float box[6];
  double fixed_load;  // fixed_point_load of the job's mesh
};

// The FP64 accumulation sums each site's (volume, volume * x, y, z) as 64-bit
// integers in units of 2^-20.  A site's sums are bounded by the mesh volume
// times the largest |coordinate| (every piece of its cell lies in one tet, at
// most one piece per tet), so with a 4x margin against pieces whose moments
// the reference decomposition overstates the sums stay exact while
// volume * max|coordinate| <= 2^41 (the [0,1000]^3 frame gives at most 1e12).
constexpr double kFixedPointMaxLoad = 2199023255552.0;  // 2^41
// Volume of the tets t0..t0+nt-1 times the largest |coordinate| of vertices
// v0..v0+nv-1 (host arrays; tets hold vertex ids relative to v0).
double fixed_point_load(const float* verts_xyz, int v0, int nv, const int32_t* tets, long long t0, int nt);

// Makes the context's problem the batch of jobs[0..n) (in that order), built
// by device copies from their buffers; the jobs' sites become the batch's.
int assemble(vd_ctx* ctx, const JobDev* jobs, int n);
// Copies the batch's current sites of the instances idx[0..n_idx) back to the
// buffers of the jobs they were assembled from.
int scatter_sites(vd_ctx* ctx, const JobDev* jobs, const int* idx, int n_idx);
hipStream_t stream(vd_ctx* ctx);
void set_error(vd_ctx* ctx, const char* msg);
// Vertex cap of the site-major launch under configuration c (its site_tcap,
// or the automatic choice).
int site_tcap_of(const vd_config& c);

}  // namespace vdi
