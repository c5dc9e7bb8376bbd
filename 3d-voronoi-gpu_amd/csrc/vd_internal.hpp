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
};

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
