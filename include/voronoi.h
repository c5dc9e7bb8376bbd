/* The reference's Lloyd driver (voronoi.h:78-84 of xh-liu-tech/3D-Voronoi-GPU)
 * with the same signature, as a header-only C++ shim over
 * vd_compute_clipped_voro_diagram (vd_api.h).
 *
 * Behaviour kept: sites updated in place in the caller's layout (SoA when
 * site_is_transposed), one record.csv row and one iter_<i>.xyz file per
 * iteration in the working directory, exit(1) on failure.  Behaviour changed:
 * sites keep their input order (the reference's grid path reorders them).
 * Environment: VD_FORM selects the formulation ("cells", the default, is the
 * reference's source of the new sites: each site's whole cell clipped by the
 * tets, voronoi.cu:1090-1109; "site" / "tet" the faster tet -> cell
 * formulation of voronoi.cu:662-701 in its two launch forms); VD_ACCUM "f64"
 * the FP64 accumulation (default the reference's FP32 atomics); VD_MOMENTS
 * "robust" the well-conditioned moment decomposition.
 */
#ifndef H_VORONOI_H
#define H_VORONOI_H

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "knearests.h"
#include "params.h"
#include "vd_api.h"
#include "voronoi_defs.h"

inline void compute_clipped_voro_diagram_GPU(const std::vector<float>& vertices, const std::vector<int>& indices,
                                             std::vector<float>& site, const int n_site, const bool site_is_transposed,
                                             int nb_Lloyd_iter = 1, int preferred_tet_k = 0) {
  vd_config cfg;
  vd_config_default(&cfg);
  cfg.form = VD_FORM_CELLS;
  const char* form = getenv("VD_FORM");
  if (form && !strcmp(form, "site")) cfg.form = VD_FORM_SITE_MAJOR;
  if (form && !strcmp(form, "tet")) cfg.form = VD_FORM_TET_MAJOR;
  const char* acc = getenv("VD_ACCUM");
  if (acc && !strcmp(acc, "f64")) cfg.accum = VD_ACCUM_F64;
  const char* mom = getenv("VD_MOMENTS");
  if (mom && !strcmp(mom, "robust")) cfg.moments = VD_MOMENTS_ROBUST;
  const int rc = vd_compute_clipped_voro_diagram(vertices.data(), (int)(vertices.size() / 3),
                                                 reinterpret_cast<const int32_t*>(indices.data()),
                                                 (int)(indices.size() / 4), site.data(), n_site,
                                                 site_is_transposed ? 1 : 0, nb_Lloyd_iter, preferred_tet_k, &cfg,
                                                 nullptr);
  if (rc != VD_OK) {
    fprintf(stderr, "compute_clipped_voro_diagram_GPU failed (%d)\n", rc);
    exit(1);
  }
}

#endif
