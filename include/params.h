/* Default sizes of the reference's build-time configuration (params.h:1-25 of
 * xh-liu-tech/3D-Voronoi-GPU, preset 0), kept under the same names for code
 * that refers to them.  In this library they are runtime settings
 * (vd_config in vd_api.h); these values are its defaults. */
#ifndef VD_PARAMS_H
#define VD_PARAMS_H

/* below this many sites the reference keeps sites in SoA layout (its
 * brute-force kNN path); the CLI keeps the same layout rule */
#define KNEARESTS_MIN_N 14000
#define _K_ 90      /* neighbours per site */
#define _MAX_P_ 64  /* plane cap of a polytope */
#define _MAX_T_ 96  /* vertex cap of a polytope */

#endif
