/*
 * vd_api.h -- C ABI of libvoronoi_hip.so, the MI355X (gfx950) clipped-3D-Voronoi
 * Lloyd engine.
 *
 * Every entry point returns an int status (VD_OK or a negative VD_ERR_*) and
 * never exits the process; vd_ctx_error() gives the message of the last
 * failure.  A context owns its device buffers and one HIP stream, and is safe
 * to use from one host thread at a time.  Plain pointers and sizes only.
 *
 * Coordinates are those of the reference: the mesh and the sites live in the
 * [0,1000]^3 frame produced by the reference loader (test_voronoi.cu:14-38,
 * 103-118); several thresholds of the algorithm are absolute in that frame.
 *
 * Which reference interface each entry point replaces is given beside it
 * (files of xh-liu-tech/3D-Voronoi-GPU).
 */
#ifndef VD_API_H
#define VD_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define VD_OK 0
#define VD_ERR_ARG (-1)          /* bad argument */
#define VD_ERR_HIP (-2)          /* HIP runtime error */
#define VD_ERR_STATE (-3)        /* call out of order (e.g. no mesh yet) */
#define VD_ERR_NOMEM (-4)        /* device allocation failed */
#define VD_ERR_PRECONDITION (-5) /* input violates a documented precondition */

/* per-pair status values: identical to the reference's enum Status
 * (voronoi_defs.h:4-12) */
#define VD_TRIANGLE_OVERFLOW 0
#define VD_VERTEX_OVERFLOW 1
#define VD_INCONSISTENT_BOUNDARY 2
#define VD_SECURITY_RADIUS_NOT_REACHED 3
#define VD_SUCCESS 4
#define VD_NEEDS_EXACT_PREDICATES 5
#define VD_NO_INTERSECTION 6

/* per-site accumulation */
#define VD_ACCUM_F32 0 /* FP32 pair sums, FP32 atomics (the reference's, voronoi.cu:577-618) */
#define VD_ACCUM_F64 1 /* FP64 pair sums, exact 64-bit fixed-point site sums (2^-20), FP64 update */

/* volume / moment decomposition */
#define VD_MOMENTS_REFERENCE 0 /* six tets per vertex around the seed (voronoi.cu:532-537) */
#define VD_MOMENTS_ROBUST 1    /* same ring around the vertex mean, edge feet on the edge line */

/* launch form of the clipping kernel (same per-pair results) */
#define VD_FORM_TET_MAJOR 0  /* one lane per (tet, rank) pair, the reference's order */
#define VD_FORM_SITE_MAJOR 1 /* one wavefront per site, bisector planes shared in LDS */
#define VD_FORM_CELLS 2      /* the cell -> tet formulation: each site's whole cell from the box
                                (kernel B, voronoi.cu:540-575), then cell x tet per pair (kernel C,
                                voronoi.cu:620-645,703-726); the reference's source of out.xyz.
                                cull_depth > 0 skips pairs whose tet lies outside the cell's ball */

typedef struct vd_config {
  int site_k;         /* neighbours per site; reference _K_ = 90 (params.h:9) */
  int tet_k;          /* candidate sites per tet; 0 = reference heuristic (voronoi.cu:936) */
  int pcap;           /* plane cap per polytope; reference _MAX_P_ = 64 (params.h:11) */
  int tcap;           /* vertex cap per polytope; reference _MAX_T_ = 96 (params.h:12) */
  int accum;          /* VD_ACCUM_* */
  int moments;        /* VD_MOMENTS_* */
  int form;           /* VD_FORM_* */
  float grid_density; /* sites per kNN grid cell */
  int record_pairs;   /* keep per-pair status and values for vd_get_pairs */
  int clip_blocks;    /* workgroups of the site-major launch; 0 = automatic */
  int site_tcap;      /* vertex cap of the site-major launch, at most 32; pairs that
                         exceed it are recomputed at tcap, so results do not depend
                         on it; 0 = automatic: the largest (at most 32) whose LDS at
                         the configured table depth lets 9 waves share a CU, 17 at
                         site_depth 40 */
  int cull_depth;     /* site-major form: a pair whose tet lies entirely beyond one of the
                         seed's first cull_depth bisectors (by more than the FP32 rounding
                         bound plus 0.01) has an empty intersection and is not clipped
                         (at most min(pcap - 4, (tcap - 3) / 2) bisectors, so that such a
                         pair's clip could not have overflowed the caps first); cell form:
                         the same test against the built cell, for cells whose clip by a
                         tet could not overflow the caps; 0 = off */
  int srnr_k;         /* tet -> cell forms: 0 = the reference's behaviour (a pair whose cell
                         does not reach the security radius with site_k neighbours is
                         accumulated as it is, voronoi.cu:694-700); > site_k (at most 256):
                         such pairs are clipped again with their seed's srnr_k nearest
                         neighbours, and that result is the one accumulated */
  int filter;         /* arithmetic filter of the conflict test (USE_ARITHMETIC_FILTER,
                         params.h:32, off in the reference): 0 off; 1 the reference's
                         (voronoi.cu:285-301: a determinant below its FP32 error bound
                         flags needs_exact_predicates); 2 such a determinant is
                         re-evaluated in FP64 from the same planes, its sign decides and
                         only its FP64 bound (voronoi.cu:325-341) flags the pair */
  int site_records;   /* site-major form: polytopes whose clipping ended go to a separate
                         volume pass through a buffer of this many records (16 bytes plus 4
                         per vertex of the launch cap, rounded up to 4 vertices); past it a
                         lane computes its volume in place, so results do not depend on it;
                         0 = automatic (n_pairs / 10 + 262144), -1 = every volume in place */
  int site_depth;     /* site-major form: bisectors per site table of the launch; a pair that
                         needs more is recomputed from scratch at site_k by the retry pass,
                         so results do not depend on it; 0 = automatic (40, and 64 for the
                         first clipping after vd_set_sites or vd_set_batch: sites not yet
                         moved by an update need deeper tables) */
  int knn_waves;      /* wavefronts per workgroup of the kNN queries: 2, 4 or 8; 0 = automatic
                         (4); results do not depend on it */
  int cells_pcap;     /* cell form: plane cap of kernel B's first launch (8..pcap); cells that
                         fail there are built again at pcap/tcap, so results do not depend on
                         it; 0 = automatic (40: with cells_tcap 64, two waves per CU instead
                         of one at 64/96) */
  int cells_tcap;     /* cell form: vertex cap of kernel B's first launch (8..tcap), likewise;
                         0 = automatic (64) */
} vd_config;

typedef struct vd_timing {
  float knn_ms;    /* grid build + site kNN + candidate kNN */
  float pairs_ms;  /* grouping of pairs by site (site-major form) */
  float clip_ms;   /* clipping + per-site accumulation */
  float update_ms; /* Lloyd update */
  float total_ms;  /* whole iteration */
  float overflow_ms; /* part of clip_ms: re-clipping of the pairs that overflowed the
                        site-major launch's vertex cap, at the configured caps, and of the
                        pairs retried with srnr_k neighbours */
} vd_timing;

typedef struct vd_ctx vd_ctx;

/* Defaults: site_k 90, tet_k 0, pcap 64, tcap 96, VD_ACCUM_F32,
 * VD_MOMENTS_REFERENCE, VD_FORM_SITE_MAJOR, grid_density 3.1
 * (knearests.cu:243), record_pairs 0, clip_blocks 0, site_tcap 0, cull_depth 16, srnr_k 0,
 * filter 0, site_records 0, site_depth 0, knn_waves 0, cells_pcap 0, cells_tcap 0. */
void vd_config_default(vd_config* cfg);

/* Context on HIP device `device`.  Replaces the device selection and buffer
 * setup of compute_clipped_voro_diagram_GPU (voronoi.cu:933-986). */
int vd_ctx_create(vd_ctx** out, int device, const vd_config* cfg);
void vd_ctx_destroy(vd_ctx* ctx);
const char* vd_ctx_error(const vd_ctx* ctx);
int vd_ctx_set_config(vd_ctx* ctx, const vd_config* cfg);

/* Mesh: n_verts xyz triples and n_tets quadruples of 0-based vertex ids, every
 * tet positively oriented (first vertex on the positive side of face c,b,d).
 * Replaces copy_tet_data (voronoi.cu:881-923). */
int vd_set_mesh(vd_ctx* ctx, const float* verts_xyz, int n_verts, const int32_t* tets, int n_tets);
/* Sites, xyz triples in input order (the engine never reorders them).  After
 * vd_set_batch, n_sites must be the batch's total (all instances, in order). */
int vd_set_sites(vd_ctx* ctx, const float* sites_xyz, int n_sites);

/* A batch of n_inst independent problems (instances) solved together: every
 * step below runs once over all of them, and each instance's result is the one
 * it would get alone.  Instance i has n_verts[i] vertices, n_tets[i] tets whose
 * vertex ids are local to it (0-based), and n_sites[i] >= 1 sites; verts_xyz,
 * tets and sites_xyz hold the instances one after another.  All instances use
 * one K = min(site_k, min_i n_sites[i] - 1) and one TK, which must come out
 * equal for every instance (set tet_k, or vd_knn_step fails).  Replaces a
 * loop over compute_clipped_voro_diagram_GPU (voronoi.cu:925-1188), one mesh
 * per call.  Results (vd_get_sites, vd_get_moments) are the instances'
 * concatenated in order; vd_set_mesh ends the batch. */
int vd_set_batch(vd_ctx* ctx, int n_inst, const int32_t* n_verts, const int32_t* n_tets, const int32_t* n_sites,
                 const float* verts_xyz, const int32_t* tets, const float* sites_xyz);
int vd_get_batch(vd_ctx* ctx, int* n_inst);
int vd_get_sites(vd_ctx* ctx, float* sites_xyz);

/* One Lloyd iteration in three steps, or n_iter whole iterations.
 * vd_knn_step replaces kn_prepare+kn_solve / knn_cuda_global_dev and
 * compute_tet_knn_dev (voronoi.cu:995-1028, 825-879); vd_clip_step replaces
 * kernel A and its accumulation (voronoi.cu:1034-1076); vd_update_step replaces
 * compute_new_site (voronoi.cu:1121-1132). */
int vd_knn_step(vd_ctx* ctx);
int vd_clip_step(vd_ctx* ctx);
int vd_update_step(vd_ctx* ctx);
int vd_lloyd(vd_ctx* ctx, int n_iter);
/* n_iter whole iterations as replays of one HIP graph holding a complete
 * iteration (kNN, grouping, clipping, update): on the first call one eager
 * iteration sizes every buffer and the next one is captured.  Results are
 * those of vd_lloyd.  The graph is dropped by vd_set_mesh, vd_set_batch,
 * vd_ctx_set_config and by vd_set_sites with a new site count.  Replaces the
 * host loop over iterations of compute_clipped_voro_diagram_GPU
 * (voronoi.cu:988-1176); vd_get_timing reports only total_ms for replays. */
int vd_lloyd_graph(vd_ctx* ctx, int n_iter);
/* Chooses the launch parameters for the loaded problem: times `iters`
 * iterations (after one untimed) for candidate values of site_tcap,
 * cull_depth, grid_density and the block sizes (clip_blocks, the site-major
 * launch's workgroups; knn_waves, the kNN queries' waves per workgroup), one
 * parameter at a time, keeps the fastest configuration (written to *chosen
 * when given) and restores the sites.  These parameters change no per-pair
 * result, so in the VD_ACCUM_F64 mode the results are bit for bit those of
 * any other choice.  No counterpart in the reference (its launch shapes are
 * constants, voronoi.cu:1034-1076). */
int vd_autotune(vd_ctx* ctx, int iters, vd_config* chosen);
/* Chooses smaller neighbour and candidate counts than the configured site_k
 * and tet_k where they change the result by at most `tol`: from the loaded
 * sites it runs `iters` iterations with the configured counts, then with each
 * candidate tet_k (from 0.875 down to 0.3 of the configured one, at least 8)
 * and site_k (the same fractions, at least 16), and keeps the fastest
 * configuration whose sites after `iters` iterations differ from the
 * configured ones by at most tol * |site| (largest relative site displacement;
 * 0 demands bit-identical sites).  That error is written to *err, the
 * configuration to *chosen and into the context; the sites are restored.
 * The reference fixes these counts (params.h:4-25, voronoi.cu:936). */
int vd_autotune_k(vd_ctx* ctx, int iters, double tol, vd_config* chosen, double* err);
int vd_sync(vd_ctx* ctx);
/* Device times (HIP events on the context's stream) of the phases of the
 * iteration in progress, or of the last completed one. */
int vd_get_timing(vd_ctx* ctx, vd_timing* t);
/* Sum of the phase times of the iterations completed since vd_timing_reset
 * (at most the last 256), and their number. */
int vd_timing_reset(vd_ctx* ctx);
int vd_get_timing_sum(vd_ctx* ctx, vd_timing* t, int* n_iters);
/* Phase times of each of the last n completed iterations (n at most the number
 * completed since vd_timing_reset, and at most 256), oldest first.  The
 * reference appends the same per-iteration times to record.csv
 * (voronoi.cu:984-1175). */
int vd_get_timing_list(vd_ctx* ctx, vd_timing* t, int n);

/* Effective sizes: K = min(site_k, n_sites-1), TK = min(tet_k or heuristic, n_sites). */
int vd_get_dims(vd_ctx* ctx, int* k, int* tk, int* n_sites, int* n_tets);
/* Readback of the last kNN step: n_sites x K neighbour ids (original order,
 * nearest first, self excluded) and n_tets x TK candidate ids. */
int vd_get_site_knn(vd_ctx* ctx, int32_t* out);
int vd_get_tet_knn(vd_ctx* ctx, int32_t* out);
/* Per-site (volume, Σvol·x, Σvol·y, Σvol·z) of the last clip step. */
int vd_get_moments(vd_ctx* ctx, double* out);
/* Per-pair status and (vol, mx, my, mz) of the last clip step; needs
 * record_pairs.  Pair index = tet * TK + rank. */
int vd_get_pairs(vd_ctx* ctx, uint8_t* status, double* vals);
/* Pairs first .. first+count-1 of vd_get_pairs (either output may be NULL),
 * for results too large to read whole.  No counterpart in the reference, which
 * keeps per-pair statuses on the device (voronoi.cu:942,979). */
int vd_get_pairs_range(vd_ctx* ctx, long long first, long long count, uint8_t* status, double* vals);
/* Rows first_tet .. first_tet+n_tets-1 of vd_get_tet_knn. */
int vd_get_tet_knn_range(vd_ctx* ctx, int first_tet, int n_tets, int32_t* out);
/* Number of pairs per status value (7 counters) of the last clip step. */
int vd_get_status_hist(vd_ctx* ctx, uint64_t* hist7);

/* ---- kNN entry points of the reference's own surface -------------------- */

/* Exact k nearest neighbours on the device, replacing knn_cuda_global_dev
 * (extern/kNN-CUDA/knncuda.h:1-12, knncuda.cu:271-349).  ref and query are
 * dim x n row-pitched device arrays (pitches in elements); the result is
 * written to knn_index_dev[j * index_pitch + q], j = 0..k-1, nearest first,
 * ties broken by the lower reference index.  Only dim == 3 is supported.
 * `stream` is a hipStream_t or NULL. */
int vd_knn_dev(const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
               size_t query_pitch, int dim, int k, int* knn_index_dev, size_t index_pitch, void* stream);

/* The kNN of vd_knn_dev for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results as vd_knn_dev,
 * but it allocates nothing: its working memory (grid tables, points, queries)
 * is a static device array of 32 MiB per GPU, held until the process ends,
 * whose address is looked up once per device at the first call on it.  That
 * first call on a device also uploads a constant table synchronously, so make
 * it outside any graph capture; every later call only enqueues kernels on
 * `stream` and returns without synchronising.  At most VD_KNN_NOALLOC_MAX_REF
 * reference points and VD_KNN_NOALLOC_MAX_QUERY queries (VD_ERR_ARG beyond,
 * with a message on stderr).  All calls on one GPU share its array: calls that
 * overlap on one GPU (different streams or host threads without ordering)
 * race and give wrong results. */
#define VD_KNN_NOALLOC_MAX_REF 262144
#define VD_KNN_NOALLOC_MAX_QUERY 1048576
int vd_knn_dev_noalloc(const float* ref_dev, int ref_nb, size_t ref_pitch, const float* query_dev, int query_nb,
                       size_t query_pitch, int dim, int k, int* knn_index_dev, size_t index_pitch, void* stream);

/* Grid kNN problem, replacing kn_prepare / kn_solve / kn_free and the getters
 * (knearests.h:3-27, knearests.cu:230-432).  Points are host xyz triples; the
 * device copy is reordered by grid cell (deterministically: by cell, then by
 * input index).  vd_kn_solve fills k neighbours per point, self excluded,
 * as indices into the reordered array, nearest first. */
typedef struct vd_kn vd_kn;
typedef struct vd_kn_info {
  int allocated_points;
  int dimx, dimy, dimz;
  int num_cell_offsets;
  float* d_stored_points;      /* device, reordered xyz triples */
  unsigned int* d_knearests;   /* device, allocated_points x k */
  unsigned int* d_permutation; /* device, reordered position -> input index */
  int* d_counters;             /* device, points per cell */
  int* d_ptrs;                 /* device, first position per cell */
} vd_kn_info;
int vd_kn_prepare(vd_kn** out, const float* points_xyz, int n, int k);
int vd_kn_solve(vd_kn* kn);
void vd_kn_free(vd_kn* kn);
int vd_kn_info_get(const vd_kn* kn, vd_kn_info* info);
int vd_kn_get_points(vd_kn* kn, float* out_xyz);
int vd_kn_get_knearests(vd_kn* kn, unsigned int* out);
int vd_kn_get_permutation(vd_kn* kn, unsigned int* out);
/* Points per grid cell, dimx * dimy * dimz counts (x fastest), the data of
 * kn_print_stats (knearests.cu:434-460). */
int vd_kn_get_cell_counts(vd_kn* kn, int* out);

/* ---- job server: continuous batching of independent jobs ---------------- */

/* Jobs (mesh, sites, iteration count) wait in arrival order until their
 * estimated device footprint fits hbm_budget_bytes, then run together: every
 * vd_server_step runs one Lloyd iteration over all running jobs as one batch
 * (vd_set_batch semantics, assembled on the device from the jobs' own
 * buffers).  A job's result is the one it gets alone with the same config.
 * cfg->tet_k must be > 0 and every job needs more than site_k sites, so that
 * all jobs share K and TK.  No counterpart in the reference, which runs one
 * mesh per process (test_voronoi.cu:150-229). */
#define VD_JOB_PENDING 0
#define VD_JOB_RUNNING 1
#define VD_JOB_DONE 2
typedef struct vd_server vd_server;
int vd_server_create(vd_server** out, int device, const vd_config* cfg, size_t hbm_budget_bytes);
void vd_server_destroy(vd_server* s);
const char* vd_server_error(const vd_server* s);
/* Copies the job's data; tets use 0-based ids into its own vertices. */
int vd_server_submit(vd_server* s, const float* verts_xyz, int n_verts, const int32_t* tets, int n_tets,
                     const float* sites_xyz, int n_sites, int n_iter, int* job_id);
/* Admits what fits, runs one iteration of the running jobs, retires the
 * finished ones (synchronous). */
int vd_server_step(vd_server* s, int* n_running, int* n_pending);
/* State (VD_JOB_*), iterations done, time from submission to admission and
 * from admission to completion (or to now), in ms of host wall clock. */
int vd_server_job(vd_server* s, int job_id, int* state, int* iters_done, double* wait_ms, double* run_ms);
/* Final sites of a finished job (n_sites xyz triples). */
int vd_server_result(vd_server* s, int job_id, float* sites_xyz);
size_t vd_server_bytes_in_use(const vd_server* s);

/* ---- whole-program entry point ------------------------------------------ */

/* The reference's Lloyd driver compute_clipped_voro_diagram_GPU
 * (voronoi.h:78-84, voronoi.cu:925-1188) with plain pointers: vertices are
 * normalised xyz triples, `site` is updated in place (SoA when
 * site_is_transposed, else AoS).  Appends one row per iteration to
 * <out_dir>/record.csv and writes <out_dir>/iter_<i>.xyz, like the reference;
 * out_dir NULL or "" means the working directory.  cfg NULL = defaults. */
int vd_compute_clipped_voro_diagram(const float* vertices, int n_verts, const int32_t* indices, int n_tets,
                                    float* site, int n_site, int site_is_transposed, int nb_lloyd_iter,
                                    int preferred_tet_k, const vd_config* cfg, const char* out_dir);

#ifdef __cplusplus
}
#endif

#endif /* VD_API_H */
