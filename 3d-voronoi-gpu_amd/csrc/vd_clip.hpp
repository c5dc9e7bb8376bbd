// Internal interface of the clipping kernels (vd_clip.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vd {

struct ClipArgs {
  const float4* sites;       // n_sites (x, y, z, _), input order
  const int32_t* site_knn;   // n_sites x K, nearest first, self excluded
  int K;
  const float4* verts;       // (x, y, z, _)
  const int4* tets;          // 4 vertex ids, positively oriented
  const float4* tet_xyz;     // site-major form: per tet its corners, 3 float4 (12 floats)
  const float4* tet_prep;    // tet-major and site-major forms (tet_prep()): per tet its 4 face
                             // planes and the 4 vertices they make, 7 float4; a pair's start
                             // loads them
  const int32_t* tet_knn;    // n_tets x TK candidate sites
  int TK;
  long long n_pairs;         // n_tets * TK
  int n_sites;
  const unsigned* pair_start;  // site-major form: n_sites + 1
  const unsigned* pair_ids;  // site-major form: n_pairs pair indices grouped by site
  const unsigned* pair_count;  // site-major and cell forms, optional: only the first pair_count[s]
                               // pairs of site s's range are clipped (cull_sites' survivors)
  const int* cell_status;      // cell form, optional: cull_sites keeps every pair of a site
                               // whose cell was not built (its pairs take the cell's status),
  const int* cell_nv;          // or whose cell's plane / triangle counts could overflow the
  const int* cell_nt;          // caps while kernel C clips it (with cell_status)
  int pcap, tcap;            // polytope caps (planes, vertices) of this launch
  int tcap_full;             // configured vertex cap (site-major form: pairs that
                             // overflow a smaller launch cap go to retry_list)
  int pcap_full;             // configured plane cap (tet-major form over a retry list:
                             // likewise for a smaller launch plane cap)
  unsigned* retry_list;      // optional, n_pairs entries
  unsigned* retry_count;     // with retry_list
  int cull_depth;            // cull_sites and the cell form: bisectors tried by the emptiness
                             // test (0 = off)
  int filter;                // arithmetic filter of the conflict test (0 off, 1 reference, 2 FP64 re-check)
  unsigned* site_counter;    // site-major form: next site to take, zero before the launch
  const int* site_order;     // site-major form and cull_sites, optional: the order in which sites are
                             // taken (a permutation; neighbouring sites next to each other share
                             // tets, which then stay in cache)
  int site_depth;            // site-major form: bisectors per site table (0 = K); a pair
                             // that needs more goes to retry_list
  // site-major form, optional: polytopes whose clipping ended are handed to
  // the volume pass as records: rec_meta (pair, seed, tet, nt | status << 8)
  // and rec_stride(tcap) triangles each; rec_count is zero before the launch.
  // Past rec_cap records a lane computes its volume in place.
  uint4* rec_meta;
  uint32_t* rec_tris;
  unsigned* rec_count;
  unsigned rec_cap;
  const unsigned* pair_list; // tet-major form: optional list of pairs to process
  const unsigned* pair_list_count;
  unsigned* srnr_list;       // optional: pairs ending security_radius_not_reached are listed
  unsigned* srnr_count;      //   here instead of accumulated
  const int32_t* knn2;       // tet-major form, optional: deeper neighbour lists, K2 per row;
  int K2;                    //   the seed s uses row knn2_row[s] instead of site_knn
  const int* knn2_row;
  float* acc_f32;            // n_sites x 4 (vol, mx, my, mz), FP32 accumulation
  unsigned long long* acc_fix;  // n_sites x 4, 2^-20 fixed point of FP64 pair sums
  uint8_t* pair_status;      // optional, n_pairs
  double* pair_vals;         // optional, n_pairs x 4
  unsigned long long* status_hist;  // optional, 7 counters
};

// Cells of kernel B, kept for kernel C: per site its status, plane and
// triangle counts, squared radius of its farthest vertex, planes [pcap] and
// triangles [tcap] (three plane ids in the low 24 bits).
struct CellStore {
  int* status;
  int* nv;
  int* nt;
  float* r2;
  float4* planes;     // n_sites x pcap
  uint32_t* tris;     // n_sites x tcap
  unsigned* redo;     // optional, n_sites + 1: kernel B's list of cells built again at the configured caps
};

size_t clip_plain_lds(int pcap, int tcap);
size_t clip_plain_cached_lds(int pcap, int tcap);
// clip_launch form of the tet-major kernel with cached vertex distances
// (tcap <= 32), used for the site-major form's overflow pairs.
constexpr int kFormTetCached = 3;
size_t cells_lds(int pcap, int tcap);
size_t cells_clip_lds(int pcap, int tcap);
// Kernel B over all sites, then kernel C over the pairs grouped by site
// (a.pair_start / a.pair_ids); a.cull_depth > 0 enables the ball test and,
// with `cull_count` (n_sites), the emptiness pass (cull_sites over pair_ids)
// between the two kernels.
// Kernel B builds every cell at caps fast_pcap / fast_tcap first (0: 40 and
// 64, at most a.pcap / a.tcap) and, with cs.redo, the cells that failed
// there again at a.pcap / a.tcap.
hipError_t cells_launch(const ClipArgs& a, const CellStore& cs, int moments, int acc_mode, hipStream_t st,
                        unsigned* pair_ids = nullptr, unsigned* cull_count = nullptr, int fast_pcap = 0,
                        int fast_tcap = 0);
// LDS of the site-major kernel with `depth` bisectors per site table.
size_t clip_site_lds(int depth, int tcap);
// Triangle words per record of the site-major form's volume pass.
__host__ __device__ inline int rec_stride(int tcap) { return (tcap + 3) & ~3; }

// form: 0 tet-major, 1 site-major; moments: 0 reference, 1 robust;
// acc_mode: 0 FP32 atomics, 1 FP64 pair sums + fixed-point site sums.
// n_blocks: workgroups of the site-major form (0 = as many as stay resident),
// and of the tet-major form when it runs over a.pair_list.  The site-major
// form with records (a.rec_meta) also launches their volume pass, unless
// volume_pass is false (the caller then launches site_volume_pass).
hipError_t clip_launch(const ClipArgs& a, int form, int moments, int acc_mode, int n_blocks, hipStream_t st,
                       bool volume_pass = true);
// The volume pass over the records of a site-major launch with arguments a.
hipError_t site_volume_pass(const ClipArgs& a, int moments, int acc_mode, hipStream_t st);

// Groups the pairs of an n_tets x TK candidate table (at most 2^32 - 1 pairs)
// by site: site s owns pair_ids[start[s] .. start[s+1]).  count must be zero; count/cursor hold
// n_sites ints, start n_sites + 1, scan_tmp scan_tmp_ints(n_sites).
hipError_t pairs_by_site(const int32_t* tet_knn, long long n_pairs, int n_sites, int* count, unsigned* start,
                         int* cursor, unsigned* pair_ids, unsigned* scan_tmp, hipStream_t st);

// Largest emptiness-test depth at which a skipped pair is what the reference
// reports at caps (pcap, tcap): before the bisector that empties a tet's
// polytope it has made at most depth - 1 cuts, so it has at most depth + 3
// planes and 2 depth + 2 triangles (a closed polytope of this form has twice
// its planes less 4 triangles), and neither cap can be reached.
inline int exact_cull_depth(int depth, int pcap, int tcap) {
  const int d = depth < pcap - 4 ? depth : pcap - 4;
  return d < (tcap - 3) / 2 ? d : (tcap - 3) / 2;
}

// The site-major form's emptiness test as a pass of its own over the pairs
// grouped by site (a.pair_start, pair_ids), with a.cull_depth bisectors (at
// most 256; at most exact_cull_depth of the configured caps): the survivors of site s are moved, in order, to the front of its
// range and their number written to pair_count[s]; a culled pair counts as
// no_intersection (a.status_hist, a.pair_status, a.pair_vals).
hipError_t cull_sites(const ClipArgs& a, unsigned* pair_ids, unsigned* pair_count, hipStream_t st);
// Per tet of the corner records tet_xyz[0..3n): its face planes (ids 0..3) and
// the FP32 vertices of its four start triangles, 7 float4 per tet
// (kTetPrepWords), so that a pair's start loads them instead of computing
// them; the values are the ones the start computes, bit for bit.
constexpr int kTetPrepWords = 7;
hipError_t tet_prep(const float4* tet_xyz, long long n, float4* out, hipStream_t st);

// Rows of the deeper neighbour lists: every site seeding a listed pair gets a
// row (row_site[r] = s, row_of[s] = r).  row_of must be all -1 and *n_rows 0.
hipError_t srnr_rows(const unsigned* list, const unsigned* n_list, long long cap, const int32_t* tet_knn, int n_sites,
                     int* row_of, int* row_site, unsigned* n_rows, hipStream_t st);

hipError_t lloyd_update(float4* sites, int n, int acc_mode, const float* acc_f32, const unsigned long long* acc_fix,
                        hipStream_t st);
hipError_t moments_out(int n, int acc_mode, const float* acc_f32, const unsigned long long* acc_fix, double* out,
                       hipStream_t st);

}  // namespace vd
