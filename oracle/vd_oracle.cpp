// C entry points of the CPU oracle (test infrastructure only; see vd_oracle.hpp).
#include "vd_oracle.hpp"

#include <omp.h>

#include <array>
#include <cstdio>
#include <queue>

namespace vdo {

void knn_brute(const float* ref, int n_ref, const float* qry, int n_q, int k, bool self, int32_t* out,
               int nthreads) {
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<uint64_t> keys;
    keys.reserve(n_ref);
#pragma omp for schedule(dynamic, 64)
    for (int q = 0; q < n_q; ++q) {
      keys.clear();
      for (int r = 0; r < n_ref; ++r) {
        if (self && r == q) continue;
        keys.push_back(knn_key(dist2(ref + 3 * r, qry + 3 * q), (uint32_t)r));
      }
      const int kk = std::min<int>(k, (int)keys.size());
      std::partial_sort(keys.begin(), keys.begin() + kk, keys.end());
      for (int j = 0; j < k; ++j) out[(int64_t)q * k + j] = j < kk ? (int32_t)(keys[j] & 0xffffffffu) : -1;
    }
  }
}

namespace {

struct Grid {
  double lo[3], h, inv_h;
  int n[3];
  std::vector<int32_t> start;  // cell -> first slot, size cells+1
  std::vector<int32_t> ids;    // slots -> reference index, by cell then index
  int cell_coord(double x, int a) const {
    int c = (int)std::floor((x - lo[a]) * inv_h);
    return std::max(0, std::min(c, n[a] - 1));
  }
};

Grid build_grid(const float* ref, int n_ref, double per_cell) {
  Grid g;
  double hi[3];
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = hi[a] = ref[a];
  }
  for (int i = 0; i < n_ref; ++i)
    for (int a = 0; a < 3; ++a) {
      g.lo[a] = std::min(g.lo[a], (double)ref[3 * i + a]);
      hi[a] = std::max(hi[a], (double)ref[3 * i + a]);
    }
  double ext = std::max({hi[0] - g.lo[0], hi[1] - g.lo[1], hi[2] - g.lo[2], 1e-30});
  double cells = std::max(1.0, n_ref / per_cell);
  double vol = 1.0;
  for (int a = 0; a < 3; ++a) vol *= std::max(hi[a] - g.lo[a], ext * 1e-3);
  g.h = std::cbrt(vol / cells);
  for (int a = 0; a < 3; ++a) g.n[a] = std::max(1, std::min(1024, (int)std::ceil((hi[a] - g.lo[a]) / g.h + 1e-9)));
  g.inv_h = 1.0 / g.h;
  const int64_t nc = (int64_t)g.n[0] * g.n[1] * g.n[2];
  std::vector<int32_t> cell(n_ref);
  g.start.assign(nc + 1, 0);
  for (int i = 0; i < n_ref; ++i) {
    const int cx = g.cell_coord(ref[3 * i], 0), cy = g.cell_coord(ref[3 * i + 1], 1),
              cz = g.cell_coord(ref[3 * i + 2], 2);
    cell[i] = (int32_t)(cx + (int64_t)g.n[0] * (cy + (int64_t)g.n[1] * cz));
    g.start[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < nc; ++c) g.start[c + 1] += g.start[c];
  g.ids.resize(n_ref);
  std::vector<int32_t> fill(g.start.begin(), g.start.end() - 1);
  for (int i = 0; i < n_ref; ++i) g.ids[fill[cell[i]]++] = i;
  return g;
}

}  // namespace

void knn_grid(const float* ref, int n_ref, const float* qry, int n_q, int k, bool self, int32_t* out,
              int nthreads) {
  if (n_ref == 0) {
    for (int64_t i = 0; i < (int64_t)n_q * k; ++i) out[i] = -1;
    return;
  }
  const Grid g = build_grid(ref, n_ref, 2.0);
  const int want_all = std::min(k, n_ref - (self ? 1 : 0));
#pragma omp parallel num_threads(nthreads)
  {
    std::priority_queue<uint64_t> heap;
#pragma omp for schedule(dynamic, 64)
    for (int q = 0; q < n_q; ++q) {
      while (!heap.empty()) heap.pop();
      const float* p = qry + 3 * q;
      int c[3];
      for (int a = 0; a < 3; ++a) c[a] = g.cell_coord(p[a], a);
      for (int R = 0;; ++R) {
        // cells at Chebyshev distance R from c
        for (int dz = -R; dz <= R; ++dz) {
          const int z = c[2] + dz;
          if (z < 0 || z >= g.n[2]) continue;
          for (int dy = -R; dy <= R; ++dy) {
            const int y = c[1] + dy;
            if (y < 0 || y >= g.n[1]) continue;
            const bool shell = (std::abs(dz) == R || std::abs(dy) == R);
            for (int dx = -R; dx <= R; dx += (shell || R == 0) ? 1 : 2 * R) {
              const int x = c[0] + dx;
              if (x < 0 || x >= g.n[0]) continue;
              const int64_t cell = x + (int64_t)g.n[0] * (y + (int64_t)g.n[1] * z);
              for (int s = g.start[cell]; s < g.start[cell + 1]; ++s) {
                const int r = g.ids[s];
                if (self && r == q) continue;
                const uint64_t key = knn_key(dist2(ref + 3 * r, p), (uint32_t)r);
                if ((int)heap.size() < k) {
                  heap.push(key);
                } else if (key < heap.top()) {
                  heap.pop();
                  heap.push(key);
                }
              }
            }
          }
        }
        // distance from p to the nearest cell not yet visited
        double L = std::numeric_limits<double>::infinity();
        for (int a = 0; a < 3; ++a) {
          if (c[a] - R > 0) L = std::min(L, (double)p[a] - (g.lo[a] + (c[a] - R) * g.h));
          if (c[a] + R < g.n[a] - 1) L = std::min(L, (g.lo[a] + (c[a] + R + 1) * g.h) - (double)p[a]);
        }
        if (L == std::numeric_limits<double>::infinity()) break;  // whole grid visited
        if ((int)heap.size() >= want_all && want_all == k) {
          float kth;
          const uint32_t bits = (uint32_t)(heap.top() >> 32);
          std::memcpy(&kth, &bits, 4);
          const double Ls = std::max(0.0, L - 1e-9 * (std::fabs((double)p[0]) + std::fabs((double)p[1]) +
                                                     std::fabs((double)p[2]) + g.h));
          if ((double)kth < Ls * Ls * (1.0 - 1e-6)) break;
        }
      }
      const int got = (int)heap.size();
      for (int j = got; j < k; ++j) out[(int64_t)q * k + j] = -1;
      for (int j = got - 1; j >= 0; --j) {
        out[(int64_t)q * k + j] = (int32_t)(heap.top() & 0xffffffffu);
        heap.pop();
      }
    }
  }
}

}  // namespace vdo

using namespace vdo;

namespace {

template <class R>
inline void load_tet(const float* verts, const int32_t* tets, int64_t t, V4<R> tv[4]) {
  for (int j = 0; j < 4; ++j) {
    const int32_t v = tets[4 * t + j];
    tv[j] = {(R)verts[3 * v], (R)verts[3 * v + 1], (R)verts[3 * v + 2], R(1)};
  }
}

// Site reduction modes.
enum { kSiteF32 = 0, kSiteFixed = 1, kSiteF64 = 2 };

struct SiteAcc {
  int mode;
  std::vector<float> f;
  std::vector<double> d;
  std::vector<int64_t> q;
  SiteAcc(int m, int n) : mode(m) {
    if (m == kSiteF32) f.assign(4 * (size_t)n, 0.f);
    if (m == kSiteF64) d.assign(4 * (size_t)n, 0.0);
    if (m == kSiteFixed) q.assign(4 * (size_t)n, 0);
  }
  void add(int s, const PairOut& o) {
    const double v[4] = {o.vol, o.mx, o.my, o.mz};
    for (int j = 0; j < 4; ++j) {
      if (mode == kSiteF32) f[4 * s + j] += (float)v[j];
      if (mode == kSiteF64) d[4 * s + j] += v[j];
      if (mode == kSiteFixed) q[4 * s + j] += (int64_t)std::llrint(v[j] * kFixScale);
    }
  }
  void write(double* out, int n) const {
    for (int64_t i = 0; i < 4 * (int64_t)n; ++i) {
      if (mode == kSiteF32) out[i] = f[i];
      if (mode == kSiteF64) out[i] = d[i];
      if (mode == kSiteFixed) out[i] = (double)q[i] / kFixScale;
    }
  }
};

void write_events(const Events& e, int64_t* ev) {
  if (!ev) return;
  ev[0] = e.pairs; ev[1] = e.clips; ev[2] = e.conflict_tests; ev[3] = e.secure_evals;
  ev[4] = e.secure_checks; ev[5] = e.new_tris; ev[6] = e.acc_tris; ev[7] = e.acc_pairs;
  ev[8] = e.max_nv; ev[9] = e.max_nt;
}

template <class R, class A>
int clip_tets_impl(const float* sites, int n_sites, const int32_t* site_knn, int k, const float* verts,
                   const int32_t* tets, int n_tets, const int32_t* tet_knn, int tk, Caps caps,
                   int32_t* pair_status, double* pair_vals, double* site_vals, int site_mode, int64_t* events,
                   int64_t* peak_hist, int nthreads, int64_t pair_begin, int64_t pair_end, int mform) {
  const int64_t npairs = (int64_t)n_tets * tk;
  if (pair_end < 0 || pair_end > npairs) pair_end = npairs;
  if (pair_begin < 0) pair_begin = 0;
  SiteAcc acc(site_mode, n_sites);
  Events total;
  const int64_t chunk = 1 << 18;
  std::vector<PairOut> buf(chunk);
  std::vector<int32_t> seeds(chunk);
  std::vector<std::array<int64_t, 1024>> hist(nthreads);
  for (auto& h : hist) h.fill(0);
  for (int64_t c0 = pair_begin; c0 < pair_end; c0 += chunk) {
    const int64_t c1 = std::min(pair_end, c0 + chunk);
#pragma omp parallel num_threads(nthreads)
    {
      Poly<R>* poly = new Poly<R>();
      poly->caps = caps;
      Events ev;
      poly->ev = &ev;
      const int tid = omp_get_thread_num();
#pragma omp for schedule(dynamic, 256)
      for (int64_t p = c0; p < c1; ++p) {
        const int64_t t = p / tk;
        const int32_t s = tet_knn[p];
        seeds[p - c0] = s;
        V4<R> tv[4];
        load_tet(verts, tets, t, tv);
        const PairOut o = clip_pair_tet<R, A>(*poly, sites, site_knn + (int64_t)s * k, k, s, tv, mform);
        buf[p - c0] = o;
        hist[tid][std::min(poly->peak_nt, 1023)]++;
      }
#pragma omp critical
      total.add(ev);
      delete poly;
    }
    for (int64_t p = c0; p < c1; ++p) {
      const PairOut& o = buf[p - c0];
      if (pair_status) pair_status[p] = o.status;
      if (pair_vals) {
        pair_vals[4 * p] = o.vol; pair_vals[4 * p + 1] = o.mx;
        pair_vals[4 * p + 2] = o.my; pair_vals[4 * p + 3] = o.mz;
      }
      if (o.accumulated) acc.add(seeds[p - c0], o);
    }
  }
  if (site_vals) acc.write(site_vals, n_sites);
  write_events(total, events);
  if (peak_hist)
    for (int i = 0; i < 1024; ++i) {
      peak_hist[i] = 0;
      for (auto& h : hist) peak_hist[i] += h[i];
    }
  return 0;
}

// Compact storage of a site cell built by kernel B.
template <class R>
struct StoredCell {
  int32_t status;
  int nv, nt;
  std::vector<V4<R>> pl;
  std::vector<std::array<uint8_t, 3>> tr;
};

template <class R, class A>
int clip_cells_impl(const float* sites, int n_sites, const int32_t* site_knn, int k, const float* verts,
                    const int32_t* tets, int n_tets, const int32_t* tet_knn, int tk, Caps caps,
                    int32_t* pair_status, double* pair_vals, double* site_vals, int site_mode, int32_t* cell_status,
                    int nthreads, int mform) {
  std::vector<StoredCell<R>> cells(n_sites);
#pragma omp parallel num_threads(nthreads)
  {
    Poly<R>* poly = new Poly<R>();
    poly->caps = caps;
#pragma omp for schedule(dynamic, 64)
    for (int s = 0; s < n_sites; ++s) {
      StoredCell<R>& sc = cells[s];
      sc.status = site_cell<R>(*poly, sites, site_knn + (int64_t)s * k, k, s);
      sc.nv = poly->nv;
      sc.nt = poly->nt;
      sc.pl.assign(poly->pl, poly->pl + poly->nv);
      sc.tr.resize(poly->nt);
      for (int t = 0; t < poly->nt; ++t) sc.tr[t] = {poly->tr[t][0], poly->tr[t][1], poly->tr[t][2]};
      if (cell_status) cell_status[s] = sc.status;
    }
    delete poly;
  }
  const int64_t npairs = (int64_t)n_tets * tk;
  std::vector<PairOut> buf(npairs);
#pragma omp parallel num_threads(nthreads)
  {
    Poly<R>* poly = new Poly<R>();
    poly->caps = caps;
#pragma omp for schedule(dynamic, 256)
    for (int64_t p = 0; p < npairs; ++p) {
      const int32_t s = tet_knn[p];
      const StoredCell<R>& sc = cells[s];
      PairOut o{};
      if (sc.status != success && sc.status != security_radius_not_reached) {
        o.status = sc.status;  // the cell could not be built; nothing to clip
      } else {
        poly->reset_links();
        poly->seed = {(R)sites[3 * s], (R)sites[3 * s + 1], (R)sites[3 * s + 2], R(1)};
        poly->status = sc.status;
        poly->nv = sc.nv;
        poly->nt = sc.nt;
        for (int i = 0; i < sc.nv; ++i) poly->pl[i] = sc.pl[i];
        for (int t = 0; t < sc.nt; ++t) poly->set_tri(t, sc.tr[t][0], sc.tr[t][1], sc.tr[t][2]);
        V4<R> tv[4];
        load_tet(verts, tets, p / tk, tv);
        o = clip_pair_cell<R, A>(*poly, tv, mform);
      }
      buf[p] = o;
    }
    delete poly;
  }
  SiteAcc acc(site_mode, n_sites);
  for (int64_t p = 0; p < npairs; ++p) {
    const PairOut& o = buf[p];
    if (pair_status) pair_status[p] = o.status;
    if (pair_vals) {
      pair_vals[4 * p] = o.vol; pair_vals[4 * p + 1] = o.mx;
      pair_vals[4 * p + 2] = o.my; pair_vals[4 * p + 3] = o.mz;
    }
    if (o.accumulated) acc.add(tet_knn[p], o);
  }
  if (site_vals) acc.write(site_vals, n_sites);
  return 0;
}

}  // namespace

extern "C" {

int vdo_version() { return 1; }

// Exact k nearest neighbours in (d2, index) order; out is n_q*k, -1 padded.
int vdo_knn(const float* ref, int n_ref, const float* qry, int n_q, int k, int self, int use_grid, int32_t* out,
            int nthreads) {
  if (k <= 0 || n_ref < 0 || n_q < 0) return -1;
  if (use_grid)
    knn_grid(ref, n_ref, qry, n_q, k, self != 0, out, nthreads);
  else
    knn_brute(ref, n_ref, qry, n_q, k, self != 0, out, nthreads);
  return 0;
}

// Tet centroids ((((v0+v1)+v2)+v3) * 0.25f, voronoi.cu:765-774), AoS out.
void vdo_tet_centroids(const float* verts, const int32_t* tets, int n_tets, float* out) {
  for (int64_t t = 0; t < n_tets; ++t)
    for (int a = 0; a < 3; ++a) {
      float c = 0.0f;
      for (int j = 0; j < 4; ++j) c += verts[3 * tets[4 * t + j] + a];
      out[3 * t + a] = c * 0.25f;
    }
}

int vdo_tet_k(int n_site, int n_tet, int preferred) { return tet_k_heuristic(n_site, n_tet, preferred); }

// Kernel A over pairs [pair_begin, pair_end) of an n_tets x tk candidate table.
//   geom: 0 = FP32 geometry, 1 = FP64 geometry.
//   pair_acc: 0 = per-pair sums in the geometry type, 1 = per-pair sums in FP64.
//   site_mode: 0 = FP32 sums in pair order, 1 = 2^-20 fixed point, 2 = FP64 sums in pair order.
// Outputs (each optional): pair_status[n_pairs], pair_vals[4*n_pairs] (vol, mx, my, mz),
// site_vals[4*n_sites], events[10], peak_hist[1024] (histogram of per-pair peak vertex count).
//   mform: 0 = the reference's moment decomposition, 1 = the robust one.
int vdo_clip_tets(int geom, int pair_acc, const float* sites, int n_sites, const int32_t* site_knn, int k,
                  const float* verts, const int32_t* tets, int n_tets, const int32_t* tet_knn, int tk, int pcap,
                  int tcap, int32_t* pair_status, double* pair_vals, double* site_vals, int site_mode,
                  int64_t* events, int64_t* peak_hist, int nthreads, int64_t pair_begin, int64_t pair_end,
                  int mform, int filter) {
  Caps caps;
  caps.filter = filter;
  caps.pcap = pcap;
  caps.tcap = tcap;
  if (pcap < 5 || pcap > 254 || tcap < 5 || tcap > Poly<float>::TMAX) return -1;
  if (geom == 0 && pair_acc == 0)
    return clip_tets_impl<float, float>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                        pair_status, pair_vals, site_vals, site_mode, events, peak_hist, nthreads,
                                        pair_begin, pair_end, mform);
  if (geom == 0 && pair_acc == 1)
    return clip_tets_impl<float, double>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                         pair_status, pair_vals, site_vals, site_mode, events, peak_hist, nthreads,
                                         pair_begin, pair_end, mform);
  if (geom == 1)
    return clip_tets_impl<double, double>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                          pair_status, pair_vals, site_vals, site_mode, events, peak_hist, nthreads,
                                          pair_begin, pair_end, mform);
  return -1;
}

// Kernels B then C (cell from the box, then clipped by each candidate tet).
int vdo_clip_cells(int geom, int pair_acc, const float* sites, int n_sites, const int32_t* site_knn, int k,
                   const float* verts, const int32_t* tets, int n_tets, const int32_t* tet_knn, int tk, int pcap,
                   int tcap, int32_t* pair_status, double* pair_vals, double* site_vals, int site_mode,
                   int32_t* cell_status, int nthreads, int mform, int filter) {
  Caps caps;
  caps.filter = filter;
  caps.pcap = pcap;
  caps.tcap = tcap;
  if (pcap < 7 || pcap > 254 || tcap < 9 || tcap > Poly<float>::TMAX) return -1;
  if (geom == 0 && pair_acc == 0)
    return clip_cells_impl<float, float>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                         pair_status, pair_vals, site_vals, site_mode, cell_status, nthreads, mform);
  if (geom == 0 && pair_acc == 1)
    return clip_cells_impl<float, double>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                          pair_status, pair_vals, site_vals, site_mode, cell_status, nthreads, mform);
  if (geom == 1)
    return clip_cells_impl<double, double>(sites, n_sites, site_knn, k, verts, tets, n_tets, tet_knn, tk, caps,
                                           pair_status, pair_vals, site_vals, site_mode, cell_status, nthreads, mform);
  return -1;
}

// Lloyd update (voronoi.cu:728-749): site = moment / volume where volume != 0.
// fp32 != 0 divides in FP32 (the reference), else in FP64 then rounds.
void vdo_lloyd_update(const double* site_vals, int n_sites, int fp32, float* sites) {
  for (int s = 0; s < n_sites; ++s) {
    const double v = site_vals[4 * s];
    if (fp32) {
      const float vf = (float)v;
      if (vf != 0.0f)
        for (int a = 0; a < 3; ++a) sites[3 * s + a] = (float)site_vals[4 * s + 1 + a] / vf;
    } else if (v != 0.0) {
      for (int a = 0; a < 3; ++a) sites[3 * s + a] = (float)(site_vals[4 * s + 1 + a] / v);
    }
  }
}

// 4x4 determinant of columns (p1,p2,p3,e) in the reference's operation order,
// exposed for the sign/closed-form tests.
double vdo_det4(const double* p1, const double* p2, const double* p3, const double* e, int fp32) {
  if (fp32) {
    V4<float> a{(float)p1[0], (float)p1[1], (float)p1[2], (float)p1[3]};
    V4<float> b{(float)p2[0], (float)p2[1], (float)p2[2], (float)p2[3]};
    V4<float> c{(float)p3[0], (float)p3[1], (float)p3[2], (float)p3[3]};
    V4<float> d{(float)e[0], (float)e[1], (float)e[2], (float)e[3]};
    return conflict_det(vertex_minors(a, b, c), d);
  }
  V4<double> a{p1[0], p1[1], p1[2], p1[3]}, b{p2[0], p2[1], p2[2], p2[3]}, c{p3[0], p3[1], p3[2], p3[3]},
      d{e[0], e[1], e[2], e[3]};
  return conflict_det(vertex_minors(a, b, c), d);
}

}  // extern "C"
