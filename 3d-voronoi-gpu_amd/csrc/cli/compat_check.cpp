// Example and check of the reference-signature kNN shims (knearests.h,
// knncuda.h): compat_check <points.f32> <n> <out_dir>
// Reads n xyz float triples, then writes to out_dir:
//   kn_points.f32  kn_get_points()       (grid order)
//   kn_perm.u32    kn_get_permutation()
//   kn_knn.u32     kn_get_knearests()    (n x _K_, grid positions)
//   dev_knn.i32    knn_cuda_global_dev(points, points, _K_ + 1) rows (k x n, self included)
// and prints kn_print_stats() on stderr.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "knearests.h"
#include "knncuda.h"

static bool dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(p, 1, bytes, f) == bytes;
  fclose(f);
  return ok;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s <points.f32> <n> <out_dir>\n", argv[0]);
    return 1;
  }
  const int n = atoi(argv[2]);
  const std::string dir = argv[3];
  std::vector<float> pts(3 * (size_t)n);
  FILE* f = fopen(argv[1], "rb");
  if (!f || fread(pts.data(), sizeof(float), pts.size(), f) != pts.size()) return 1;
  fclose(f);

  kn_problem* kn = kn_prepare(reinterpret_cast<float3*>(pts.data()), n);
  kn_solve(kn);
  kn_print_stats(kn);
  float3* sp = kn_get_points(kn);
  unsigned int* perm = kn_get_permutation(kn);
  unsigned int* knn = kn_get_knearests(kn);
  bool ok = dump(dir + "/kn_points.f32", sp, sizeof(float3) * n) && dump(dir + "/kn_perm.u32", perm, 4 * (size_t)n) &&
            dump(dir + "/kn_knn.u32", knn, 4 * (size_t)n * _K_);
  free(sp);
  free(perm);
  free(knn);
  kn_free(&kn);

  // knn_cuda_global_dev on pitched SoA rows, as voronoi.cu:996-1010 calls it
  const size_t pitch = (size_t)((n + 63) / 64) * 64;
  const int k = _K_ + 1;
  std::vector<float> soa(3 * pitch, 0.f);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) soa[a * pitch + i] = pts[3 * i + a];
  float* d_pts = nullptr;
  int* d_idx = nullptr;
  if (hipMalloc(&d_pts, sizeof(float) * soa.size()) != hipSuccess ||
      hipMalloc(&d_idx, sizeof(int) * pitch * k) != hipSuccess)
    return 1;
  if (hipMemcpy(d_pts, soa.data(), sizeof(float) * soa.size(), hipMemcpyHostToDevice) != hipSuccess) return 1;
  ok = ok && knn_cuda_global_dev(d_pts, n, pitch, d_pts, n, pitch, 3, k, d_idx, pitch);
  std::vector<int> idx(pitch * k);
  if (hipMemcpy(idx.data(), d_idx, sizeof(int) * idx.size(), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  std::vector<int> rows((size_t)k * n);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < n; ++i) rows[(size_t)j * n + i] = idx[j * pitch + i];
  ok = ok && dump(dir + "/dev_knn.i32", rows.data(), sizeof(int) * rows.size());
  (void)hipFree(d_pts);
  (void)hipFree(d_idx);
  printf(ok ? "compat_check ok\n" : "compat_check failed\n");
  return ok ? 0 : 1;
}
