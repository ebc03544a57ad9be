// Drives libhodlr3d_asan.so (the library's host code under AddressSanitizer
// and UndefinedBehaviorSanitizer) through its C ABI.  Without a GPU: the
// argument checks, error codes and clean-up paths.  With one: a small
// workload over most entry points (create, assemble, products in every
// precision and with several right-hand sides, transposes, entries, the mixed
// copy, solves, save / load, graphs' absence aside, tuning, timers, an
// injected allocation failure, moves, destroy).  Exit 0 = no error detected
// and every status as expected.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "hodlr3d.h"

static int fails = 0;
#define EXPECT(cond)                                                       \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::printf("FAIL line %d: %s (%s)\n", __LINE__, #cond, h3d_last_error()); \
      ++fails;                                                             \
    }                                                                      \
  } while (0)

static double rel(const std::vector<double>& a, const std::vector<double>& b) {
  double d = 0, n = 0;
  for (size_t i = 0; i < a.size(); ++i) d += (a[i] - b[i]) * (a[i] - b[i]), n += b[i] * b[i];
  return std::sqrt(d / n);
}

int main(int argc, char** argv) {
  const std::string tmp = argc > 1 ? argv[1] : "/tmp";
  std::mt19937_64 g(5);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  const int64_t N = 6000;
  std::vector<double> xyz(3 * N);
  for (auto& v : xyz) v = u(g);
  // ---- argument checks (no GPU needed)
  h3d_handle* h = nullptr;
  EXPECT(h3d_create(nullptr, N, 4, 1.0, 8, nullptr, nullptr, &h) == H3D_ERR_INVALID && !h);
  EXPECT(h3d_create(xyz.data(), 0, 4, 1.0, 8, nullptr, nullptr, &h) == H3D_ERR_INVALID);
  EXPECT(h3d_create(xyz.data(), N, 4, -1.0, 8, nullptr, nullptr, &h) == H3D_ERR_INVALID);
  h3d_kernel bad{50, 1.0, 0.0, nullptr};
  EXPECT(h3d_create(xyz.data(), N, 4, 1.0, 8, &bad, nullptr, &h) == H3D_ERR_INVALID);
  h3d_options opt{{0, 0, 0}, 9, 0, 0};
  EXPECT(h3d_create(xyz.data(), N, 4, 1.0, 8, nullptr, &opt, &h) == H3D_ERR_INVALID);
  std::vector<double> nan_xyz(xyz);
  nan_xyz[7] = NAN;
  EXPECT(h3d_create(nan_xyz.data(), N, 4, 1.0, 8, nullptr, nullptr, &h) == H3D_ERR_NONFINITE);
  EXPECT(h3d_nlevels(N, 4) == 3 && h3d_nlevels(0, 4) == -1);
  EXPECT(h3d_load((tmp + "/does-not-exist.h3d").c_str(), nullptr, &h) == H3D_ERR_IO);
  {
    FILE* f = std::fopen((tmp + "/foreign.h3d").c_str(), "wb");
    const char junk[64] = "not a saved operator";
    std::fwrite(junk, 1, sizeof(junk), f);
    std::fclose(f);
    EXPECT(h3d_load((tmp + "/foreign.h3d").c_str(), nullptr, &h) == H3D_ERR_IO);
  }
  EXPECT(h3d_assemble(nullptr) == H3D_ERR_INVALID);
  EXPECT(h3d_matvec(nullptr, nullptr, nullptr, 1, 0, 0) == H3D_ERR_INVALID);
  EXPECT(h3d_get_matrix(nullptr, nullptr, 1, nullptr, 1, nullptr, 0) == H3D_ERR_INVALID);
  for (int s = 0; s <= 9; ++s) EXPECT(h3d_status_string(s) != nullptr);
  EXPECT(h3d_plugin_name() != nullptr);
  h3d_destroy(nullptr);

  // ---- a workload, when a GPU is there
  const int rc = h3d_create(xyz.data(), N, 4, 1.0, 8, nullptr, nullptr, &h);
  if (rc == H3D_ERR_HIP) {
    std::printf("no GPU: argument checks only (%d failures)\n", fails);
    return fails ? 1 : 0;
  }
  EXPECT(rc == H3D_OK);
  if (rc != H3D_OK) return 1;
  EXPECT(h3d_matvec(h, xyz.data(), xyz.data(), 1, 0, 0) == H3D_ERR_STATE);  // before assemble
  EXPECT(h3d_assemble(h) == H3D_OK);
  EXPECT(h3d_assemble(h) == H3D_ERR_STATE);
  std::vector<double> b(3 * N), y(3 * N), y64(N), yt(N), y2(N);
  std::normal_distribution<double> nd;
  for (auto& v : b) v = nd(g);
  EXPECT(h3d_matvec(h, b.data(), y64.data(), 1, H3D_PREC_FP64, H3D_HOST) == H3D_OK);
  EXPECT(h3d_matvec_transpose(h, b.data(), yt.data(), 1, H3D_PREC_FP64, H3D_HOST) == H3D_OK);
  EXPECT(h3d_prepare_precision(h, H3D_PREC_FP32) == H3D_OK);
  EXPECT(h3d_matvec(h, b.data(), y.data(), 3, H3D_PREC_FP32 | H3D_ACC_FP64, H3D_HOST) == H3D_OK);
  EXPECT(rel(std::vector<double>(y.begin(), y.begin() + N), y64) < 1e-6);
  h3d_mixed_stats ms;
  EXPECT(h3d_compress(h, 1e-7, &ms) == H3D_OK);
  EXPECT(h3d_matvec(h, b.data(), y2.data(), 1, H3D_PREC_MIXED, H3D_HOST) == H3D_OK);
  EXPECT(rel(y2, y64) < 1e-6);
  std::vector<int64_t> rows{0, 5, 17, N - 1}, cols{5, 9, 1000};
  std::vector<double> ent(rows.size() * cols.size());
  EXPECT(h3d_get_matrix(h, rows.data(), 4, cols.data(), 3, ent.data(), H3D_HOST) == H3D_OK);
  EXPECT(ent[1] == 0.0);  // K(5, 5): the diagonal
  std::vector<int64_t> oob{N};
  EXPECT(h3d_get_matrix(h, oob.data(), 1, cols.data(), 1, ent.data(), H3D_HOST) == H3D_ERR_INVALID);
  std::vector<double> d(rows.size());
  EXPECT(h3d_direct(h, b.data(), rows.data(), 4, d.data(), H3D_HOST) == H3D_OK);
  h3d_stats st;
  EXPECT(h3d_get_stats(h, &st) == H3D_OK && st.find_memory > 0 && st.max_rank > 0);
  int tun[6];
  EXPECT(h3d_get_tuning(h, H3D_PREC_MIXED, tun) == H3D_OK);
  EXPECT(h3d_set_tuning_state(h, H3D_PREC_MIXED, tun) == H3D_OK);
  double kms[15];
  int cnt = 0;
  EXPECT(h3d_kernel_timing(h, 1, nullptr, nullptr) == H3D_OK);
  EXPECT(h3d_matvec(h, b.data(), y2.data(), 1, H3D_PREC_FP64, H3D_HOST) == H3D_OK);
  EXPECT(h3d_kernel_timing(h, 0, kms, &cnt) == H3D_OK && cnt == 1);
  // solves
  EXPECT(h3d_set_operator(h, 1e-4, 1.0) == H3D_OK);
  h3d_solve_info info;
  EXPECT(h3d_solve_ir(h, b.data(), y2.data(), 30, 300, 1e-5, H3D_PREC_FP32, 1e-10, 10, 0, &info,
                      H3D_HOST) == H3D_OK);
  EXPECT(info.rel_residual <= 1e-10);
  std::vector<double> hist(101);
  double err = 0;
  int its = 0;
  EXPECT(h3d_gmres(h, b.data(), y2.data(), 100, 1e-10, &err, &its, hist.data(), H3D_HOST) ==
         H3D_OK);
  EXPECT(h3d_set_operator(h, 1.0, 0.0) == H3D_OK);
  // save / load
  const std::string path = tmp + "/op.h3d";
  EXPECT(h3d_save(h, path.c_str()) == H3D_OK);
  h3d_handle* h2 = nullptr;
  EXPECT(h3d_load(path.c_str(), nullptr, &h2) == H3D_OK);
  if (h2) {
    EXPECT(h3d_matvec(h2, b.data(), y2.data(), 1, H3D_PREC_FP64, H3D_HOST) == H3D_OK);
    EXPECT(rel(y2, y64) == 0.0);
    h3d_destroy(h2);
  }
  // no-allocation product: device pointers are required, host ones are refused by HIP
  // (not exercised here); injected allocation failures
  EXPECT(h3d_debug_fail_alloc(1) == H3D_OK);
  EXPECT(h3d_prepare_precision(h, H3D_PREC_BF16) == H3D_ERR_OOM);
  EXPECT(h3d_debug_fail_alloc(-1) == H3D_OK);
  EXPECT(h3d_matvec(h, b.data(), y2.data(), 1, H3D_PREC_FP64, H3D_HOST) == H3D_OK);
  EXPECT(rel(y2, y64) == 0.0);
  EXPECT(h3d_move_precision(h, H3D_PREC_BF16) == H3D_OK);
  EXPECT(h3d_matvec(h, b.data(), y2.data(), 1, H3D_PREC_BF16, H3D_HOST) == H3D_OK);
  EXPECT(rel(y2, y64) < 3e-2);
  EXPECT(h3d_release_precision(h, H3D_PREC_MIXED) == H3D_OK);
  h3d_destroy(h);
  std::printf("GPU workload done (%d failures)\n", fails);
  std::fflush(stdout);
  // End here, without the HIP runtime's own teardown: under ASan, chunks of
  // device memory freed during the run wait in the sanitizer's quarantine, and
  // draining it from the runtime's exit-time destructors, after the device
  // runtime has unloaded, trips ASan's own check.  Every handle is destroyed
  // above, so the library's free paths have run.
  std::_Exit(fails ? 1 : 0);
}
