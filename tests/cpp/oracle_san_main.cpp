// Runs the CPU oracle (oracle/h3d_oracle.cpp) through its C entry points under
// AddressSanitizer and UndefinedBehaviorSanitizer: tree, assembly, both
// product forms, GMRES, findMemory, per-box assembly, a block's ACA, and the
// degenerate inputs the reference leaves undefined (empty boxes, coincident
// points, a zero kernel).  Exit status 0 = no error detected.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
void* orc_create(const double* xyz, int64_t N, int leaf1d, double L, const double* center,
                 int nlevels, int tol_pow, int qchoice, double scale, double diag,
                 const double* diag_vec, int keep_uv);
void orc_destroy(void* h);
int orc_get_nlevels(void* h);
double orc_assemble(void* h);
double orc_assemble_boxes(void* h, int level, int64_t n, const int32_t* boxes);
void orc_matvec(void* h, const double* q, double* y, int mode, const uint8_t* leafmask);
double orc_find_memory(void* h);
int orc_max_rank(void* h);
int orc_gmres(void* h, const double* b, int maxit, double thr, double* x, double* err, int* its,
              double* hist);
int orc_aca_positions(void* h, int rs, int n, int cs, int m, int32_t* I, int32_t* J, double* U,
                      double* V);
}

static int run(const std::vector<double>& xyz, int leaf, int tol, int q, double scale,
               bool keep) {
  const int64_t N = (int64_t)xyz.size() / 3;
  void* h = orc_create(xyz.data(), N, leaf, 1.0, nullptr, -1, tol, q, scale, 0.0, nullptr,
                       keep ? 1 : 0);
  if (!h) return 1;
  orc_assemble(h);
  std::vector<double> b(N), y(N), x(N), hist(41);
  std::mt19937_64 g(7);
  std::normal_distribution<double> nd;
  for (auto& v : b) v = nd(g);
  orc_matvec(h, b.data(), y.data(), 0, nullptr);
  orc_matvec(h, b.data(), y.data(), keep ? 1 : 0, nullptr);
  double err = 0;
  int its = 0;
  orc_gmres(h, b.data(), 40, 1e-8, x.data(), &err, &its, hist.data());
  const double fm = orc_find_memory(h);
  const int mr = orc_max_rank(h);
  int nmin = N < 64 ? (int)N : 64;
  std::vector<int32_t> I(nmin), J(nmin);
  std::vector<double> U(nmin * nmin), V(nmin * nmin);
  orc_aca_positions(h, 0, nmin / 2, nmin / 2, nmin - nmin / 2, I.data(), J.data(), U.data(),
                    V.data());
  std::printf("N %lld levels %d findMemory %.0f maxrank %d gmres its %d\n", (long long)N,
              orc_get_nlevels(h), fm, mr, its);
  orc_destroy(h);
  return 0;
}

int main() {
  std::mt19937_64 g(3);
  std::uniform_real_distribution<double> u(-1.0, 1.0), c(0.05, 0.95);
  std::vector<double> a(3 * 3000), oct(3 * 2000);
  for (auto& v : a) v = u(g);
  for (auto& v : oct) v = c(g);  // one octant: many empty boxes
  std::vector<double> dup(a);
  dup.insert(dup.end(), a.begin(), a.begin() + 3 * 300);  // 300 coincident points
  int rc = 0;
  rc |= run(a, 4, 8, 7, 1.0, true);
  rc |= run(oct, 4, 7, 7, 1.0, false);
  rc |= run(dup, 4, 8, 3, 1.0, true);
  rc |= run(dup, 4, 8, 100, 1.0, false);
  rc |= run(a, 4, 8, 7, 0.0, false);  // zero kernel: rank 0 everywhere
  return rc;
}
