// A driver making the calls of the reference's examples/testHODLR3D.cpp:103-154
// through include/hodlr3d.hpp, with a small vector type of the Eigen::VectorXd
// shape (Zero, +, -, norm, operator()) standing in for Eigen in this test.
//
//   ref_driver cubeRootN leaf1d L TOL_POW Qchoice seed out.bin
//
// Points: the reference's lattice generator (examples/testHODLR3D.cpp:15-31).
// b: N/500 ones at distinct indices drawn with a seeded generator (the
// reference seeds with the time).  Writes N, nLevels, the relative error of Hb
// against the sum of kernel columns, findMemory, getMaxRank, assTime,
// matVecTime, the indices and Hb to out.bin for the test to check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

namespace testvec {
struct VectorXd {
  std::vector<double> v;
  VectorXd() = default;
  explicit VectorXd(long n) : v((size_t)n, 0.0) {}
  static VectorXd Zero(long n) { return VectorXd(n); }
  long size() const { return (long)v.size(); }
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  void resize(long n) { v.resize((size_t)n); }
  double& operator()(long i) { return v[(size_t)i]; }
  double operator()(long i) const { return v[(size_t)i]; }
  VectorXd operator+(const VectorXd& o) const {
    VectorXd r(size());
    for (long i = 0; i < size(); ++i) r.v[i] = v[i] + o.v[i];
    return r;
  }
  VectorXd operator-(const VectorXd& o) const {
    VectorXd r(size());
    for (long i = 0; i < size(); ++i) r.v[i] = v[i] - o.v[i];
    return r;
  }
  double norm() const {
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
  }
};
}  // namespace testvec

#define HODLR3D_VECTOR testvec::VectorXd
#include "hodlr3d.hpp"

// examples/testHODLR3D.cpp:15-31: x_k = -L + 2L(k+1)/(n+1), z fastest, then x, then y
static void set_Uniform_Nodes(int n, double L, std::vector<pts3D>& particles) {
  std::vector<double> t(n);
  for (int k = 0; k < n; ++k) t[k] = -L + 2.0 * L * (k + 1.0) / (n + 1.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) particles.push_back(pts3D{t[k], t[j], t[i]});
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s cubeRootN leaf1d L TOL_POW Qchoice seed out.bin\n", argv[0]);
    return 2;
  }
  const int cubeRootN = std::atoi(argv[1]), nParticlesInLeafAlong1D = std::atoi(argv[2]);
  const double L = std::atof(argv[3]);
  const int TOL_POW = std::atoi(argv[4]), Qchoice = std::atoi(argv[5]);
  const unsigned seed = (unsigned)std::atoi(argv[6]);
  try {
    std::vector<pts3D> particles;
    set_Uniform_Nodes(cubeRootN, L, particles);
    HODLR3D* H = new HODLR3D(nParticlesInLeafAlong1D, L, TOL_POW, Qchoice, particles);
    int N = H->N;
    H->assemble();

    testvec::VectorXd b = testvec::VectorXd::Zero(N);
    int n = N / 500;
    std::mt19937 gen(seed);
    std::set<int> s;
    while ((int)s.size() < n) s.insert((int)(gen() % (unsigned)N));
    std::set<int>::iterator it;
    for (it = s.begin(); it != s.end(); it++) b(*it) = 1.0;

    testvec::VectorXd Hb;
    H->MatVecProduct(b, Hb);
    double timeAss = H->A->assTime;
    double timeMatVecProduct = H->A->matVecTime;

    testvec::VectorXd true_Ab = testvec::VectorXd::Zero(N);
    for (it = s.begin(); it != s.end(); it++) true_Ab = true_Ab + H->mykernel->getCol(*it);
    double err = (true_Ab - Hb).norm() / true_Ab.norm();

    double sum;
    H->A->findMemory(sum);
    int maxRank = H->A->getMaxRank();
    std::printf("N %d nLevels %d err %.3e memory %.0f maxrank %d assTime %.3e matVecTime %.3e\n",
                N, H->nLevels, err, sum, maxRank, timeAss, timeMatVecProduct);

    FILE* f = std::fopen(argv[7], "wb");
    if (!f) return 3;
    double head[8] = {(double)N, (double)H->nLevels, err, sum, (double)maxRank, timeAss,
                      timeMatVecProduct, (double)s.size()};
    std::fwrite(head, sizeof(double), 8, f);
    for (it = s.begin(); it != s.end(); it++) {
      double v = *it;
      std::fwrite(&v, sizeof(double), 1, f);
    }
    std::fwrite(Hb.data(), sizeof(double), (size_t)N, f);
    std::fclose(f);
    delete H;
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  return 0;
}
