// hodlr3d.hpp -- header-only C++ mirror of the reference facade over the C ABI.
//
// Same class names and argument meaning as SAFRAN-LAB/HODLR3D:
//   pts3D                          include/definitions.hpp:7-9
//   userkernel(Qchoice)            include/kernel.hpp:37-115 (catalogue + i == j rule)
//   HODLR3D(leaf1d, L, TOL_POW, Qchoice, particles), assemble(), MatVecProduct(b, out)
//                                  include/HODLR3D.hpp:12-22, src/HODLR3D.cpp:3-27
//   classGMRES::gmres(M, b, maxIterations, threshold, x, error, noOfIterations, e)
//                                  include/gmres.hpp:13-27, src/gmres.cpp:40-94
// Vectors are std::vector<double> instead of Eigen::VectorXd; errors throw
// std::runtime_error with the library's message instead of undefined behaviour.
#ifndef HODLR3D_HPP_
#define HODLR3D_HPP_

#include <stdexcept>
#include <string>
#include <vector>

#include "hodlr3d.h"

#ifndef HODLR3D_HAVE_PTS3D
struct pts3D {
  double x, y, z;
};
#endif

namespace hodlr3d_detail {
inline void check(int rc) {
  if (rc != H3D_OK) throw std::runtime_error(std::string("hodlr3d: ") + h3d_last_error());
}
}  // namespace hodlr3d_detail

// Kernel choice: entry(i,j) = scale * phi_Qchoice(r_ij), diag on i == j.
class userkernel {
 public:
  int Qchoice;
  double scale, diag;
  std::vector<double> diag_vec;
  explicit userkernel(int q = 7, double scale_ = -1.0, double diag_ = 0.0)
      : Qchoice(q), scale(scale_ < 0 ? (q == 16 ? 0.0 : 1.0) : scale_), diag(diag_) {}
};

class HODLR3D {
 public:
  h3d_handle* h = nullptr;
  int N = 0;
  int nLevels = 0;
  userkernel mykernel;

  HODLR3D(int nParticlesInLeafAlong1D, double L, int TOL_POW, int Qchoice,
          std::vector<pts3D>& particles)
      : HODLR3D(nParticlesInLeafAlong1D, L, TOL_POW, userkernel(Qchoice), particles) {}

  HODLR3D(int nParticlesInLeafAlong1D, double L, int TOL_POW, const userkernel& k,
          std::vector<pts3D>& particles)
      : N((int)particles.size()), mykernel(k) {
    h3d_kernel kk{k.Qchoice, k.scale, k.diag, k.diag_vec.empty() ? nullptr : k.diag_vec.data()};
    static_assert(sizeof(pts3D) == 3 * sizeof(double), "pts3D must be three doubles");
    hodlr3d_detail::check(h3d_create(reinterpret_cast<const double*>(particles.data()),
                                     (int64_t)particles.size(), nParticlesInLeafAlong1D, L,
                                     TOL_POW, &kk, nullptr, &h));
    h3d_stats s;
    hodlr3d_detail::check(h3d_get_stats(h, &s));
    nLevels = s.nlevels;
  }
  HODLR3D(const HODLR3D&) = delete;
  HODLR3D& operator=(const HODLR3D&) = delete;
  ~HODLR3D() { h3d_destroy(h); }

  void assemble() { hodlr3d_detail::check(h3d_assemble(h)); }

  // out = A b (original ordering), fp64 factors unless prec says otherwise
  void MatVecProduct(const std::vector<double>& b, std::vector<double>& out,
                     int prec = H3D_PREC_FP64) {
    if ((int)b.size() != N) throw std::runtime_error("hodlr3d: b has the wrong length");
    out.resize(N);
    hodlr3d_detail::check(h3d_matvec(h, b.data(), out.data(), 1, prec, H3D_HOST));
  }

  double findMemory() {
    h3d_stats s;
    hodlr3d_detail::check(h3d_get_stats(h, &s));
    return s.find_memory;
  }
  int getMaxRank() {
    h3d_stats s;
    hodlr3d_detail::check(h3d_get_stats(h, &s));
    return s.max_rank;
  }
};

class classGMRES {
 public:
  void gmres(HODLR3D* M, const std::vector<double>& b, int maxIterations, double threshold,
             std::vector<double>& x, double& error, int& noOfIterations, std::vector<double>& e) {
    x.resize(b.size());
    std::vector<double> hist(maxIterations + 1);
    hodlr3d_detail::check(h3d_gmres(M->h, b.data(), x.data(), maxIterations, threshold, &error,
                                     &noOfIterations, hist.data(), H3D_HOST));
    e.assign(hist.begin(), hist.begin() + noOfIterations + 1);
  }
};

#endif  // HODLR3D_HPP_
