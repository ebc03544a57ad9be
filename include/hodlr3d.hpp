// hodlr3d.hpp -- header-only C++ mirror of the reference's classes over the C ABI.
//
// Same class names, members and argument meaning as SAFRAN-LAB/HODLR3D, so that
// a driver written against the reference (examples/testHODLR3D.cpp:103-154)
// compiles against this header:
//   pts3D                              include/definitions.hpp:7-9
//   kernel::getMatrixEntry / getRow / getCol / getMatrix
//                                      include/kernel.hpp:12-33, src/kernel.cpp:7-62
//   userkernel(particles, Qchoice)     include/kernel.hpp:37-115
//   HODLR3D(leaf1d, L, TOL_POW, Qchoice, particles), assemble(), MatVecProduct(b, out),
//   members N, mykernel, A             include/HODLR3D.hpp:12-22, src/HODLR3D.cpp:3-27
//   A->assTime, A->matVecTime, A->findMemory(sum), A->getMaxRank()
//                                      include/HODLR3DTree.hpp:46, src/HODLR3DTree.cpp:1757-1818
//   classGMRES::gmres(M, b, maxIterations, threshold, x, error, noOfIterations, e)
//                                      include/gmres.hpp:13-27, src/gmres.cpp:40-94
// Vectors: any type with size(), data() and resize(n) -- std::vector<double> by
// default; define HODLR3D_VECTOR (e.g. Eigen::VectorXd, after including Eigen)
// before this header for the type getRow/getCol return, and HODLR3D_MATRIX
// (a type with resize(rows, cols) and operator()(r, c), e.g. Eigen::MatrixXd)
// for getMatrix.  Errors throw std::runtime_error with the library's message
// instead of the reference's undefined behaviour.
//
// Timers: A->assTime / A->matVecTime follow the reference (seconds, accumulated
// over every product, never reset): assTime the regeneration of kernel entries
// inside products (the on-the-fly near field), matVecTime the low-rank BLAS-2
// applies (the far field).  To keep them separable the products of this class
// run every kernel in line (h3d_kernel_timing); setReferenceTimers(false)
// restores the overlapped near field and stops the timers.
#ifndef HODLR3D_HPP_
#define HODLR3D_HPP_

#include <cstdint>
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
// Column-major dense block with the Eigen::MatrixXd members getMatrix uses.
struct Matrix {
  int64_t nr = 0, nc = 0;
  std::vector<double> v;
  void resize(int64_t r, int64_t c) {
    nr = r;
    nc = c;
    v.assign((size_t)(r * c), 0.0);
  }
  int64_t rows() const { return nr; }
  int64_t cols() const { return nc; }
  double& operator()(int64_t r, int64_t c) { return v[(size_t)(r + c * nr)]; }
  double operator()(int64_t r, int64_t c) const { return v[(size_t)(r + c * nr)]; }
};
}  // namespace hodlr3d_detail

#ifndef HODLR3D_VECTOR
#define HODLR3D_VECTOR std::vector<double>
#endif
#ifndef HODLR3D_MATRIX
#define HODLR3D_MATRIX hodlr3d_detail::Matrix
#endif

// include/kernel.hpp:12-33: entries K(i, j) of the particles, here evaluated by
// the library on the GPU (h3d_get_matrix) once the HODLR3D owning the kernel
// has built its handle.
class kernel {
 public:
  double a = 0.0;
  std::vector<pts3D> particles;
  explicit kernel(const std::vector<pts3D>& p) : particles(p) {}
  virtual ~kernel() = default;

  virtual double getMatrixEntry(const unsigned i, const unsigned j) {
    const int64_t r = i, c = j;
    double v = 0.0;
    hodlr3d_detail::check(h3d_get_matrix(bound(), &r, 1, &c, 1, &v, H3D_HOST));
    return v;
  }
  HODLR3D_VECTOR getRow(const int j, std::vector<int> col_indices) {
    std::vector<int64_t> r(1, j), c(col_indices.begin(), col_indices.end());
    return gather(r, c);
  }
  HODLR3D_VECTOR getCol(const int k, std::vector<int> row_indices) {
    std::vector<int64_t> r(row_indices.begin(), row_indices.end()), c(1, k);
    return gather(r, c);
  }
  HODLR3D_VECTOR getCol(const int k) {
    std::vector<int64_t> r(particles.size()), c(1, k);
    for (size_t i = 0; i < r.size(); ++i) r[i] = (int64_t)i;
    return gather(r, c);
  }
  HODLR3D_MATRIX getMatrix(std::vector<int> row_indices, std::vector<int> col_indices) {
    std::vector<int64_t> r(row_indices.begin(), row_indices.end());
    std::vector<int64_t> c(col_indices.begin(), col_indices.end());
    std::vector<double> buf(r.size() * c.size());
    if (!buf.empty())
      hodlr3d_detail::check(h3d_get_matrix(bound(), r.data(), (int64_t)r.size(), c.data(),
                                           (int64_t)c.size(), buf.data(), H3D_HOST));
    HODLR3D_MATRIX m;
    m.resize((int64_t)r.size(), (int64_t)c.size());
    for (size_t q = 0; q < c.size(); ++q)
      for (size_t p = 0; p < r.size(); ++p) m((int64_t)p, (int64_t)q) = buf[p + q * r.size()];
    return m;
  }

  h3d_handle* h = nullptr;  // set by the owning HODLR3D

 private:
  h3d_handle* bound() const {
    if (!h) throw std::runtime_error("hodlr3d: kernel entries need the HODLR3D that owns it");
    return h;
  }
  HODLR3D_VECTOR gather(const std::vector<int64_t>& r, const std::vector<int64_t>& c) {
    HODLR3D_VECTOR out;
    out.resize(r.size() * c.size());
    if (r.size() * c.size())
      hodlr3d_detail::check(h3d_get_matrix(bound(), r.data(), (int64_t)r.size(), c.data(),
                                           (int64_t)c.size(), out.data(), H3D_HOST));
    return out;
  }
};

// include/kernel.hpp:37-115: the Qchoice catalogue (0..16) or the plugin built
// into the library (H3D_QCHOICE_PLUGIN), entry(i,j) = scale * phi(x_i, x_j)
// for i != j and diag (or diag_vec[i]) for i == j.  scale/diag default to the
// reference's dispatch (1 and 0; Qchoice 16: h3 = 1/N and Kii).
class userkernel : public kernel {
 public:
  int Qchoice;
  double Kii = 0.0, h3 = 1.0;  // filled from the library for Qchoice 16
  double scale, diag;
  std::vector<double> diag_vec;
  userkernel(const std::vector<pts3D>& particles, int Qchoice_)
      : kernel(particles), Qchoice(Qchoice_), scale(Qchoice_ == 16 ? 0.0 : 1.0), diag(0.0) {}
};

class HODLR3D;

// The members of the reference's HODLR3DTree that drivers read (H->A->...).
class HODLR3DTree {
 public:
  double assTime = 0.0, matVecTime = 0.0;
  int nLevels = 0;
  int N = 0;
  void findMemory(double& sum) { sum = stats().find_memory; }
  int getMaxRank() { return stats().max_rank; }

 private:
  friend class HODLR3D;
  h3d_handle* h = nullptr;
  h3d_stats stats() {
    h3d_stats s;
    hodlr3d_detail::check(h3d_get_stats(h, &s));
    return s;
  }
  void refresh_timers() {
    const h3d_stats s = stats();
    assTime = s.ref_ass_time;
    matVecTime = s.ref_matvec_time;
  }
};

class HODLR3D {
 public:
  int N = 0;
  int nLevels = 0;
  userkernel* mykernel = nullptr;
  HODLR3DTree* A = nullptr;

  // src/HODLR3D.cpp:3-15
  HODLR3D(int nParticlesInLeafAlong1D, double L, int TOL_POW, int Qchoice,
          std::vector<pts3D>& particles)
      : HODLR3D(nParticlesInLeafAlong1D, L, TOL_POW, userkernel(particles, Qchoice)) {}

  // The same with a kernel's scale, diagonal or per-point diagonal set.
  HODLR3D(int nParticlesInLeafAlong1D, double L, int TOL_POW, const userkernel& k) {
    static_assert(sizeof(pts3D) == 3 * sizeof(double), "pts3D must be three doubles");
    mykernel = new userkernel(k);
    A = new HODLR3DTree();
    N = (int)k.particles.size();
    h3d_kernel kk{k.Qchoice, k.scale, k.diag, k.diag_vec.empty() ? nullptr : k.diag_vec.data()};
    h3d_handle* hh = nullptr;
    const int rc = h3d_create(reinterpret_cast<const double*>(k.particles.data()), (int64_t)N,
                              nParticlesInLeafAlong1D, L, TOL_POW, &kk, nullptr, &hh);
    if (rc != H3D_OK) {
      delete mykernel;
      delete A;
      hodlr3d_detail::check(rc);
    }
    h = hh;
    mykernel->h = h;
    A->h = h;
    A->N = N;
    h3d_config c;
    hodlr3d_detail::check(h3d_get_config(h, &c));
    nLevels = A->nLevels = c.nlevels;
    if (k.Qchoice == 16) {
      mykernel->h3 = c.kernel_scale;
      mykernel->Kii = c.kernel_diag;
    }
    setReferenceTimers(true);
  }
  HODLR3D(const HODLR3D&) = delete;
  HODLR3D& operator=(const HODLR3D&) = delete;
  // src/HODLR3D.cpp:29-32
  ~HODLR3D() {
    h3d_destroy(h);
    delete mykernel;
    delete A;
  }

  // src/HODLR3D.cpp:17-19
  void assemble() { hodlr3d_detail::check(h3d_assemble(h)); }

  // src/HODLR3D.cpp:21-27: out = A b in the original ordering (fp64 factors
  // unless `prec` says otherwise)
  template <class Vec>
  void MatVecProduct(const Vec& b, Vec& out, int prec = H3D_PREC_FP64) {
    if ((int64_t)b.size() != (int64_t)N) throw std::runtime_error("hodlr3d: b has the wrong length");
    out.resize(N);
    hodlr3d_detail::check(h3d_matvec(h, b.data(), out.data(), 1, prec, H3D_HOST));
    A->refresh_timers();
  }

  // Reference timers on (every kernel in line, timed) or off (overlapped near field).
  void setReferenceTimers(bool on) { hodlr3d_detail::check(h3d_kernel_timing(h, on ? 1 : 0, nullptr, nullptr)); }

  h3d_handle* handle() const { return h; }

 private:
  h3d_handle* h = nullptr;
};

class classGMRES {
 public:
  // src/gmres.cpp:40-94 on the GPU
  template <class Vec>
  void gmres(HODLR3D* M, const Vec& b, int maxIterations, double threshold, Vec& x,
             double& error, int& noOfIterations, std::vector<double>& e) {
    x.resize(b.size());
    std::vector<double> hist(maxIterations + 1);
    hodlr3d_detail::check(h3d_gmres(M->handle(), b.data(), x.data(), maxIterations, threshold,
                                     &error, &noOfIterations, hist.data(), H3D_HOST));
    e.assign(hist.begin(), hist.begin() + noOfIterations + 1);
  }
};

#endif  // HODLR3D_HPP_
