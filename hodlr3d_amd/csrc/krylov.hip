// Krylov solvers on the GPU: the reference's unrestarted GMRES and GMRES(m)
// inside fp64 iterative refinement.
//
// Vector work (dots, axpys, norms) runs on the device with a deterministic
// two-stage reduction; the Givens rotations, the rotated Hessenberg matrix and
// its back substitution (src/gmres.cpp:5-26, :88-93) run in fp64 on the device
// too.  One double crosses to the host per Arnoldi step: the residual estimate
// |beta_{k+1}| / ||b|| that decides whether the cycle stops (src/gmres.cpp:80-86).
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"
#include "stream.hpp"

namespace h3d {

namespace {

constexpr int kDotBlocks = 512;

// partial[b] = sum over the block's grid-stride slice of a[i]*b[i]
__global__ __launch_bounds__(256) void dot_partial_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b, int64_t n,
                                                          double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += a[i] * b[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] = sum(partial); mode 1: out[0] = sqrt(sum)
__global__ void dot_final_kernel(const double* __restrict__ partial, int np, double* __restrict__ out,
                                 int mode) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = mode == 1 ? sqrt(red[0]) : red[0];
}

// y = y - (*alpha) * x
__global__ void axpy_dev_kernel(double* __restrict__ y, const double* __restrict__ x,
                                const double* __restrict__ alpha, int64_t n) {
  const double a = *alpha;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] - a * x[i];
}

// dst = src / (*den)
__global__ void scale_div_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                 const double* __restrict__ den, int64_t n) {
  const double d = *den;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] / d;
}

// dst = a - b
__global__ void sub_kernel(double* __restrict__ dst, const double* __restrict__ a,
                           const double* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = a[i] - b[i];
}

// y += x
__global__ void add_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + x[i];
}

// x += sum_c y[c] Q[:, c]  (Q column-major, ld = n)
__global__ void combine_basis_kernel(double* __restrict__ x, const double* __restrict__ Q,
                                     const double* __restrict__ y, int K, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c < K; ++c) acc += Q[(int64_t)c * n + i] * y[c];
    x[i] = x[i] + acc;
  }
}

struct Vec {
  Engine& e;
  DevBuf<double> pbuf, sbuf;
  double* partial;
  double* scal;  // device scalars
  int64_t n;
  unsigned grid;
  Vec(Engine& en, int nscal) : e(en), pbuf(kDotBlocks), sbuf(nscal) {
    n = e.N;
    partial = pbuf.p;
    scal = sbuf.p;
    grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  }
  // scal[k] = a.b (mode 0) or ||a|| (mode 1, a == b)
  void dot(const double* a, const double* b, int k, int mode = 0) {
    dot_partial_kernel<<<kDotBlocks, 256, 0, e.stream>>>(a, b, n, partial);
    dot_final_kernel<<<1, 256, 0, e.stream>>>(partial, kDotBlocks, scal + k, mode);
    H3D_LAUNCH_CHECK();
  }
  double get(int k) {
    double v;
    H3D_HIP(hipMemcpyAsync(&v, scal + k, sizeof(double), hipMemcpyDeviceToHost, e.stream));
    H3D_HIP(hipStreamSynchronize(e.stream));
    return v;
  }
  void axpy(double* y, const double* x, int k) {
    axpy_dev_kernel<<<grid, 256, 0, e.stream>>>(y, x, scal + k, n);
    H3D_LAUNCH_CHECK();
  }
  void div(double* dst, const double* src, int k) {
    scale_div_kernel<<<grid, 256, 0, e.stream>>>(dst, src, scal + k, n);
    H3D_LAUNCH_CHECK();
  }
  void sub(double* dst, const double* a, const double* b) {
    sub_kernel<<<grid, 256, 0, e.stream>>>(dst, a, b, n);
    H3D_LAUNCH_CHECK();
  }
};

// Device state of one GMRES cycle of up to m steps: the rotated Hessenberg
// matrix H ((m+1) x m, column-major), the rotations cs / sn, the rotated
// right-hand side beta (m+1), the least-squares solution y and the residual
// estimate of every step.  Allocated once per solve.
struct Cycle {
  int m;
  DevBuf<double> H, cs, sn, beta, y, err;
  explicit Cycle(int m_)
      : m(m_), H((int64_t)(m_ + 1) * m_), cs(m_), sn(m_), beta(m_ + 1), y(m_), err(m_) {}
};

// beta = (r0, 0, ..., 0) with r0 = *rnorm
__global__ void cycle_init_kernel(double* __restrict__ beta, int m, const double* __restrict__ rnorm) {
  for (int i = threadIdx.x; i <= m; i += blockDim.x) beta[i] = i == 0 ? *rnorm : 0.0;
}

// Arnoldi step k (src/gmres.cpp:14-26): column k of H from the MGS
// coefficients h[0..k] and ||w|| in scal[0..k+1], rotated by the earlier
// rotations; rotation k zeroes H[k+1, k]; beta rotated; err[k] = |beta[k+1]| /
// bnorm.  One thread, fp64, no contraction (the host form's arithmetic).
__global__ void givens_kernel(const double* __restrict__ scal, double* __restrict__ H,
                              double* __restrict__ cs, double* __restrict__ sn,
                              double* __restrict__ beta, double* __restrict__ err, int k, int m,
                              double bnorm) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  double* h = H + (int64_t)k * (m + 1);
  for (int i = 0; i <= k + 1; ++i) h[i] = scal[i];
  for (int i = 0; i < k; ++i) {
    const double tmp = cs[i] * h[i] + sn[i] * h[i + 1];
    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
    h[i] = tmp;
  }
  const double tt = sqrt(h[k] * h[k] + h[k + 1] * h[k + 1]);
  const double c = h[k] / tt, sgn = h[k + 1] / tt;
  h[k] = c * h[k] + sgn * h[k + 1];
  h[k + 1] = 0.0;
  cs[k] = c;
  sn[k] = sgn;
  beta[k + 1] = -sn[k] * beta[k];
  beta[k] = cs[k] * beta[k];
  err[k] = fabs(beta[k + 1]) / bnorm;
}

// y = R^-1 beta[0..K] for the upper triangular leading (K+1) x (K+1) block
// of H (back substitution; one thread, fp64, no contraction).
__global__ void back_subst_kernel(const double* __restrict__ H, const double* __restrict__ beta,
                                  double* __restrict__ y, int K, int m) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  for (int a = K; a >= 0; --a) {
    double s = beta[a];
    for (int c = a + 1; c <= K; ++c) s -= H[(int64_t)c * (m + 1) + a] * y[c];
    y[a] = s / H[(int64_t)a * (m + 1) + a];
  }
}

// One GMRES cycle (x0 = 0) on A (or A^T) with the given precision, up to m
// steps (m <= cy.m); stops when |beta_{k+1}| <= thr * bnorm.  Returns the
// steps taken and the final residual estimate; d += the correction.
int gmres_cycle(Engine& e, Vec& v, Cycle& cy, double* Q, const double* rhs, double bnorm, int m,
                double thr, int prec, bool transpose, double* d, double* est_out,
                std::vector<double>* hist) {
  const int64_t n = e.N;
  const int mm = cy.m;
  // Q[:,0] = rhs / ||rhs||  (scal slot mm+1 holds ||rhs||)
  v.dot(rhs, rhs, mm + 1, 1);
  v.div(Q, rhs, mm + 1);
  cycle_init_kernel<<<1, 256, 0, e.stream>>>(cy.beta.p, mm, v.scal + mm + 1);
  H3D_LAUNCH_CHECK();
  int K = m - 1;
  double error = 0.0;
  for (int k = 0; k < m; ++k) {
    double* qk = Q + (int64_t)k * n;
    double* q = Q + (int64_t)(k + 1) * n;
    matvec_device(e, qk, q, 1, prec, transpose);
    for (int i = 0; i <= k; ++i) {  // modified Gram-Schmidt (src/gmres.cpp:29-38)
      v.dot(q, Q + (int64_t)i * n, i);
      v.axpy(q, Q + (int64_t)i * n, i);
    }
    v.dot(q, q, k + 1, 1);
    v.div(q, q, k + 1);
    givens_kernel<<<1, 64, 0, e.stream>>>(v.scal, cy.H.p, cy.cs.p, cy.sn.p, cy.beta.p, cy.err.p,
                                          k, mm, bnorm);
    H3D_LAUNCH_CHECK();
    H3D_HIP(hipMemcpyAsync(&error, cy.err.p + k, sizeof(double), hipMemcpyDeviceToHost,
                           e.stream));
    H3D_HIP(hipStreamSynchronize(e.stream));
    if (hist) hist->push_back(error);
    if (error <= thr) {
      K = k;
      break;
    }
  }
  back_subst_kernel<<<1, 64, 0, e.stream>>>(cy.H.p, cy.beta.p, cy.y.p, K, mm);
  combine_basis_kernel<<<v.grid, 256, 0, e.stream>>>(d, Q, cy.y.p, K + 1, n);
  H3D_LAUNCH_CHECK();
  *est_out = error;
  return K + 1;
}

}  // namespace

// classGMRES::gmres (src/gmres.cpp:40-94) on the fp64 operator.
int gmres_reference(Engine& e, const double* d_b, double* d_x, int maxit, double thr, double* err,
                    int* its, double* hist) {
  const int64_t n = e.N;
  if (maxit < 1) throw Error(H3D_ERR_INVALID, "maxit must be positive");
  prepare_mode(e, H3D_PREC_FP64);
  Vec v(e, maxit + 4);
  Cycle cy(maxit);
  DevBuf<double> Q((int64_t)(maxit + 1) * n), tmp(n), r(n);
  zero_kernel<<<v.grid, 256, 0, e.stream>>>(d_x, n);
  H3D_LAUNCH_CHECK();
  matvec_device(e, d_x, tmp.p, 1, H3D_PREC_FP64, false);  // A x0 with x0 = 0 (:47)
  v.sub(r.p, d_b, tmp.p);
  v.dot(d_b, d_b, 0, 1);
  const double bnorm = v.get(0);
  v.dot(r.p, r.p, 0, 1);
  const double rnorm = v.get(0);
  std::vector<double> h;
  h.push_back(rnorm / bnorm);
  double est = 0;
  int k = gmres_cycle(e, v, cy, Q.p, r.p, bnorm, maxit, thr, H3D_PREC_FP64, false, d_x, &est, &h);
  H3D_HIP(hipStreamSynchronize(e.stream));
  *err = est;
  *its = k;
  for (size_t i = 0; i < h.size(); ++i) hist[i] = h[i];
  return (int)h.size();
}

int solve_ir(Engine& e, const double* d_b, double* d_x, int restart, int max_inner,
             double inner_tol, int inner_prec, double tol, int max_outer, bool transpose,
             h3d_solve_info* info) {
  const int64_t n = e.N;
  if (restart < 1 || max_inner < 1 || max_outer < 1)
    throw Error(H3D_ERR_INVALID, "restart, max_inner and max_outer must be positive");
  // Outer residual: fp64 factors when resident, otherwise fp64 arithmetic over
  // the fp32-stored factors.
  const int outer_prec = e.fac_bytes[0] > 0 ? H3D_PREC_FP64 : (H3D_PREC_FP32 | H3D_ACC_FP64);
  prepare_mode(e, inner_prec);  // copies and scratch before the solve's own buffers
  prepare_mode(e, outer_prec);
  Vec v(e, restart + 4);
  Cycle cy(restart);
  DevBuf<double> Q((int64_t)(restart + 1) * n), Ax(n), r(n), rr(n), d(n);
  v.dot(d_b, d_b, 0, 1);
  const double bnorm = v.get(0);
  if (!(bnorm > 0)) {
    zero_kernel<<<v.grid, 256, 0, e.stream>>>(d_x, n);
    H3D_LAUNCH_CHECK();
    H3D_HIP(hipStreamSynchronize(e.stream));
    if (info) *info = h3d_solve_info{};
    return 0;
  }
  zero_kernel<<<v.grid, 256, 0, e.stream>>>(d_x, n);
  H3D_LAUNCH_CHECK();
  H3D_HIP(hipMemcpyAsync(r.p, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, e.stream));
  double rel = 1.0;
  int outer = 0, inner_total = 0;
  h3d_solve_info loc{};
  while (outer < max_outer && rel > tol) {
    // inner: restarted GMRES on A d = r, x0 = 0, to inner_tol relative
    zero_kernel<<<v.grid, 256, 0, e.stream>>>(d.p, n);
    H3D_LAUNCH_CHECK();
    H3D_HIP(hipMemcpyAsync(rr.p, r.p, sizeof(double) * n, hipMemcpyDeviceToDevice, e.stream));
    v.dot(r.p, r.p, 0, 1);
    const double r0 = v.get(0);
    int used = 0;
    while (used < max_inner) {
      const int m = std::min(restart, max_inner - used);
      double est = 0;
      int k = gmres_cycle(e, v, cy, Q.p, rr.p, r0, m, inner_tol, inner_prec, transpose, d.p, &est,
                          nullptr);
      used += k;
      // converged by the cycle's own estimate: the outer step recomputes the
      // true residual, so no inner one is formed
      if (k < m && est <= inner_tol) break;
      // restart from the true inner residual rr = r - A_inner d
      matvec_device(e, d.p, Ax.p, 1, inner_prec, transpose);
      v.sub(rr.p, r.p, Ax.p);
      v.dot(rr.p, rr.p, 0, 1);
      if (v.get(0) <= inner_tol * r0) break;
    }
    inner_total += used;
    // x += d; r = b - A x with the outer (fp64) operator
    add_kernel<<<v.grid, 256, 0, e.stream>>>(d_x, d.p, n);
    H3D_LAUNCH_CHECK();
    matvec_device(e, d_x, Ax.p, 1, outer_prec, transpose);
    v.sub(r.p, d_b, Ax.p);
    v.dot(r.p, r.p, 0, 1);
    rel = v.get(0) / bnorm;
    if (outer < 64) loc.history[outer] = rel;
    ++outer;
  }
  loc.outer_iterations = outer;
  loc.outer_prec = outer_prec;
  loc.inner_iterations = inner_total;
  loc.rel_residual = rel;
  if (info) *info = loc;
  return rel <= tol ? 0 : H3D_ERR_NOT_CONVERGED;
}

}  // namespace h3d
