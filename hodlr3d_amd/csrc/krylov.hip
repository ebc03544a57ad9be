// Krylov solvers on the GPU: the reference's unrestarted GMRES and GMRES(m)
// inside fp64 iterative refinement.
//
// Vector work (dots, axpys, norms) runs on the device with a deterministic
// two-stage reduction; only the Hessenberg column (k+2 doubles) crosses to the
// host per Arnoldi step for the Givens rotations (src/gmres.cpp:5-26).
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

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
  double* partial;
  double* scal;  // device scalars
  int64_t n;
  unsigned grid;
  Vec(Engine& en, int nscal) : e(en) {
    n = e.N;
    partial = dalloc<double>(kDotBlocks);
    scal = dalloc<double>(nscal);
    grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  }
  ~Vec() {
    dfree(partial);
    dfree(scal);
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

// One GMRES cycle (x0 = 0) on A (or A^T) with the given precision, up to m
// iterations; stops when |beta_{k+1}| <= thr * bnorm.  Returns iterations and
// the final residual estimate; d receives the correction.
int gmres_cycle(Engine& e, Vec& v, double* Q, const double* rhs, double rnorm_host, double bnorm,
                int m, double thr, int prec, bool transpose, double* d, double* est_out,
                std::vector<double>* hist) {
  const int64_t n = e.N;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), beta(m + 1, 0.0);
  // Q[:,0] = rhs / ||rhs||  (scal slot m+1 holds ||rhs||)
  v.dot(rhs, rhs, m + 1, 1);
  v.div(Q, rhs, m + 1);
  beta[0] = rnorm_host;
  int K = m - 1;
  double error = rnorm_host / bnorm;
  std::vector<double> hv(m + 2);
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
    H3D_HIP(hipMemcpyAsync(hv.data(), v.scal, sizeof(double) * (k + 2), hipMemcpyDeviceToHost,
                           e.stream));
    H3D_HIP(hipStreamSynchronize(e.stream));
    // Givens (src/gmres.cpp:14-26)
    for (int i = 0; i < k; ++i) {
      double tmp = cs[i] * hv[i] + sn[i] * hv[i + 1];
      hv[i + 1] = -sn[i] * hv[i] + cs[i] * hv[i + 1];
      hv[i] = tmp;
    }
    double tt = std::sqrt(hv[k] * hv[k] + hv[k + 1] * hv[k + 1]);
    double cs_k = hv[k] / tt, sn_k = hv[k + 1] / tt;
    hv[k] = cs_k * hv[k] + sn_k * hv[k + 1];
    hv[k + 1] = 0.0;
    for (int i = 0; i <= k + 1; ++i) H[(size_t)k * (m + 1) + i] = hv[i];
    cs[k] = cs_k;
    sn[k] = sn_k;
    beta[k + 1] = -sn[k] * beta[k];
    beta[k] = cs[k] * beta[k];
    error = std::fabs(beta[k + 1]) / bnorm;
    if (hist) hist->push_back(error);
    if (error <= thr) {
      K = k;
      break;
    }
  }
  // back substitution on the rotated (upper triangular) Hessenberg matrix
  std::vector<double> y(K + 1);
  for (int a = K; a >= 0; --a) {
    double s = beta[a];
    for (int c = a + 1; c <= K; ++c) s -= H[(size_t)c * (m + 1) + a] * y[c];
    y[a] = s / H[(size_t)a * (m + 1) + a];
  }
  double* d_y = dalloc<double>(K + 1);
  H3D_HIP(hipMemcpyAsync(d_y, y.data(), sizeof(double) * (K + 1), hipMemcpyHostToDevice, e.stream));
  combine_basis_kernel<<<v.grid, 256, 0, e.stream>>>(d, Q, d_y, K + 1, n);
  H3D_LAUNCH_CHECK();
  H3D_HIP(hipStreamSynchronize(e.stream));
  dfree(d_y);
  *est_out = error;
  return K + 1;
}

}  // namespace

// classGMRES::gmres (src/gmres.cpp:40-94) on the fp64 operator.
int gmres_reference(Engine& e, const double* d_b, double* d_x, int maxit, double thr, double* err,
                    int* its, double* hist) {
  const int64_t n = e.N;
  Vec v(e, maxit + 4);
  double* Q = dalloc<double>((int64_t)(maxit + 1) * n);
  double* tmp = dalloc<double>(n);
  double* r = dalloc<double>(n);
  H3D_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, e.stream));
  matvec_device(e, d_x, tmp, 1, H3D_PREC_FP64, false);  // A x0 with x0 = 0 (:47)
  v.sub(r, d_b, tmp);
  v.dot(d_b, d_b, 0, 1);
  const double bnorm = v.get(0);
  v.dot(r, r, 0, 1);
  const double rnorm = v.get(0);
  std::vector<double> h;
  h.push_back(rnorm / bnorm);
  double est = 0;
  int k = gmres_cycle(e, v, Q, r, rnorm, bnorm, maxit, thr, H3D_PREC_FP64, false, d_x, &est, &h);
  *err = est;
  *its = k;
  for (size_t i = 0; i < h.size(); ++i) hist[i] = h[i];
  dfree(Q);
  dfree(tmp);
  dfree(r);
  return (int)h.size();
}

int solve_ir(Engine& e, const double* d_b, double* d_x, int restart, int max_inner,
             double inner_tol, int inner_prec, double tol, int max_outer, bool transpose,
             h3d_solve_info* info) {
  const int64_t n = e.N;
  if (restart < 1 || max_inner < 1 || max_outer < 1)
    throw Error(H3D_ERR_INVALID, "restart, max_inner and max_outer must be positive");
  Vec v(e, restart + 4);
  double* Q = dalloc<double>((int64_t)(restart + 1) * n);
  double* Ax = dalloc<double>(n);
  double* r = dalloc<double>(n);
  double* rr = dalloc<double>(n);
  double* d = dalloc<double>(n);
  v.dot(d_b, d_b, 0, 1);
  const double bnorm = v.get(0);
  if (!(bnorm > 0)) {
    H3D_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, e.stream));
    if (info) *info = h3d_solve_info{};
    dfree(Q), dfree(Ax), dfree(r), dfree(rr), dfree(d);
    return 0;
  }
  // Outer residual: fp64 factors when resident, otherwise fp64 arithmetic over
  // the fp32-stored factors.
  const int outer_prec = e.fac_bytes[0] > 0 ? H3D_PREC_FP64 : (H3D_PREC_FP32 | H3D_ACC_FP64);
  H3D_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, e.stream));
  H3D_HIP(hipMemcpyAsync(r, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, e.stream));
  double rel = 1.0;
  int outer = 0, inner_total = 0;
  h3d_solve_info loc{};
  while (outer < max_outer && rel > tol) {
    // inner: restarted GMRES on A d = r, x0 = 0, to inner_tol relative
    H3D_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, e.stream));
    H3D_HIP(hipMemcpyAsync(rr, r, sizeof(double) * n, hipMemcpyDeviceToDevice, e.stream));
    v.dot(r, r, 0, 1);
    const double r0 = v.get(0);
    double cur = r0;
    int used = 0;
    while (used < max_inner && cur > inner_tol * r0) {
      const int m = std::min(restart, max_inner - used);
      double est = 0;
      int k = gmres_cycle(e, v, Q, rr, cur, r0, m, inner_tol, inner_prec, transpose, d, &est,
                          nullptr);
      used += k;
      // converged by the cycle's own estimate: the outer step recomputes the
      // true residual, so no inner one is formed
      if (k < m && est <= inner_tol) break;
      // restart from the true inner residual rr = r - A_inner d
      matvec_device(e, d, Ax, 1, inner_prec, transpose);
      v.sub(rr, r, Ax);
      v.dot(rr, rr, 0, 1);
      cur = v.get(0);
    }
    inner_total += used;
    // x += d; r = b - A x with the fp64 operator
    add_kernel<<<v.grid, 256, 0, e.stream>>>(d_x, d, n);
    H3D_LAUNCH_CHECK();
    matvec_device(e, d_x, Ax, 1, outer_prec, transpose);
    v.sub(r, d_b, Ax);
    v.dot(r, r, 0, 1);
    rel = v.get(0) / bnorm;
    if (outer < 64) loc.history[outer] = rel;
    ++outer;
  }
  loc.outer_iterations = outer;
  loc.outer_prec = outer_prec;
  loc.inner_iterations = inner_total;
  loc.rel_residual = rel;
  if (info) *info = loc;
  dfree(Q);
  dfree(Ax);
  dfree(r);
  dfree(rr);
  dfree(d);
  return rel <= tol ? 0 : H3D_ERR_NOT_CONVERGED;
}

}  // namespace h3d
