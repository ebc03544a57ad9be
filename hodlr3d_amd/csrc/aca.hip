// Batched, pivot-exact ACA on the GPU and the layout of its factors.
//
// One workgroup runs LowRank::ACA_only_nodesCUR (src/ACA.cpp:127-238) for one
// block (rows = target box, columns = source box).  The decisions follow the
// reference step by step:
//   * first row: the first row whose largest entry (over all columns) has
//     magnitude > 1e-16 (ACA.cpp:152-161);
//   * argmax over the remaining rows/columns, first maximum wins
//     (maxAbsVector, ACA.cpp:10-23);
//   * residual row/column: kernel entry minus sum_l u_l[i] v_l[s], l outer,
//     no FMA (ACA.cpp:181-193); u = column / pivot;
//   * reject when ||u|| or ||v|| < 1e-16 (ACA.cpp:196-200), or when the
//     pivot is exactly zero (there the reference divides by zero);
//   * normS^2 = normS^2 + (|u||v|)^2 + sum_{l<=r} 2|u_l.u v_l.v| including the
//     new pair (ACA.cpp:206-212); stop when |u||v| <= tol normS or r = min(n,m);
//   * next row = argmax of the raw residual column over the remaining rows.
// Norms and dots use the canonical order (element s to lane s%256, lanes
// summed in increasing s, then a pairwise tree over 256 lanes) that the CPU
// oracle also uses, so ranks and pivots compare bit for bit.  This file is
// compiled with -ffp-contract=off.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "kernels.hpp"

namespace h3d {

namespace {

struct AcaDesc {
  int32_t rs, n, cs, m;
  int64_t uw, vw;   // workspace offsets (doubles): u_l at uw + l*n, v_l at vw + l*m
  int32_t rcap;     // rank capacity of the workspace
  int32_t pw;       // pivot scratch offset (I at pw, J at pw + rcap)
};

struct Arg {
  double a;   // |value|
  double v;   // value
  int idx;    // -1: none
};

__device__ inline bool arg_better(const Arg& x, const Arg& y) {  // y beats x?
  if (y.idx < 0) return false;
  if (x.idx < 0) return true;
  if (y.a > x.a) return true;
  return (y.a == x.a) && (y.idx < x.idx);
}

__device__ inline Arg shfl_xor_arg(const Arg& x, int d) {
  Arg o;
  o.a = __shfl_xor(x.a, d);
  o.v = __shfl_xor(x.v, d);
  o.idx = __shfl_xor(x.idx, d);
  return o;
}

// Workgroup-wide argmax with first-maximum tie break; result in all threads.
template <int T>
__device__ Arg block_argmax(Arg x, double* lds) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    Arg o = shfl_xor_arg(x, d);
    if (arg_better(x, o)) x = o;
  }
  if constexpr (T == 64) {
    return x;
  } else {
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
      lds[3 * w] = x.a;
      lds[3 * w + 1] = x.v;
      lds[3 * w + 2] = (double)x.idx;
    }
    __syncthreads();
    Arg best;
    best.idx = -1;
    best.a = 0;
    best.v = 0;
    for (int q = 0; q < T / 64; ++q) {
      Arg o;
      o.a = lds[3 * q];
      o.v = lds[3 * q + 1];
      o.idx = (int)lds[3 * q + 2];
      if (arg_better(best, o)) best = o;
    }
    __syncthreads();
    return best;
  }
}

// Canonical 256-lane reduction of NV sums.  a[q][p] is virtual lane
// threadIdx.x + T*p of sum q.  The result is returned in every thread.
template <int T, int NV>
__device__ void canon_reduce(double (&a)[NV][256 / T], double* lds, double (&out)[NV]) {
  const int t = threadIdx.x;
  double v[NV];
  if constexpr (T == 64) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x0 = a[q][0] + a[q][2];  // d = 128
      double x1 = a[q][1] + a[q][3];
      v[q] = x0 + x1;                 // d = 64
    }
  } else {
    static_assert(T == 256, "T must be 64 or 256");
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q * 256 + t] = a[q][0];
    __syncthreads();
    if (t < 128) {
#pragma unroll
      for (int q = 0; q < NV; ++q) lds[q * 256 + t] = lds[q * 256 + t] + lds[q * 256 + t + 128];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = (t < 64) ? lds[q * 256 + t] + lds[q * 256 + t + 64] : 0.0;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = v[q] + __shfl_down(v[q], d);
  if constexpr (T == 64) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = readlane_d(v[q], 0);
  } else {
    __syncthreads();
    if (t == 0)
#pragma unroll
      for (int q = 0; q < NV; ++q) lds[q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = lds[q];
    __syncthreads();
  }
}

template <class KF>
__device__ inline double kentry(const KernelDev& kd, double xa, double ya, double za, double xb,
                                double yb, double zb) {
  return kd.scale * kpair<KF>(kd.q, xa, ya, za, xb, yb, zb);
}

__device__ inline bool bit_get(const uint32_t* m, int s) { return (m[s >> 5] >> (s & 31)) & 1u; }

// x - sum_{l < r} c[l] * Z[l * ld], subtracted in l order (the reference's
// order, no FMA), with sixteen loads of Z issued before their subtractions so
// that a lane has many memory requests in flight.
__device__ inline double sub_sum(double x, const double* c, const double* __restrict__ Z,
                                 int64_t ld, int r) {
  int l = 0;
  for (; l + 16 <= r; l += 16) {
    double z[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) z[q] = Z[(int64_t)(l + q) * ld];
#pragma unroll
    for (int q = 0; q < 16; ++q) x -= c[l + q] * z[q];
  }
  for (; l + 4 <= r; l += 4) {
    double z[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) z[q] = Z[(int64_t)(l + q) * ld];
#pragma unroll
    for (int q = 0; q < 4; ++q) x -= c[l + q] * z[q];
  }
  for (; l < r; ++l) x -= c[l] * Z[(int64_t)l * ld];
  return x;
}

template <int T, class KF>
__global__ __launch_bounds__(T) void aca_kernel(const AcaDesc* __restrict__ descs,
                                                const double* __restrict__ px,
                                                const double* __restrict__ py,
                                                const double* __restrict__ pz, KernelDev kd,
                                                double tol, double* __restrict__ Uws,
                                                double* __restrict__ Vws,
                                                int32_t* __restrict__ pivws,
                                                int32_t* __restrict__ rank_out,
                                                int32_t* __restrict__ status_out) {
  extern __shared__ double smem[];
  constexpr int VPT = 256 / T;
  const AcaDesc d = descs[blockIdx.x];
  const int t = threadIdx.x;
  const int n = d.n, m = d.m;
  double* red = smem;           // 8 x 256 (T = 256; the one-wave kernel reduces in registers)
  double* amx = red + (T == 64 ? 0 : 2048);  // 3 x (T/64)
  double* uil = amx + 16;       // rcap
  double* vjl = uil + d.rcap;   // rcap
  uint32_t* rmask = (uint32_t*)(vjl + d.rcap);
  const int nwr = (n + 31) >> 5, nwc = (m + 31) >> 5;
  uint32_t* cmask = rmask + nwr;
  for (int w = t; w < nwr; w += T) rmask[w] = 0xffffffffu;
  for (int w = t; w < nwc; w += T) cmask[w] = 0xffffffffu;
  double* U = Uws + d.uw;
  double* V = Vws + d.vw;
  int32_t* PI = pivws + d.pw;
  int32_t* PJ = PI + d.rcap;
  const double* rx = px + d.rs;
  const double* ry = py + d.rs;
  const double* rz = pz + d.rs;
  const double* cx = px + d.cs;
  const double* cy = py + d.cs;
  const double* cz = pz + d.cs;
  const int mn = min(n, m);
  int nonfinite = 0;
  __syncthreads();

  // ---- first pair (ACA.cpp:152-171)
  int i = -1, jj = -1;
  double piv = 0.0, vv = 0.0, uu = 0.0;
  double acc[2][VPT];
  for (int l = 0; l < n && i < 0; ++l) {
    const double xi = rx[l], yi = ry[l], zi = rz[l];
#pragma unroll
    for (int p = 0; p < VPT; ++p) acc[0][p] = 0.0;
    Arg best;
    best.idx = -1;
    best.a = 0;
    best.v = 0;
    for (int base = 0; base < m; base += 256)
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        const int s = base + t + T * p;
        if (s >= m) continue;
        double val = kentry<KF>(kd, xi, yi, zi, cx[s], cy[s], cz[s]);
        V[s] = val;
        nonfinite |= !isfinite(val);
        acc[0][p] = acc[0][p] + val * val;
        double a = fabs(val);
        if (best.idx < 0 || a > best.a) {
          best.a = a;
          best.v = val;
          best.idx = s;
        }
      }
    Arg g = block_argmax<T>(best, amx);
    if (fabs(g.v) > 1e-16) {
      i = l;
      jj = g.idx;
      piv = g.v;
      double o[2];
      double a2[2][VPT];
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        a2[0][p] = acc[0][p];
        a2[1][p] = 0.0;
      }
      canon_reduce<T, 2>(a2, red, o);
      vv = o[0];
    }
  }
  if (i < 0) {
    if (t == 0) {
      rank_out[blockIdx.x] = 0;
      status_out[blockIdx.x] = nonfinite ? 1 : 0;
    }
    return;
  }
  {
    const double xj = cx[jj], yj = cy[jj], zj = cz[jj];
#pragma unroll
    for (int p = 0; p < VPT; ++p) acc[0][p] = 0.0;
    Arg best;
    best.idx = -1;
    best.a = 0;
    best.v = 0;
    for (int base = 0; base < n; base += 256)
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        const int s = base + t + T * p;
        if (s >= n) continue;
        double col = kentry<KF>(kd, rx[s], ry[s], rz[s], xj, yj, zj);
        nonfinite |= !isfinite(col);
        double u = col / piv;
        U[s] = u;
        acc[0][p] = acc[0][p] + u * u;
        if (s != i) {
          double a = fabs(col);
          if (best.idx < 0 || a > best.a) {
            best.a = a;
            best.v = col;
            best.idx = s;
          }
        }
      }
    Arg g = block_argmax<T>(best, amx);
    double o[2];
#pragma unroll
    for (int p = 0; p < VPT; ++p) acc[1][p] = 0.0;
    canon_reduce<T, 2>(acc, red, o);
    uu = o[0];
    if (t == 0) {
      PI[0] = i;
      PJ[0] = jj;
      rmask[i >> 5] &= ~(1u << (i & 31));
      cmask[jj >> 5] &= ~(1u << (jj & 31));
    }
    __syncthreads();
    i = g.idx;  // next row (may be -1 when n == 1)
  }
  int r = 1;
  int status = 0;
  double normS = 0.0;
  double nu = sqrt(uu), nv = sqrt(vv);

  // ---- main loop (ACA.cpp:173-214)
  while (nu * nv > tol * normS && r < mn) {
    if (r == d.rcap) {
      status |= 2;  // workspace capacity reached before convergence
      break;
    }
    for (int l = t; l < r; l += T) uil[l] = U[(int64_t)l * n + i];
    __syncthreads();
    const double xi = rx[i], yi = ry[i], zi = rz[i];
    double* vr = V + (int64_t)r * m;
#pragma unroll
    for (int p = 0; p < VPT; ++p) acc[0][p] = 0.0;
    Arg best;
    best.idx = -1;
    best.a = 0;
    best.v = 0;
    for (int base = 0; base < m; base += 256)
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        const int s = base + t + T * p;
        if (s >= m) continue;
        double val = kentry<KF>(kd, xi, yi, zi, cx[s], cy[s], cz[s]);
        val = sub_sum(val, uil, V + s, m, r);
        vr[s] = val;
        nonfinite |= !isfinite(val);
        acc[0][p] = acc[0][p] + val * val;
        if (bit_get(cmask, s)) {
          double a = fabs(val);
          if (best.idx < 0 || a > best.a) {
            best.a = a;
            best.v = val;
            best.idx = s;
          }
        }
      }
    Arg g = block_argmax<T>(best, amx);
    jj = g.idx;
    piv = g.v;
    // A residual row that vanishes on every remaining column (a block with
    // coincident points, rank reached exactly): the reference would divide by
    // this zero pivot (NaN factors); stop as its small-norm test does.
    if (piv == 0.0) break;
    for (int l = t; l < r; l += T) vjl[l] = V[(int64_t)l * m + jj];
    __syncthreads();
    const double xj = cx[jj], yj = cy[jj], zj = cz[jj];
    double* ur = U + (int64_t)r * n;
#pragma unroll
    for (int p = 0; p < VPT; ++p) acc[1][p] = 0.0;
    best.idx = -1;
    best.a = 0;
    best.v = 0;
    for (int base = 0; base < n; base += 256)
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        const int s = base + t + T * p;
        if (s >= n) continue;
        double col = kentry<KF>(kd, rx[s], ry[s], rz[s], xj, yj, zj);
        col = sub_sum(col, vjl, U + s, n, r);
        nonfinite |= !isfinite(col);
        double u = col / piv;
        ur[s] = u;
        acc[1][p] = acc[1][p] + u * u;
        if (s != i && bit_get(rmask, s)) {
          double a = fabs(col);
          if (best.idx < 0 || a > best.a) {
            best.a = a;
            best.v = col;
            best.idx = s;
          }
        }
      }
    Arg gi = block_argmax<T>(best, amx);
    double o[2];
    canon_reduce<T, 2>(acc, red, o);
    vv = o[0];
    uu = o[1];
    double nu2 = sqrt(uu), nv2 = sqrt(vv);
    if (nu2 < 1e-16 || nv2 < 1e-16) break;
    if (t == 0) {
      PI[r] = i;
      PJ[r] = jj;
      rmask[i >> 5] &= ~(1u << (i & 31));
      cmask[jj >> 5] &= ~(1u << (jj & 31));
    }
    ++r;
    nu = nu2;
    nv = nv2;
    double ns2 = normS * normS + (nu * nv) * (nu * nv);
    // dots of the previous vectors with the new pair, in l order
    // four previous vectors at a time (each dot still summed per lane in s
    // order and reduced canonically; ns2 gathers them in l order)
    int l0 = 0;
    for (; l0 + 4 <= r - 1; l0 += 4) {
      double a8[8][VPT];
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int p = 0; p < VPT; ++p) a8[q][p] = 0.0;
      // two rows per lane and iteration: all ten loads first, then the sums
      // in row order
      auto dots4 = [&](const double* __restrict__ Z, const double* __restrict__ zr, int len,
                       int q0) {
        for (int base = 0; base < len; base += 512)
#pragma unroll
          for (int p = 0; p < VPT; ++p) {
            const int s0 = base + t + T * p, s1 = s0 + 256;
            double x[2] = {0.0, 0.0}, w[2][4];
            const bool in0 = s0 < len, in1 = s1 < len;
            if (in0) {
              x[0] = zr[s0];
#pragma unroll
              for (int j = 0; j < 4; ++j) w[0][j] = Z[(int64_t)(l0 + j) * len + s0];
            }
            if (in1) {
              x[1] = zr[s1];
#pragma unroll
              for (int j = 0; j < 4; ++j) w[1][j] = Z[(int64_t)(l0 + j) * len + s1];
            }
            if (in0)
#pragma unroll
              for (int j = 0; j < 4; ++j) a8[q0 + 2 * j][p] = a8[q0 + 2 * j][p] + w[0][j] * x[0];
            if (in1)
#pragma unroll
              for (int j = 0; j < 4; ++j) a8[q0 + 2 * j][p] = a8[q0 + 2 * j][p] + w[1][j] * x[1];
          }
      };
      dots4(U, ur, n, 0);
      dots4(V, vr, m, 1);
      double dd[8];
      canon_reduce<T, 8>(a8, red, dd);
#pragma unroll
      for (int j = 0; j < 4; ++j) ns2 += 2 * fabs(dd[2 * j] * dd[2 * j + 1]);
    }
    for (int l = l0; l < r - 1; ++l) {
      double a2[2][VPT];
#pragma unroll
      for (int p = 0; p < VPT; ++p) {
        a2[0][p] = 0.0;
        a2[1][p] = 0.0;
      }
      const double* ul = U + (int64_t)l * n;
      const double* vl = V + (int64_t)l * m;
      for (int base = 0; base < n; base += 256)
#pragma unroll
        for (int p = 0; p < VPT; ++p) {
          const int s = base + t + T * p;
          if (s < n) a2[0][p] = a2[0][p] + ul[s] * ur[s];
        }
      for (int base = 0; base < m; base += 256)
#pragma unroll
        for (int p = 0; p < VPT; ++p) {
          const int s = base + t + T * p;
          if (s < m) a2[1][p] = a2[1][p] + vl[s] * vr[s];
        }
      double dd[2];
      canon_reduce<T, 2>(a2, red, dd);
      ns2 += 2 * fabs(dd[0] * dd[1]);
    }
    ns2 += 2 * fabs(uu * vv);  // l = r-1: u.u and v.v
    normS = sqrt(ns2);
    i = gi.idx;
    __syncthreads();
  }
  if (t == 0) {
    rank_out[blockIdx.x] = r;
    status_out[blockIdx.x] = (nonfinite ? 1 : 0) | status;
  }
}

// Moves one finished block from the ACA workspace into its chunk: u-vectors
// as consecutive columns of the target box's U (column-major, contiguous),
// v-vectors transposed to row-major m x r, pivots I then J.  Every stored
// precision is written from the fp64 values.
struct WriteDesc {
  int64_t uw, vw;  // workspace offsets (doubles)
  int64_t uo, vo;  // chunk offsets (elements) of the block's U columns / V
  int64_t po;      // pivot offset
  int32_t n, m, r, pw, rcap, pad;
};

__global__ void write_kernel(const WriteDesc* __restrict__ ds, const double* __restrict__ Uws,
                             const double* __restrict__ Vws, const int32_t* __restrict__ pivws,
                             double* __restrict__ U64, double* __restrict__ V64,
                             float* __restrict__ U32, float* __restrict__ V32,
                             uint16_t* __restrict__ U16, uint16_t* __restrict__ V16,
                             int32_t* __restrict__ piv) {
  const WriteDesc d = ds[blockIdx.x];
  const int ld = aca_ldu(d.n), ldv = d.m + (d.m & 1);
  const int64_t nu = (int64_t)d.r * d.n;
  for (int64_t q = threadIdx.x; q < nu; q += blockDim.x) {
    const int64_t l = q / d.n;
    const int64_t o = d.uo + l * ld + (q - l * d.n);
    const double v = Uws[d.uw + q];
    if (U64) U64[o] = v;
    if (U32) U32[o] = (float)v;
    if (U16) U16[o] = f_to_bf16((float)v);
  }
  const int64_t nv = (int64_t)d.r * d.m;
  for (int64_t q = threadIdx.x; q < nv; q += blockDim.x) {
    const int64_t l = q / d.m;
    const int64_t o = d.vo + l * ldv + (q - l * d.m);
    const double v = Vws[d.vw + q];
    if (V64) V64[o] = v;
    if (V32) V32[o] = (float)v;
    if (V16) V16[o] = f_to_bf16((float)v);
  }
  for (int l = threadIdx.x; l < d.r; l += blockDim.x) {
    piv[d.po + l] = pivws[d.pw + l];
    piv[d.po + d.r + l] = pivws[d.pw + d.rcap + l];
  }
}

size_t aca_smem(int T, int rcap, int n, int m) {
  return sizeof(double) * ((T == 64 ? 0 : 2048) + 16 + 2 * (size_t)rcap) +
         sizeof(uint32_t) * (size_t)(((n + 31) >> 5) + ((m + 31) >> 5)) + 16;
}

// Grow-only device workspace of one ACA pass index, reused by every batch:
// allocating tens of GB per pass and batch took seconds of the assembly.  It
// also holds the pass's block descriptors, ranks and statuses and the write
// descriptors, so that no device memory is freed while the assembly runs
// (hipFree waits for the whole device, which would stall a pass running on
// another stream).
struct AcaWorkspace {
  double* U = nullptr;
  double* V = nullptr;
  int32_t* P = nullptr;
  AcaDesc* D = nullptr;
  int32_t* R = nullptr;
  int32_t* S = nullptr;
  WriteDesc* WD = nullptr;
  int64_t nU = 0, nV = 0, nP = 0, nD = 0, nWD = 0;
  template <class T>
  static void grow(T*& p, int64_t& cap, int64_t n) {
    if (n <= cap) return;
    dfree(p);
    cap = 0;
    p = dalloc<T>(n);
    cap = n;
  }
  void reserve(int64_t u, int64_t v, int64_t p) {
    grow(U, nU, u);
    grow(V, nV, v);
    grow(P, nP, p);
  }
  void reserve_desc(int64_t n) {
    if (n <= nD) return;
    dfree(D);
    dfree(R);
    dfree(S);
    nD = 0;
    D = dalloc<AcaDesc>(n);
    R = dalloc<int32_t>(n);
    S = dalloc<int32_t>(n);
    nD = n;
  }
  ~AcaWorkspace() {
    dfree(U);
    dfree(V);
    dfree(P);
    dfree(D);
    dfree(R);
    dfree(S);
    dfree(WD);
  }
};

// One ACA pass over a batch of blocks; U, V, P point into the workspace of
// its pass index and stay valid until the batch's factors are written.
struct Pass {
  double* U = nullptr;
  double* V = nullptr;
  int32_t* P = nullptr;
  std::vector<AcaDesc> ds;
  std::vector<int64_t> ids;  // CSR block index within the level
  std::vector<int32_t> rank, status;
  // launched on `st`; per class (one-wave / four-wave kernel) the position
  // of its blocks in ds and their range in the workspace's D, R, S
  hipStream_t st = nullptr;
  AcaWorkspace* ws = nullptr;  // holds its descriptors, ranks and statuses
  std::vector<AcaDesc> sub[2];
  std::vector<int64_t> where[2];
  int64_t off[2] = {0, 0};
};

// Rank capacity of the ACA workspace: a block that would need more stops
// with status 2 and is run again from scratch with the next capacity.
int cap_for(int pass, int n, int m) {
  const int mn = std::min(n, m);
  const int big = std::max(n, m);
  const int first = big > 32768 ? 256 : (big > 256 ? 128 : 96);
  int c = pass == 0 ? first : (pass == 1 ? 512 : mn);
  return std::min(c, mn);
}

void launch_pass(Engine& e, Pass& P, AcaWorkspace& ws, hipStream_t st) {
  const int64_t nb = (int64_t)P.ds.size();
  P.st = st;
  P.ws = &ws;
  P.rank.assign(nb, 0);
  P.status.assign(nb, 0);
  if (nb == 0) return;
  int64_t wsU = 0, wsV = 0, pw = 0;
  for (auto& d : P.ds) {
    wsU = std::max<int64_t>(wsU, d.uw + (int64_t)d.rcap * d.n);
    wsV = std::max<int64_t>(wsV, d.vw + (int64_t)d.rcap * d.m);
    pw = std::max<int64_t>(pw, d.pw + 2 * (int64_t)d.rcap);
  }
  ws.reserve(wsU, wsV, pw);
  ws.reserve_desc(nb);
  P.U = ws.U;
  P.V = ws.V;
  P.P = ws.P;
  const bool fast = e.kern.q == 7 || e.kern.q == 16;
  // two launches: small blocks on one wave, large blocks on four
  int64_t off = 0;
  for (int cls = 0; cls < 2; ++cls) {
    std::vector<AcaDesc>& sub = P.sub[cls];
    std::vector<int64_t>& where = P.where[cls];
    sub.clear();
    where.clear();
    size_t smax = 0;
    for (int64_t q = 0; q < nb; ++q) {
      const AcaDesc& d = P.ds[q];
      const int c = std::max(d.n, d.m) <= 256 ? 0 : 1;
      if (c != cls) continue;
      sub.push_back(d);
      where.push_back(q);
      smax = std::max(smax, aca_smem(cls == 0 ? 64 : 256, d.rcap, d.n, d.m));
    }
    P.off[cls] = off;
    if (sub.empty()) continue;
    if (smax > 160 * 1024) throw Error(H3D_ERR_INVALID, "ACA block too large for LDS");
    const bool plug = e.kern.q == kQPlugin;
    const void* kfun = cls == 0 ? (fast   ? (const void*)aca_kernel<64, InvR>
                                   : plug ? (const void*)aca_kernel<64, Plugin>
                                          : (const void*)aca_kernel<64, AnyQ>)
                                : (fast   ? (const void*)aca_kernel<256, InvR>
                                   : plug ? (const void*)aca_kernel<256, Plugin>
                                          : (const void*)aca_kernel<256, AnyQ>);
    if (smax > 64 * 1024)
      H3D_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smax));
    AcaDesc* d_ds = ws.D + off;
    H3D_HIP(hipMemcpyAsync(d_ds, sub.data(), sizeof(AcaDesc) * sub.size(), hipMemcpyHostToDevice, st));
    for (size_t o = 0; o < sub.size(); o += 1000000) {
      const unsigned g = (unsigned)std::min<size_t>(1000000, sub.size() - o);
      AcaDesc* dd = d_ds + o;
      int32_t* rr = ws.R + off + o;
      int32_t* ss = ws.S + off + o;
      void* args[] = {&dd, &e.d_px, &e.d_py, &e.d_pz, &e.kern, &e.tol, &P.U, &P.V, &P.P, &rr, &ss};
      H3D_HIP(hipLaunchKernel(kfun, dim3(g), dim3(cls == 0 ? 64 : 256), args, smax, st));
    }
    off += (int64_t)sub.size();
  }
}

// Waits for a launched pass and reads its ranks and statuses.
void collect_pass(Pass& P) {
  if (!P.ws) return;  // no blocks
  const AcaWorkspace& ws = *P.ws;
  for (int cls = 0; cls < 2; ++cls) {
    const size_t n = P.sub[cls].size();
    if (!n) continue;
    std::vector<int32_t> rk(n), sv(n);
    H3D_HIP(hipMemcpyAsync(rk.data(), ws.R + P.off[cls], sizeof(int32_t) * n, hipMemcpyDeviceToHost, P.st));
    H3D_HIP(hipMemcpyAsync(sv.data(), ws.S + P.off[cls], sizeof(int32_t) * n, hipMemcpyDeviceToHost, P.st));
    H3D_HIP(hipStreamSynchronize(P.st));
    for (size_t q = 0; q < n; ++q) {
      P.rank[P.where[cls][q]] = rk[q];
      P.status[P.where[cls][q]] = sv[q];
    }
  }
}

// End of the batch of complete target boxes of level j starting at k0 whose
// first-pass workspace fits the budget.
int64_t batch_end(const Engine& e, int j, int64_t k0, int64_t budget) {
  const Level& lv = e.lev[j];
  const int64_t lo = level_offset(j);
  int64_t k1 = k0, bytes = 0;
  while (k1 < lv.nboxes) {
    int64_t need = 0;
    const int n = e.bcount[lo + k1];
    for (int32_t b = lv.iloff[k1]; b < lv.iloff[k1 + 1]; ++b) {
      const int m = e.bcount[lo + lv.ilidx[b]];
      need += (int64_t)cap_for(0, n, m) * (n + m) * 8;
    }
    if (k1 > k0 && (bytes + need > budget || lv.iloff[k1 + 1] - lv.iloff[k0] > 2000000)) break;
    bytes += need;
    ++k1;
  }
  return k1;
}

// Blocks of level j with a nonempty target in boxes [k0, k1), and their
// descriptors for `pass`.
std::vector<int64_t> batch_blocks(const Engine& e, int j, int64_t k0, int64_t k1) {
  const Level& lv = e.lev[j];
  const int64_t lo = level_offset(j);
  std::vector<int64_t> todo;
  for (int64_t k = k0; k < k1; ++k) {
    const int n = e.bcount[lo + k];
    for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b)
      if (n > 0 && e.bcount[lo + lv.ilidx[b]] > 0) todo.push_back(b);
  }
  return todo;
}

void fill_pass(const Engine& e, int j, const std::vector<int64_t>& todo, int pass, Pass& P) {
  const Level& lv = e.lev[j];
  const int64_t lo = level_offset(j);
  int64_t wsU = 0, wsV = 0;
  int32_t pw = 0;
  for (int64_t b : todo) {
    int32_t k = (int32_t)(std::upper_bound(lv.iloff.begin(), lv.iloff.end(), (int32_t)b) -
                          lv.iloff.begin() - 1);
    const int32_t ki = lv.ilidx[b];
    AcaDesc d;
    d.rs = e.bstart[lo + k];
    d.n = e.bcount[lo + k];
    d.cs = e.bstart[lo + ki];
    d.m = e.bcount[lo + ki];
    d.rcap = cap_for(pass, d.n, d.m);
    d.uw = wsU;
    d.vw = wsV;
    d.pw = pw;
    wsU += (int64_t)d.rcap * d.n;
    wsV += (int64_t)d.rcap * d.m;
    pw += 2 * d.rcap;
    P.ds.push_back(d);
    P.ids.push_back(b);
  }
}

// ACA of the batch [k0, k1) of level j (pass 0 given when it was launched
// earlier), chunk layout and factor write-out.
void assemble_batch(Engine& e, int j, int64_t k0, int64_t k1, AcaWorkspace (&ws)[3],
                    Pass* early) {
  hipStream_t st = e.stream;
  Level& lv = e.lev[j];
  const int64_t lo = level_offset(j);
  std::vector<int64_t> todo = batch_blocks(e, j, k0, k1);
  std::vector<Pass> passes;
  for (int pass = 0; pass < 3 && !todo.empty(); ++pass) {
    if (pass == 0 && early) {
      passes.push_back(std::move(*early));  // launched before, on its own stream and workspace
      collect_pass(passes.back());
    } else {
      passes.emplace_back();
      fill_pass(e, j, todo, pass, passes.back());
      launch_pass(e, passes.back(), ws[pass], st);
      collect_pass(passes.back());
    }
    Pass& P = passes.back();
    std::vector<int64_t> retry;
    for (size_t q = 0; q < P.ids.size(); ++q) {
      const int64_t b = P.ids[q];
      if (P.status[q] & 2) {
        retry.push_back(b);
        P.rank[q] = -1;  // not final
        continue;
      }
      if (P.status[q] & 1) throw Error(H3D_ERR_NONFINITE, "non-finite kernel entry in ACA");
      lv.rank[b] = P.rank[q];
      lv.status[b] = P.status[q];
    }
    todo.swap(retry);
  }
  if (!todo.empty()) throw Error(H3D_ERR_INTERNAL, "ACA did not finish within capacity");
    // chunk layout for boxes [k0, k1)
    Chunk ch;
    for (int64_t k = k0; k < k1; ++k) {
      const int n = e.bcount[lo + k];
      int32_t R = 0;
      for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
        const int r = lv.rank[b];
        const int m = e.bcount[lo + lv.ilidx[b]];
        lv.col[b] = R;
        R += r;
        lv.voff[b] = ch.nV;
        ch.nV += (int64_t)r * (m + (m & 1));
        lv.pivoff[b] = ch.npiv;
        ch.npiv += 2 * (int64_t)r;
      }
      lv.Rt[k] = R;
      lv.uoff[k] = aca_ustart(ch.nU, n);
      ch.nU = lv.uoff[k] + (int64_t)aca_ldu(n) * R;
      lv.chunk[k] = (int32_t)e.chunks.size();
    }
    for (int p = 0; p < 3; ++p) {
      if (!(e.store_mask & (1 << p))) continue;
      const size_t bytes = (size_t)(ch.nU + ch.nV + kChunkTail) * elem_size(p);
      dmalloc(&ch.data[p], bytes);
      H3D_HIP(hipMemsetAsync(ch.data[p], 0, bytes, st));
      e.fac_bytes[p] += (ch.nU + ch.nV) * elem_size(p);
    }
    ch.piv = dalloc<int32_t>(std::max<int64_t>(ch.npiv, 1));
    for (Pass& P : passes) {
      std::vector<WriteDesc> wd;
      for (size_t q = 0; q < P.ids.size(); ++q) {
        if (P.rank[q] <= 0) continue;
        const int64_t b = P.ids[q];
        int32_t k = (int32_t)(std::upper_bound(lv.iloff.begin(), lv.iloff.end(), (int32_t)b) -
                              lv.iloff.begin() - 1);
        WriteDesc w;
        w.uw = P.ds[q].uw;
        w.vw = P.ds[q].vw;
        w.n = P.ds[q].n;
        w.m = P.ds[q].m;
        w.r = P.rank[q];
        w.pw = P.ds[q].pw;
        w.rcap = P.ds[q].rcap;
        w.pad = 0;
        w.uo = lv.uoff[k] + (int64_t)lv.col[b] * aca_ldu(w.n);
        w.vo = lv.voff[b];
        w.po = lv.pivoff[b];
        wd.push_back(w);
      }
      if (!wd.empty()) {
        AcaWorkspace::grow(ws[0].WD, ws[0].nWD, (int64_t)wd.size());
        WriteDesc* d_wd = ws[0].WD;
        H3D_HIP(hipMemcpyAsync(d_wd, wd.data(), sizeof(WriteDesc) * wd.size(),
                               hipMemcpyHostToDevice, st));
        double* U64 = (double*)ch.data[0];
        double* V64 = U64 ? U64 + ch.nU : nullptr;
        float* U32 = (float*)ch.data[1];
        float* V32 = U32 ? U32 + ch.nU : nullptr;
        uint16_t* U16 = (uint16_t*)ch.data[2];
        uint16_t* V16 = U16 ? U16 + ch.nU : nullptr;
        for (size_t off = 0; off < wd.size(); off += 1000000) {
          const unsigned g = (unsigned)std::min<size_t>(1000000, wd.size() - off);
          write_kernel<<<g, 256, 0, st>>>(d_wd + off, P.U, P.V, P.P, U64, V64, U32, V32, U16,
                                          V16, ch.piv);
          H3D_LAUNCH_CHECK();
        }
        H3D_HIP(hipStreamSynchronize(st));
      }
    }
    e.chunks.push_back(ch);
}

}  // namespace

void assemble_all(Engine& e) {
  auto t0 = std::chrono::steady_clock::now();
  const int nL = e.nL;
  size_t freeb = 0, totb = 0;
  H3D_HIP(hipMemGetInfo(&freeb, &totb));
  int64_t budget = e.ws_budget > 0 ? e.ws_budget
                                   : std::min<int64_t>((int64_t)(freeb / 8), int64_t(24) << 30);
  budget = std::max<int64_t>(budget, int64_t(64) << 20);

  e.T = 0;
  e.max_rank = 0;
  e.n_blocks = e.rank_sum = e.factor_elems = e.u_elems = e.v_elems = 0;
  e.find_memory = 0;
  e.nblk_all = 0;
  for (int p = 0; p < 3; ++p) e.fac_bytes[p] = 0;
  for (int j = 1; j <= nL; ++j) {  // block ids and per-level arrays (levels may finish out of order)
    Level& lv = e.lev[j];
    const int64_t nb = lv.nboxes;
    const int64_t nblk = (int64_t)lv.ilidx.size();
    lv.bbase = e.nblk_all;
    e.nblk_all += nblk;
    lv.rank.assign(nblk, 0);
    lv.status.assign(nblk, 0);
    lv.col.assign(nblk, 0);
    lv.voff.assign(nblk, 0);
    lv.pivoff.assign(nblk, 0);
    lv.Rt.assign(nb, 0);
    lv.toff.assign(nb, 0);
    lv.chunk.assign(nb, -1);
    lv.uoff.assign(nb, 0);
  }
  AcaWorkspace ws[3];  // per pass index, freed when the assembly ends
  // The first pass of level 1 (a few huge blocks: one workgroup each would
  // hold a handful of CUs for seconds) runs on a second stream while the
  // other levels are assembled; level 1 is finished last.  Its factors land
  // in a chunk of their own, so the layout does not depend on the order.
  AcaWorkspace wsE;
  Pass early;
  hipStream_t side = nullptr;
  bool use_early = false;
  std::vector<int64_t> early_todo;
  if (nL >= 2 && batch_end(e, 1, 0, budget) == e.lev[1].nboxes) {
    early_todo = batch_blocks(e, 1, 0, e.lev[1].nboxes);
    int big = 0;
    for (int64_t b : early_todo)
      big = std::max<int>(big, e.bcount[level_offset(1) + e.lev[1].ilidx[b]]);
    use_early = !early_todo.empty() && big >= 8192;
  }
  try {
    // the first-pass workspace of every later batch is reserved now (no frees
    // while the early pass runs)
    {
      int64_t u = 0, v = 0, pw = 0;
      for (int j = use_early ? 2 : 1; j <= nL; ++j)
        for (int64_t k0 = 0; k0 < e.lev[j].nboxes;) {
          const int64_t k1 = batch_end(e, j, k0, budget);
          Pass P;
          fill_pass(e, j, batch_blocks(e, j, k0, k1), 0, P);
          int64_t wu = 0, wv = 0, wp = 0;
          for (auto& d : P.ds) {
            wu = std::max<int64_t>(wu, d.uw + (int64_t)d.rcap * d.n);
            wv = std::max<int64_t>(wv, d.vw + (int64_t)d.rcap * d.m);
            wp = std::max<int64_t>(wp, d.pw + 2 * (int64_t)d.rcap);
          }
          u = std::max(u, wu);
          v = std::max(v, wv);
          pw = std::max(pw, wp);
          ws[0].reserve_desc((int64_t)P.ds.size());
          AcaWorkspace::grow(ws[0].WD, ws[0].nWD, (int64_t)P.ds.size());
          k0 = k1;
        }
      ws[0].reserve(u, v, pw);
    }
    if (use_early) {  // launched after every allocation made ahead of time
      H3D_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      fill_pass(e, 1, early_todo, 0, early);
      launch_pass(e, early, wsE, side);
    }
    for (int jj = 1; jj <= nL; ++jj) {
      const int j = use_early ? (jj < nL ? jj + 1 : 1) : jj;
      for (int64_t k0 = 0; k0 < e.lev[j].nboxes;) {
        const int64_t k1 = batch_end(e, j, k0, budget);
        const bool first = use_early && j == 1 && k0 == 0;
        assemble_batch(e, j, k0, k1, ws, first ? &early : nullptr);
        k0 = k1;
      }
    }
  } catch (...) {
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    throw;
  }
  if (side) {
    H3D_HIP(hipStreamSynchronize(side));
    (void)hipStreamDestroy(side);
  }
  for (int j = 1; j <= nL; ++j) {  // t offsets in level order
    Level& lv = e.lev[j];
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      lv.toff[k] = e.T;
      if ((int64_t)e.T + lv.Rt[k] > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
      e.T += lv.Rt[k];
    }
  }
  finish_assembly(e);
  e.t_assemble_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Statistics, the per-block device table, blocks grouped by source box, the t
// index per block and the chunk tables, from the per-level factor bookkeeping
// (after assembly, or after loading a saved representation).
void finish_assembly(Engine& e) {
  hipStream_t st = e.stream;
  const int nL = e.nL;
  e.max_rank = 0;
  e.n_blocks = e.rank_sum = e.factor_elems = e.u_elems = e.v_elems = 0;
  e.find_memory = 0;
  for (int j = 1; j <= nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int64_t n = e.bcount[lo + k];
      for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
        const int64_t r = lv.rank[b], m = e.bcount[lo + lv.ilidx[b]];
        e.max_rank = std::max<int>(e.max_rank, (int)r);
        e.factor_elems += r * (n + m);
        e.u_elems += r * n;
        e.v_elems += r * m;
        e.rank_sum += r;
        e.find_memory += (double)r * n + (double)r * r + (double)m * r;
      }
    }
    e.n_blocks += (int64_t)lv.ilidx.size();
  }
  // per-block device table, blocks grouped by source box, t index per block
  {
    std::vector<BlockDev> blk(e.nblk_all);
    std::vector<int32_t> btoff(e.nblk_all);
    e.by_src.clear();
    e.by_src.reserve(e.nblk_all);
    for (int j = 1; j <= nL; ++j) {
      const Level& lv = e.lev[j];
      const int64_t lo = level_offset(j);
      std::vector<std::vector<int32_t>> bys(lv.nboxes);
      for (int64_t k = 0; k < lv.nboxes; ++k)
        for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
          const int32_t ki = lv.ilidx[b];
          BlockDev d;
          d.voff = lv.voff[b];
          d.col = lv.col[b];
          d.r = lv.rank[b];
          d.m = e.bcount[lo + ki];
          d.src = e.bstart[lo + ki];
          d.chunk = lv.chunk[k];
          d.pad = 0;
          blk[lv.bbase + b] = d;
          btoff[lv.bbase + b] = lv.toff[k] + lv.col[b];
          bys[ki].push_back((int32_t)(lv.bbase + b));
        }
      for (auto& v : bys) e.by_src.insert(e.by_src.end(), v.begin(), v.end());
    }
    dfree(e.d_blk);
    dfree(e.d_btoff);
    dfree(e.d_by_src);
    e.d_blk = dalloc<BlockDev>(std::max<int64_t>(e.nblk_all, 1));
    e.d_btoff = dalloc<int32_t>(std::max<int64_t>(e.nblk_all, 1));
    e.d_by_src = dalloc<int32_t>(std::max<int64_t>(e.nblk_all, 1));
    if (e.nblk_all) {
      H3D_HIP(hipMemcpyAsync(e.d_blk, blk.data(), sizeof(BlockDev) * e.nblk_all, hipMemcpyHostToDevice, st));
      H3D_HIP(hipMemcpyAsync(e.d_btoff, btoff.data(), sizeof(int32_t) * e.nblk_all, hipMemcpyHostToDevice, st));
      H3D_HIP(hipMemcpyAsync(e.d_by_src, e.by_src.data(), sizeof(int32_t) * e.nblk_all, hipMemcpyHostToDevice, st));
    }
    H3D_HIP(hipStreamSynchronize(st));
  }
  // near-field pair count and findMemory's dense part (src/HODLR3DTree.cpp:1770-1777)
  {
    const int64_t lo = level_offset(nL);
    const int64_t nleaf = e.lev[nL].nboxes;
    std::vector<int32_t> nbr(nleaf * 26);
    H3D_HIP(hipMemcpy(nbr.data(), e.lev[nL].d_nbr, sizeof(int32_t) * nleaf * 26,
                      hipMemcpyDeviceToHost));
    e.near_pairs = 0;
    for (int64_t k = 0; k < nleaf; ++k) {
      int64_t n = e.bcount[lo + k];
      int64_t s = n;
      for (int q = 0; q < 26; ++q)
        if (nbr[k * 26 + q] >= 0) s += e.bcount[lo + nbr[k * 26 + q]];
      e.near_pairs += n * s;
    }
    e.find_memory += (double)e.near_pairs;
  }
  upload_chunk_table(e);
  e.assembled = true;
}

// Kernel entries for h3d_get_matrix, in this translation unit so that they
// are formed without contraction, exactly as the ACA and the oracle form them.
namespace {

// One thread per entry of the column-major block: K(i, j) for rows i and
// columns j in the caller's order; i == j takes the diagonal value.
template <class KF>
__global__ __launch_bounds__(256) void entries_kernel(
    const int64_t* __restrict__ rows, int64_t nr, const int64_t* __restrict__ cols, int64_t n,
    int64_t base, const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ pz, const int32_t* __restrict__ iperm, KernelDev kd,
    const double* __restrict__ diag_vec, double* __restrict__ out) {
  const int64_t idx = base + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= n) return;
  const int64_t c = idx / nr, r = idx - c * nr;
  const int64_t i = rows[r], j = cols[c];
  double v;
  if (i == j) {
    v = diag_vec ? diag_vec[i] : kd.diag;
  } else {
    const int64_t a = iperm[i], b = iperm[j];
    v = kd.scale * kpair<KF>(kd.q, px[a], py[a], pz[a], px[b], py[b], pz[b]);
  }
  out[idx] = v;
}

}  // namespace

void get_entries(Engine& e, const int64_t* h_rows, int64_t nrows, const int64_t* h_cols,
                 int64_t ncols, double* d_out) {
  if (nrows <= 0 || ncols <= 0) return;
  for (int64_t r = 0; r < nrows; ++r)
    if (h_rows[r] < 0 || h_rows[r] >= e.N) throw Error(H3D_ERR_INVALID, "row out of range");
  for (int64_t c = 0; c < ncols; ++c)
    if (h_cols[c] < 0 || h_cols[c] >= e.N) throw Error(H3D_ERR_INVALID, "column out of range");
  hipStream_t st = e.stream;
  DevBuf<int64_t> d_rows(nrows), d_cols(ncols);
  H3D_HIP(hipMemcpyAsync(d_rows.p, h_rows, sizeof(int64_t) * nrows, hipMemcpyHostToDevice, st));
  H3D_HIP(hipMemcpyAsync(d_cols.p, h_cols, sizeof(int64_t) * ncols, hipMemcpyHostToDevice, st));
  const int64_t n = nrows * ncols;
  const int64_t per = int64_t(1) << 30;  // entries per launch (2^22 workgroups)
  for (int64_t base = 0; base < n; base += per) {
    const unsigned g = (unsigned)((std::min(per, n - base) + 255) / 256);
    const bool invr = e.kern.q == 7 || e.kern.q == 16;
    if (invr)
      entries_kernel<InvR><<<g, 256, 0, st>>>(d_rows.p, nrows, d_cols.p, n, base, e.d_px, e.d_py,
                                              e.d_pz, e.d_iperm, e.kern, e.d_diag, d_out);
    else if (e.kern.q == kQPlugin)
      entries_kernel<Plugin><<<g, 256, 0, st>>>(d_rows.p, nrows, d_cols.p, n, base, e.d_px,
                                                e.d_py, e.d_pz, e.d_iperm, e.kern, e.d_diag, d_out);
    else
      entries_kernel<AnyQ><<<g, 256, 0, st>>>(d_rows.p, nrows, d_cols.p, n, base, e.d_px, e.d_py,
                                              e.d_pz, e.d_iperm, e.kern, e.d_diag, d_out);
    H3D_LAUNCH_CHECK();
  }
  H3D_HIP(hipStreamSynchronize(st));
}

}  // namespace h3d
