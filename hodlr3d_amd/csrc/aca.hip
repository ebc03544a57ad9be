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
//   * reject when ||u|| or ||v|| < 1e-16 (ACA.cpp:196-200);
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
  double dx = xa - xb, dy = ya - yb, dz = za - zb;
  double R2 = dx * dx + dy * dy + dz * dz;
  return kd.scale * KF::f64(kd.q, R2);
}

__device__ inline bool bit_get(const uint32_t* m, int s) { return (m[s >> 5] >> (s & 31)) & 1u; }

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
  double* red = smem;           // 2 x 256
  double* amx = red + 512;      // 3 x (T/64)
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
        for (int l = 0; l < r; ++l) val -= uil[l] * V[(int64_t)l * m + s];
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
        for (int l = 0; l < r; ++l) col -= vjl[l] * U[(int64_t)l * n + s];
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
    for (int l = 0; l < r - 1; ++l) {
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

// Copies the first r vectors and pivots of each block into compact staging.
struct ExtractDesc {
  int64_t uw, vw;     // workspace offsets
  int64_t su, sv;     // staging offsets (doubles)
  int32_t n, m, r, pw, rcap, sp;  // sp: staging pivot offset (ints)
};

__global__ void extract_kernel(const ExtractDesc* __restrict__ ds, const double* __restrict__ Uws,
                               const double* __restrict__ Vws, const int32_t* __restrict__ pivws,
                               double* __restrict__ stU, double* __restrict__ stV,
                               int32_t* __restrict__ stP) {
  const ExtractDesc d = ds[blockIdx.x];
  const int64_t nu = (int64_t)d.r * d.n, nv = (int64_t)d.r * d.m;
  for (int64_t q = threadIdx.x; q < nu; q += blockDim.x) stU[d.su + q] = Uws[d.uw + q];
  for (int64_t q = threadIdx.x; q < nv; q += blockDim.x) stV[d.sv + q] = Vws[d.vw + q];
  for (int q = threadIdx.x; q < d.r; q += blockDim.x) {
    stP[d.sp + q] = pivws[d.pw + q];
    stP[d.sp + d.r + q] = pivws[d.pw + d.rcap + q];
  }
}

// Moves one block's vectors from staging into the per-level matvec layout.
struct LayoutDesc {
  const double* su;  // staging u_l (r x n)
  const double* sv;  // staging v_l (r x m)
  int64_t uo;        // U: element offset of the block's first column
  int64_t vo;        // V: element offset of row 0, first column of the block
  int32_t n, m, r, Rp;
  int32_t tdo;       // tdst offset of the block's first column
  int32_t tval;      // t index of the block's first rank
};

__global__ void layout_kernel(const LayoutDesc* __restrict__ ds, double* __restrict__ Ulev,
                              double* __restrict__ Vlev, int32_t* __restrict__ tdst) {
  const LayoutDesc d = ds[blockIdx.x];
  const int64_t nu = (int64_t)d.r * d.n;
  for (int64_t q = threadIdx.x; q < nu; q += blockDim.x) Ulev[d.uo + q] = d.su[q];
  const int64_t nv = (int64_t)d.r * d.m;
  for (int64_t q = threadIdx.x; q < nv; q += blockDim.x) {
    int l = (int)(q % d.r);
    int64_t s = q / d.r;
    Vlev[d.vo + s * d.Rp + l] = d.sv[(int64_t)l * d.m + s];
  }
  for (int l = threadIdx.x; l < d.r; l += blockDim.x) tdst[d.tdo + l] = d.tval + l;
}

size_t aca_smem(int rcap, int n, int m) {
  return sizeof(double) * (512 + 16 + 2 * (size_t)rcap) +
         sizeof(uint32_t) * (size_t)(((n + 31) >> 5) + ((m + 31) >> 5)) + 16;
}

struct Staged {
  double* su = nullptr;
  double* sv = nullptr;
  int32_t* sp = nullptr;  // I then J
};

}  // namespace

void assemble_all(Engine& e) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = e.stream;
  const int nL = e.nL;
  size_t freeb = 0, totb = 0;
  H3D_HIP(hipMemGetInfo(&freeb, &totb));
  int64_t budget = e.ws_budget > 0 ? e.ws_budget
                                   : std::min<int64_t>((int64_t)(freeb / 4), int64_t(32) << 30);
  budget = std::max<int64_t>(budget, int64_t(64) << 20);

  e.T = 0;
  e.nU = e.nV = 0;
  e.max_rank = 0;
  e.n_blocks = e.rank_sum = e.factor_elems = 0;
  std::vector<double*> Ulev(nL + 1, nullptr), Vlev(nL + 1, nullptr);
  std::vector<int64_t> nUlev(nL + 1, 0), nVlev(nL + 1, 0);
  std::vector<int32_t*> tdlev(nL + 1, nullptr);
  std::vector<int64_t> ntd(nL + 1, 0);

  for (int j = 1; j <= nL; ++j) {
    Level& lv = e.lev[j];
    const int64_t nb = lv.nboxes;
    const int64_t nblk = (int64_t)lv.ilidx.size();
    const int64_t lo = level_offset(j);
    lv.rank.assign(nblk, 0);
    lv.status.assign(nblk, 0);
    std::vector<int32_t> tgt(nblk);
    for (int64_t k = 0; k < nb; ++k)
      for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) tgt[b] = (int32_t)k;
    std::vector<Staged> stg(nblk);
    std::vector<std::pair<void*, int>> chunks;  // allocations to free after layout

    // pending blocks with their rank capacity
    std::vector<int64_t> todo;
    for (int64_t b = 0; b < nblk; ++b) {
      int n = e.bcount[lo + tgt[b]], m = e.bcount[lo + lv.ilidx[b]];
      if (n > 0 && m > 0) todo.push_back(b);
    }
    for (int pass = 0; pass < 3 && !todo.empty(); ++pass) {
      std::vector<int64_t> retry;
      for (int cls = 0; cls < 2; ++cls) {
        std::vector<int64_t> list;
        for (int64_t b : todo) {
          int n = e.bcount[lo + tgt[b]], m = e.bcount[lo + lv.ilidx[b]];
          int c = std::max(n, m) <= 256 ? 0 : 1;
          if (c == cls) list.push_back(b);
        }
        size_t pos = 0;
        while (pos < list.size()) {
          // gather a batch within the workspace budget
          std::vector<AcaDesc> ds;
          std::vector<int64_t> ids;
          int64_t wsU = 0, wsV = 0;
          int32_t pw = 0;
          size_t smax = 0;
          while (pos < list.size()) {
            int64_t b = list[pos];
            int32_t k = tgt[b], ki = lv.ilidx[b];
            int n = e.bcount[lo + k], m = e.bcount[lo + ki];
            int cap = pass == 0 ? 64 : (pass == 1 ? 512 : std::min(n, m));
            cap = std::min(cap, std::min(n, m));
            int64_t need = (int64_t)cap * (n + m) * 8;
            if (!ds.empty() && (wsU + wsV) * 8 + need > budget) break;
            if (ds.size() >= 1000000) break;
            AcaDesc d;
            d.rs = e.bstart[lo + k];
            d.n = n;
            d.cs = e.bstart[lo + ki];
            d.m = m;
            d.uw = wsU;
            d.vw = wsV;
            d.rcap = cap;
            d.pw = pw;
            wsU += (int64_t)cap * n;
            wsV += (int64_t)cap * m;
            pw += 2 * cap;
            smax = std::max(smax, aca_smem(cap, n, m));
            ds.push_back(d);
            ids.push_back(b);
            ++pos;
          }
          const int64_t nbat = (int64_t)ds.size();
          AcaDesc* d_ds = dalloc<AcaDesc>(nbat);
          double* d_U = dalloc<double>(wsU);
          double* d_V = dalloc<double>(wsV);
          int32_t* d_pw = dalloc<int32_t>(pw);
          int32_t* d_rank = dalloc<int32_t>(nbat);
          int32_t* d_stat = dalloc<int32_t>(nbat);
          H3D_HIP(hipMemcpyAsync(d_ds, ds.data(), sizeof(AcaDesc) * nbat, hipMemcpyHostToDevice, st));
          if (smax > 160 * 1024)
            throw Error(H3D_ERR_INVALID, "ACA block too large for LDS");
          const bool fast = e.kern.q == 7 || e.kern.q == 16;
          const void* kfun =
              cls == 0 ? (fast ? (const void*)aca_kernel<64, InvR> : (const void*)aca_kernel<64, AnyQ>)
                       : (fast ? (const void*)aca_kernel<256, InvR> : (const void*)aca_kernel<256, AnyQ>);
          if (smax > 64 * 1024)
            H3D_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smax));
          void* args[] = {&d_ds, &e.d_px, &e.d_py, &e.d_pz, &e.kern, &e.tol, &d_U, &d_V, &d_pw,
                          &d_rank, &d_stat};
          H3D_HIP(hipLaunchKernel(kfun, dim3((unsigned)nbat), dim3(cls == 0 ? 64 : 256), args, smax, st));
          H3D_LAUNCH_CHECK();
          std::vector<int32_t> rk(nbat), ss(nbat);
          H3D_HIP(hipMemcpyAsync(rk.data(), d_rank, sizeof(int32_t) * nbat, hipMemcpyDeviceToHost, st));
          H3D_HIP(hipMemcpyAsync(ss.data(), d_stat, sizeof(int32_t) * nbat, hipMemcpyDeviceToHost, st));
          H3D_HIP(hipStreamSynchronize(st));
          // stage accepted blocks
          std::vector<ExtractDesc> xs;
          int64_t su = 0, sv = 0;
          int32_t sp = 0;
          std::vector<std::array<int64_t, 3>> soff;
          for (int64_t q = 0; q < nbat; ++q) {
            int64_t b = ids[q];
            if (ss[q] & 2) {
              retry.push_back(b);
              continue;
            }
            if (ss[q] & 1) throw Error(H3D_ERR_NONFINITE, "non-finite kernel entry in ACA");
            lv.rank[b] = rk[q];
            lv.status[b] = ss[q];
            if (rk[q] == 0) continue;
            ExtractDesc x;
            x.uw = ds[q].uw;
            x.vw = ds[q].vw;
            x.su = su;
            x.sv = sv;
            x.n = ds[q].n;
            x.m = ds[q].m;
            x.r = rk[q];
            x.pw = ds[q].pw;
            x.rcap = ds[q].rcap;
            x.sp = sp;
            soff.push_back({su, sv, sp});
            su += (int64_t)x.r * x.n;
            sv += (int64_t)x.r * x.m;
            sp += 2 * x.r;
            xs.push_back(x);
          }
          if (!xs.empty()) {
            double* cu = dalloc<double>(su);
            double* cv = dalloc<double>(sv);
            int32_t* cp = dalloc<int32_t>(sp);
            chunks.push_back({cu, 0});
            chunks.push_back({cv, 0});
            chunks.push_back({cp, 0});
            ExtractDesc* d_x = dalloc<ExtractDesc>((int64_t)xs.size());
            H3D_HIP(hipMemcpyAsync(d_x, xs.data(), sizeof(ExtractDesc) * xs.size(),
                                   hipMemcpyHostToDevice, st));
            extract_kernel<<<(unsigned)xs.size(), 256, 0, st>>>(d_x, d_U, d_V, d_pw, cu, cv, cp);
            H3D_LAUNCH_CHECK();
            size_t qq = 0;
            for (int64_t q = 0; q < nbat; ++q) {
              int64_t b = ids[q];
              if ((ss[q] & 2) || rk[q] == 0) continue;
              stg[b].su = cu + soff[qq][0];
              stg[b].sv = cv + soff[qq][1];
              stg[b].sp = cp + soff[qq][2];
              ++qq;
            }
            H3D_HIP(hipStreamSynchronize(st));
            dfree(d_x);
          }
          dfree(d_ds);
          dfree(d_U);
          dfree(d_V);
          dfree(d_pw);
          dfree(d_rank);
          dfree(d_stat);
        }
      }
      todo.swap(retry);
    }
    if (!todo.empty()) throw Error(H3D_ERR_INTERNAL, "ACA did not finish within capacity");

    // ---- layout of this level
    lv.Rt.assign(nb, 0);
    lv.Rs.assign(nb, 0);
    lv.bcol_t.assign(nblk, 0);
    lv.bcol_s.assign(nblk, 0);
    lv.pivoff.assign(nblk + 1, 0);
    for (int64_t b = 0; b < nblk; ++b) {
      int r = lv.rank[b];
      lv.bcol_t[b] = lv.Rt[tgt[b]];
      lv.Rt[tgt[b]] += r;
      lv.bcol_s[b] = lv.Rs[lv.ilidx[b]];
      lv.Rs[lv.ilidx[b]] += r;
      lv.pivoff[b + 1] = lv.pivoff[b] + r;
      e.max_rank = std::max(e.max_rank, r);
      int n = e.bcount[lo + tgt[b]], m = e.bcount[lo + lv.ilidx[b]];
      e.factor_elems += (int64_t)r * (n + m);
      e.rank_sum += r;
      e.find_memory += (double)r * n + (double)r * r + (double)m * r;
    }
    e.n_blocks += nblk;
    for (int64_t k = 0; k < nb; ++k) lv.Rs[k] = (lv.Rs[k] + 3) & ~3;  // pad to 4 columns
    lv.uoff.assign(nb, 0);
    lv.voff.assign(nb, 0);
    lv.toff.assign(nb, 0);
    lv.tcol.assign(nb, 0);
    int64_t uo = 0, vo = 0;
    int64_t tc = 0;
    for (int64_t k = 0; k < nb; ++k) {
      int n = e.bcount[lo + k];
      lv.uoff[k] = uo;
      uo += (int64_t)n * lv.Rt[k];
      lv.voff[k] = vo;
      vo += (int64_t)n * lv.Rs[k];
      lv.toff[k] = e.T;
      e.T += lv.Rt[k];
      lv.tcol[k] = (int32_t)tc;
      tc += lv.Rs[k];
    }
    nUlev[j] = uo;
    nVlev[j] = vo;
    ntd[j] = tc;
    if (e.T < 0 || tc > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large (t index)");
    Ulev[j] = dalloc<double>(std::max<int64_t>(uo, 1));
    Vlev[j] = dalloc<double>(std::max<int64_t>(vo, 1));
    tdlev[j] = dalloc<int32_t>(std::max<int64_t>(tc, 1));
    H3D_HIP(hipMemsetAsync(Vlev[j], 0, sizeof(double) * std::max<int64_t>(vo, 1), st));
    H3D_HIP(hipMemsetAsync(tdlev[j], 0xff, sizeof(int32_t) * std::max<int64_t>(tc, 1), st));
    std::vector<LayoutDesc> ld;
    ld.reserve(nblk);
    std::vector<int32_t> hpI(lv.pivoff[nblk]), hpJ(lv.pivoff[nblk]);
    for (int64_t b = 0; b < nblk; ++b) {
      int r = lv.rank[b];
      if (r == 0) continue;
      int32_t k = tgt[b], ki = lv.ilidx[b];
      LayoutDesc x;
      x.su = stg[b].su;
      x.sv = stg[b].sv;
      x.n = e.bcount[lo + k];
      x.m = e.bcount[lo + ki];
      x.r = r;
      x.Rp = lv.Rs[ki];
      x.uo = lv.uoff[k] + (int64_t)lv.bcol_t[b] * x.n;
      x.vo = lv.voff[ki] + lv.bcol_s[b];
      x.tdo = lv.tcol[ki] + lv.bcol_s[b];
      x.tval = lv.toff[k] + lv.bcol_t[b];
      ld.push_back(x);
    }
    if (!ld.empty()) {
      LayoutDesc* d_ld = dalloc<LayoutDesc>((int64_t)ld.size());
      H3D_HIP(hipMemcpyAsync(d_ld, ld.data(), sizeof(LayoutDesc) * ld.size(), hipMemcpyHostToDevice, st));
      layout_kernel<<<(unsigned)ld.size(), 256, 0, st>>>(d_ld, Ulev[j], Vlev[j], tdlev[j]);
      H3D_LAUNCH_CHECK();
      // pivots: copy each block's staged pivots to the compact per-level arrays
      lv.d_pivI = dalloc<int32_t>(std::max<int64_t>(lv.pivoff[nblk], 1));
      lv.d_pivJ = dalloc<int32_t>(std::max<int64_t>(lv.pivoff[nblk], 1));
      for (int64_t b = 0; b < nblk; ++b) {
        int r = lv.rank[b];
        if (r == 0) continue;
        H3D_HIP(hipMemcpyAsync(lv.d_pivI + lv.pivoff[b], stg[b].sp, sizeof(int32_t) * r,
                               hipMemcpyDeviceToDevice, st));
        H3D_HIP(hipMemcpyAsync(lv.d_pivJ + lv.pivoff[b], stg[b].sp + r, sizeof(int32_t) * r,
                               hipMemcpyDeviceToDevice, st));
      }
      H3D_HIP(hipStreamSynchronize(st));
      dfree(d_ld);
    }
    for (auto& c : chunks) (void)hipFree(c.first);
  }
  // The factors stay in one allocation per level; tdst (one int per V
  // column) is merged into a single array with level offsets folded in.
  int64_t ntdall = 0;
  e.nU = e.nV = 0;
  for (int j = 1; j <= nL; ++j) {
    ntdall += ntd[j];
    e.nUlev[j] = nUlev[j];
    e.nVlev[j] = nVlev[j];
    e.nU += nUlev[j];
    e.nV += nVlev[j];
    e.fac[0].U[j] = Ulev[j];
    e.fac[0].V[j] = Vlev[j];
  }
  e.d_tdst = dalloc<int32_t>(std::max<int64_t>(ntdall, 1));
  int64_t ot = 0;
  for (int j = 1; j <= nL; ++j) {
    Level& lv = e.lev[j];
    if (ntd[j])
      H3D_HIP(hipMemcpyAsync(e.d_tdst + ot, tdlev[j], sizeof(int32_t) * ntd[j],
                             hipMemcpyDeviceToDevice, st));
    H3D_HIP(hipStreamSynchronize(st));
    dfree(tdlev[j]);
    for (auto& x : lv.tcol) x += (int32_t)ot;
    ot += ntd[j];
  }
  e.fac[0].bytes = (e.nU + e.nV) * 8;
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
  e.assembled = true;
  e.t_assemble_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace h3d
