// Fast matvec: batched low-rank applies streamed from HBM, the on-the-fly
// near field, and their transposes.
//
// y = A b is computed as (src/HODLR3DTree.cpp:1647-1805 restated for stored
// factors U = K(:,J) R^-1, V^T = L^-1 K(I,:), which equal the ACA vectors):
//   gather    x[s] = b[perm[s]]                               (assignCharges)
//   phase 1   t = V^T x, per source box, one wave per 64 columns
//   phase 2   y_j = U t, per target box, one wave per <= 64 rows
//   near      y_L += sum over self + face/edge leaves of K(i,j) x_j
//   combine   y[i] = sum_j y_j[iperm[i]]                (collect + reorder)
// Phase 1 lanes own columns and phase 2 lanes own rows, so neither needs a
// cross-lane reduction and every wave instruction reads a contiguous run of
// factors.  Each factor element is read exactly once per matvec.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "engine.hpp"
#include "kernels.hpp"

namespace h3d {

namespace {

constexpr int kP1Rows = 1024;  // rows of one phase-1 item before splitting
constexpr int kWavesPerBlock = 4;

struct LevPtr {
  const void* p[8];
};

template <class S>
struct Load;
template <>
struct Load<double> {
  __device__ static inline double f(const double* p) { return __builtin_nontemporal_load(p); }
};
template <>
struct Load<float> {
  __device__ static inline float f(const float* p) { return __builtin_nontemporal_load(p); }
};
template <>
struct Load<uint16_t> {
  __device__ static inline float f(const uint16_t* p) {
    return bf16_to_f(__builtin_nontemporal_load(p));
  }
};

template <class A>
__device__ inline A rl(A v, int l);
template <>
__device__ inline double rl<double>(double v, int l) { return readlane_d(v, l); }
template <>
__device__ inline float rl<float>(float v, int l) { return readlane_f(v, l); }

template <class A>
__device__ inline A wave_sum_xor(A v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__global__ void gather_kernel(const double* __restrict__ b, const int32_t* __restrict__ perm,
                              int64_t N, double* __restrict__ xp, float* __restrict__ xf) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  double v = b[perm[s]];
  xp[s] = v;
  if (xf) xf[s] = (float)v;
}

// ---------------------------------------------------------------- phase 1
template <class S, class A>
__global__ __launch_bounds__(256) void phase1_kernel(const P1Item* __restrict__ items, int nitems,
                                                     LevPtr Vl, const A* __restrict__ x,
                                                     const int32_t* __restrict__ tdst,
                                                     A* __restrict__ t, A* __restrict__ part) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const P1Item it = items[w];
  const S* V = (const S*)Vl.p[it.level] + it.voff;
  const bool active = lane < it.nc;
  const int c = it.c0 + (active ? lane : 0);
  const int R = it.R;
  A acc = 0;
  for (int s = it.s0; s < it.s1; s += 64) {
    const int cnt = min(64, it.s1 - s);
    const A xv = (lane < cnt) ? x[it.xoff + s + lane] : (A)0;
    const S* row = V + (int64_t)s * R + c;
    int u = 0;
    for (; u + 8 <= cnt; u += 8) {
      A v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (A)Load<S>::f(row + (int64_t)(u + q) * R);
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += v[q] * rl<A>(xv, u + q);
    }
    for (; u < cnt; ++u) acc += (A)Load<S>::f(row + (int64_t)u * R) * rl<A>(xv, u);
  }
  if (!active) return;
  if (it.pslot < 0) {
    const int dst = tdst[it.tcol + c];
    if (dst >= 0) t[dst] = acc;
  } else {
    part[(int64_t)it.pslot * 64 + lane] = acc;
  }
}

// Sums the row-split partials of a column chunk in row order.  direct != 0:
// the destination is tcol + c0 + lane; otherwise tdst[tcol + c0 + lane].
template <class A>
__global__ void reduce_kernel(const P1Reduce* __restrict__ rs, int n, const A* __restrict__ part,
                              const int32_t* __restrict__ tdst, A* __restrict__ t, int direct) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= n) return;
  const int lane = threadIdx.x & 63;
  const P1Reduce r = rs[w];
  if (lane >= r.nc) return;
  A acc = 0;
  for (int q = 0; q < r.nsplit; ++q) acc += part[(int64_t)(r.pbase + q) * 64 + lane];
  const int idx = r.tcol + r.c0 + lane;
  const int dst = direct ? idx : tdst[idx];
  if (dst >= 0) t[dst] = acc;
}

// ---------------------------------------------------------------- phase 2
template <class S, class A>
__global__ __launch_bounds__(256) void phase2_kernel(const P2Item* __restrict__ items, int nitems,
                                                     LevPtr Ul, const A* __restrict__ t,
                                                     double* __restrict__ ylev, int64_t N) {
  __shared__ A red[kWavesPerBlock][64];
  const int wv = threadIdx.x >> 6;
  const int w = blockIdx.x * kWavesPerBlock + wv;
  const int lane = threadIdx.x & 63;
  const bool valid = w < nitems;
  P2Item it;
  if (valid) it = items[w];
  int G = 1, RB = 1;
  A acc = 0;
  if (valid) {
    RB = it.nr;
    G = 64 / RB;
    const int row = lane % RB, g = lane / RB;
    if (g < G) {
      const S* U = (const S*)Ul.p[it.level] + it.uoff + it.i0 + row;
      const A* tt = t + it.toff;
      const int64_t ld = it.ld;
      int c = g;
      for (; c + 3 * G < it.R; c += 4 * G) {
        A v0 = (A)Load<S>::f(U + (int64_t)c * ld);
        A v1 = (A)Load<S>::f(U + (int64_t)(c + G) * ld);
        A v2 = (A)Load<S>::f(U + (int64_t)(c + 2 * G) * ld);
        A v3 = (A)Load<S>::f(U + (int64_t)(c + 3 * G) * ld);
        acc += v0 * tt[c];
        acc += v1 * tt[c + G];
        acc += v2 * tt[c + 2 * G];
        acc += v3 * tt[c + 3 * G];
      }
      for (; c < it.R; c += G) acc += (A)Load<S>::f(U + (int64_t)c * ld) * tt[c];
    }
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (!valid) return;
  if (lane < RB) {
    A s = 0;
    for (int g = 0; g < G; ++g) s += red[wv][g * RB + lane];
    ylev[(int64_t)(it.level - 1) * N + it.yoff + it.i0 + lane] = (double)s;
  }
}

// --------------------------------------------------------- transpose phases
// Transpose phase 1 (items reuse P1Item: voff = U offset of the target box,
// R = its leading dimension, xoff = its first position, tcol = its t offset):
// t[toff + c] = sum_i U[c*ld + i] x[i], lanes over rows, one wave reduction
// per column.
template <class S, class A>
__global__ __launch_bounds__(256) void tphase1_kernel(const P1Item* __restrict__ items, int nitems,
                                                      LevPtr Ul, const A* __restrict__ x,
                                                      A* __restrict__ t, A* __restrict__ part) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const P1Item it = items[w];
  const S* U = (const S*)Ul.p[it.level] + it.voff;
  const int64_t ld = it.R;
  A mine = 0;
  for (int s = it.s0; s < it.s1; s += 64) {
    const bool rv = s + lane < it.s1;
    const A xv = rv ? x[it.xoff + s + lane] : (A)0;
    const int rr = rv ? s + lane : it.s0;
    for (int cc = 0; cc < it.nc; ++cc) {
      A p = (A)Load<S>::f(U + (int64_t)(it.c0 + cc) * ld + rr) * xv;
      p = wave_sum_xor(p);
      if (lane == cc) mine += p;
    }
  }
  if (lane >= it.nc) return;
  if (it.pslot < 0)
    t[it.tcol + it.c0 + lane] = mine;
  else
    part[(int64_t)it.pslot * 64 + lane] = mine;
}

// Transpose phase 2 (items reuse P2Item: uoff = V offset of the source box,
// ld = R = its padded column count, toff = its tdst offset): y[s] = sum_c
// V[s*Rp + c] t[tdst[c]], lanes over columns, one wave reduction per row.
template <class S, class A>
__global__ __launch_bounds__(256) void tphase2_kernel(const P2Item* __restrict__ items, int nitems,
                                                      LevPtr Vl, const A* __restrict__ t,
                                                      const int32_t* __restrict__ tdst,
                                                      double* __restrict__ ylev, int64_t N) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const P2Item it = items[w];
  const S* V = (const S*)Vl.p[it.level] + it.uoff;
  const int Rp = it.ld;
  A mine = 0;
  for (int c0 = 0; c0 < it.R; c0 += 64) {
    const int c = c0 + lane;
    A tv = 0;
    if (c < it.R) {
      const int d = tdst[it.toff + c];
      tv = d >= 0 ? t[d] : (A)0;
    }
    const int cc = c < it.R ? c : c0;
    for (int r = 0; r < it.nr; ++r) {
      A p = (A)Load<S>::f(V + (int64_t)(it.i0 + r) * Rp + cc) * tv;
      p = wave_sum_xor(p);
      if (lane == r) mine += p;
    }
  }
  if (lane < it.nr) ylev[(int64_t)(it.level - 1) * N + it.yoff + it.i0 + lane] = (double)mine;
}

// ---------------------------------------------------------------- near field
// fp64: exact entries (the reference's arithmetic).  The i == j term uses the
// diagonal (reference: 0).
template <class KF>
__global__ __launch_bounds__(256) void near64_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const int32_t* __restrict__ perm, KernelDev kd, const double* __restrict__ x,
    double* __restrict__ yL) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  const bool act = lane < it.nr;
  const int ip = rs + it.i0 + (act ? lane : 0);
  const double xi = px[ip], yi = py[ip], zi = pz[ip];
  double acc = 0.0;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    for (int s0 = 0; s0 < m; s0 += 64) {
      const int cnt = min(64, m - s0);
      const int jp = cs + s0 + (lane < cnt ? lane : 0);
      const double sx = px[jp], sy = py[jp], sz = pz[jp], sq = x[jp];
      for (int u = 0; u < cnt; ++u) {
        const double X = readlane_d(sx, u), Y = readlane_d(sy, u), Z = readlane_d(sz, u),
                     Q = readlane_d(sq, u);
        const int jpos = cs + s0 + u;
        if (jpos == ip) continue;
        double dx = xi - X, dy = yi - Y, dz = zi - Z;
        double R2 = dx * dx + dy * dy + dz * dz;
        acc += kd.scale * KF::f64(kd.q, R2) * Q;
      }
    }
  }
  if (!act) return;
  const double dg = kd.diag_vec ? kd.diag_vec[perm[ip]] : kd.diag;
  acc += dg * x[ip];
  yL[ip] += acc;
}

// fp32: coordinates as fp32 offsets from the leaf centre plus the fp32
// difference of the two leaf centres, so near pairs keep ~1e-7 relative
// accuracy in r.
__global__ __launch_bounds__(256) void near32_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const float* __restrict__ lx, const float* __restrict__ ly, const float* __restrict__ lz,
    const double* __restrict__ leafc, const int32_t* __restrict__ perm, KernelDev kd,
    const float* __restrict__ x, double* __restrict__ yL) {
  const int w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  const bool act = lane < it.nr;
  const int ip = rs + it.i0 + (act ? lane : 0);
  const float xi = lx[ip], yi = ly[ip], zi = lz[ip];
  const double cxk = leafc[3 * (int64_t)it.leaf], cyk = leafc[3 * (int64_t)it.leaf + 1],
               czk = leafc[3 * (int64_t)it.leaf + 2];
  const float sc = (float)kd.scale;
  float acc = 0.0f;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    const float dcx = (float)(cxk - leafc[3 * (int64_t)nn]);
    const float dcy = (float)(cyk - leafc[3 * (int64_t)nn + 1]);
    const float dcz = (float)(czk - leafc[3 * (int64_t)nn + 2]);
    const float ax = xi + dcx, ay = yi + dcy, az = zi + dcz;
    float part = 0.0f;
    for (int s0 = 0; s0 < m; s0 += 64) {
      const int cnt = min(64, m - s0);
      const int jp = cs + s0 + (lane < cnt ? lane : 0);
      const float sx = lx[jp], sy = ly[jp], sz = lz[jp], sq = x[jp];
      for (int u = 0; u < cnt; ++u) {
        const float X = readlane_f(sx, u), Y = readlane_f(sy, u), Z = readlane_f(sz, u),
                    Q = readlane_f(sq, u);
        float dx = ax - X, dy = ay - Y, dz = az - Z;
        float R2 = dx * dx + dy * dy + dz * dz;
        float v = (cs + s0 + u == ip) ? 0.0f : phi32(kd.q, R2);
        part += v * Q;
      }
    }
    acc += part;
  }
  if (!act) return;
  const double dg = kd.diag_vec ? kd.diag_vec[perm[ip]] : kd.diag;
  yL[ip] += (double)(sc * acc) + dg * (double)x[ip];
}

__global__ void combine_kernel(const double* __restrict__ ylev, int nL, int64_t N,
                               const int32_t* __restrict__ iperm, double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t s = iperm[i];
  double acc = 0.0;
  for (int j = 0; j < nL; ++j) acc = acc + ylev[(int64_t)j * N + s];
  y[i] = acc;
}

template <class S>
__global__ void convert_kernel(const double* __restrict__ src, int64_t n, S* __restrict__ dst);
template <>
__global__ void convert_kernel<float>(const double* __restrict__ src, int64_t n,
                                      float* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}
template <>
__global__ void convert_kernel<uint16_t>(const double* __restrict__ src, int64_t n,
                                         uint16_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = f_to_bf16((float)src[i]);
}

inline unsigned wave_grid(int64_t items) {
  return (unsigned)((items + kWavesPerBlock - 1) / kWavesPerBlock);
}

template <class T>
T* upload(const std::vector<T>& v, hipStream_t st) {
  T* d = dalloc<T>(std::max<int64_t>((int64_t)v.size(), 1));
  if (!v.empty())
    H3D_HIP(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
  return d;
}

LevPtr levptr(void* const* a) {
  LevPtr p{};
  for (int j = 0; j < 8; ++j) p.p[j] = a[j];
  return p;
}

}  // namespace

void build_plan(Engine& e) {
  const int nL = e.nL;
  hipStream_t st = e.stream;
  e.p1.clear();
  e.p1r.clear();
  e.p2.clear();
  e.nitems.clear();
  e.tp2.clear();
  int32_t pslot = 0;
  std::vector<P1Item> tp1;
  std::vector<P1Reduce> tp1r;
  int32_t tslot = 0;
  for (int j = 1; j <= nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      if (n == 0) continue;
      // forward phase 1 over source box k
      if (lv.Rs[k] > 0) {
        const int nsplit = (n + kP1Rows - 1) / kP1Rows;
        for (int c0 = 0; c0 < lv.Rs[k]; c0 += 64) {
          const int nc = std::min(64, lv.Rs[k] - c0);
          if (nsplit > 1) e.p1r.push_back({pslot, nsplit, lv.tcol[k], c0, nc, 0});
          for (int q = 0; q < nsplit; ++q) {
            P1Item it{};
            it.voff = lv.voff[k];
            it.level = j;
            it.R = lv.Rs[k];
            it.c0 = c0;
            it.nc = nc;
            it.xoff = e.bstart[lo + k];
            it.s0 = q * kP1Rows;
            it.s1 = std::min(n, (q + 1) * kP1Rows);
            it.tcol = lv.tcol[k];
            it.pslot = nsplit > 1 ? pslot++ : -1;
            e.p1.push_back(it);
          }
          // transpose phase 2 over source box k
        }
        for (int i0 = 0; i0 < n; i0 += 64) {
          P2Item it{};
          it.uoff = lv.voff[k];
          it.ld = lv.Rs[k];
          it.i0 = i0;
          it.nr = std::min(64, n - i0);
          it.R = lv.Rs[k];
          it.toff = lv.tcol[k];
          it.yoff = e.bstart[lo + k];
          it.level = j;
          e.tp2.push_back(it);
        }
      }
      // forward phase 2 and transpose phase 1 over target box k
      if (lv.Rt[k] > 0) {
        for (int i0 = 0; i0 < n; i0 += 64) {
          P2Item it{};
          it.uoff = lv.uoff[k];
          it.ld = n;
          it.i0 = i0;
          it.nr = std::min(64, n - i0);
          it.R = lv.Rt[k];
          it.toff = lv.toff[k];
          it.yoff = e.bstart[lo + k];
          it.level = j;
          e.p2.push_back(it);
        }
        const int nsplit = (n + kP1Rows - 1) / kP1Rows;
        for (int c0 = 0; c0 < lv.Rt[k]; c0 += 64) {
          const int nc = std::min(64, lv.Rt[k] - c0);
          if (nsplit > 1) tp1r.push_back({tslot, nsplit, lv.toff[k], c0, nc, 0});
          for (int q = 0; q < nsplit; ++q) {
            P1Item it{};
            it.voff = lv.uoff[k];
            it.level = j;
            it.R = n;
            it.c0 = c0;
            it.nc = nc;
            it.xoff = e.bstart[lo + k];
            it.s0 = q * kP1Rows;
            it.s1 = std::min(n, (q + 1) * kP1Rows);
            it.tcol = lv.toff[k];
            it.pslot = nsplit > 1 ? tslot++ : -1;
            tp1.push_back(it);
          }
        }
      }
    }
  }
  {
    const int64_t lo = level_offset(nL);
    for (int64_t k = 0; k < e.lev[nL].nboxes; ++k) {
      const int n = e.bcount[lo + k];
      for (int i0 = 0; i0 < n; i0 += 64) e.nitems.push_back({(int32_t)k, i0, std::min(64, n - i0), 0});
    }
  }
  // Larger items first: less tail at the end of the launch.
  std::stable_sort(e.p2.begin(), e.p2.end(), [](const P2Item& a, const P2Item& b) {
    return (int64_t)a.R * a.nr > (int64_t)b.R * b.nr;
  });
  e.npartial = std::max(pslot, tslot);
  dfree(e.d_p1);
  dfree(e.d_p1r);
  dfree(e.d_p2);
  dfree(e.d_near);
  dfree(e.d_tp1x);
  dfree(e.d_tp1r);
  dfree(e.d_tp2);
  e.d_p1 = upload(e.p1, st);
  e.d_p1r = upload(e.p1r, st);
  e.d_p2 = upload(e.p2, st);
  e.d_near = upload(e.nitems, st);
  e.tp1_items = tp1;
  e.tp1_red = tp1r;
  e.d_tp1x = upload(tp1, st);
  e.d_tp1r = upload(tp1r, st);
  e.d_tp2 = upload(e.tp2, st);
  H3D_HIP(hipStreamSynchronize(st));
}

void ensure_scratch(Engine& e, int nrhs) {
  if (e.max_nrhs >= nrhs && e.d_xp) return;
  dfree(e.d_xp);
  dfree(e.d_t);
  dfree(e.d_part);
  dfree(e.d_ylev);
  dfree(e.d_io);
  dfree(e.d_xf);
  e.d_xp = dalloc<double>(e.N * nrhs);
  e.d_xf = dalloc<float>(e.N * nrhs);
  e.d_t = dalloc<double>(std::max<int64_t>((int64_t)e.T * nrhs, 1));
  e.d_part = dalloc<double>(std::max<int64_t>((int64_t)e.npartial * 64 * nrhs, 1));
  e.d_ylev = dalloc<double>(e.N * e.nL * nrhs);
  e.d_io = dalloc<double>(2 * e.N * nrhs);
  e.max_nrhs = nrhs;
}

void prepare_precision(Engine& e, int prec) {
  if (prec == H3D_PREC_FP64 || e.fac[prec].bytes > 0) return;
  hipStream_t st = e.stream;
  int64_t bytes = 0;
  for (int j = 1; j <= e.nL; ++j) {
    for (int side = 0; side < 2; ++side) {
      const int64_t n = side == 0 ? e.nUlev[j] : e.nVlev[j];
      const double* src = (const double*)(side == 0 ? e.fac[0].U[j] : e.fac[0].V[j]);
      void* dst = nullptr;
      if (n > 0) {
        if (prec == H3D_PREC_FP32) {
          float* d = dalloc<float>(n);
          convert_kernel<float><<<4096, 256, 0, st>>>(src, n, d);
          dst = d;
          bytes += n * 4;
        } else {
          uint16_t* d = dalloc<uint16_t>(n);
          convert_kernel<uint16_t><<<4096, 256, 0, st>>>(src, n, d);
          dst = d;
          bytes += n * 2;
        }
        H3D_LAUNCH_CHECK();
      }
      (side == 0 ? e.fac[prec].U[j] : e.fac[prec].V[j]) = dst;
    }
  }
  H3D_HIP(hipStreamSynchronize(st));
  e.fac[prec].bytes = std::max<int64_t>(bytes, 1);
}

namespace {

template <class S, class A>
void run_forward(Engine& e, const A* x, A* t, A* part, int nitems1, int nitems2) {
  hipStream_t st = e.stream;
  int prec = std::is_same<S, double>::value ? 0 : (std::is_same<S, float>::value ? 1 : 2);
  LevPtr Vl = levptr(e.fac[prec].V), Ul = levptr(e.fac[prec].U);
  if (nitems1)
    phase1_kernel<S, A><<<wave_grid(nitems1), 256, 0, st>>>(e.d_p1, nitems1, Vl, x, e.d_tdst, t,
                                                            part);
  H3D_LAUNCH_CHECK();
  if (!e.p1r.empty())
    reduce_kernel<A><<<wave_grid(e.p1r.size()), 256, 0, st>>>(e.d_p1r, (int)e.p1r.size(), part,
                                                              e.d_tdst, t, 0);
  H3D_LAUNCH_CHECK();
  if (nitems2)
    phase2_kernel<S, A><<<wave_grid(nitems2), 256, 0, st>>>(e.d_p2, nitems2, Ul, t, e.d_ylev, e.N);
  H3D_LAUNCH_CHECK();
}

template <class S, class A>
void run_transpose(Engine& e, const A* x, A* t, A* part) {
  hipStream_t st = e.stream;
  int prec = std::is_same<S, double>::value ? 0 : (std::is_same<S, float>::value ? 1 : 2);
  LevPtr Vl = levptr(e.fac[prec].V), Ul = levptr(e.fac[prec].U);
  const int n1 = (int)e.tp1_items.size(), n2 = (int)e.tp2.size();
  if (n1)
    tphase1_kernel<S, A><<<wave_grid(n1), 256, 0, st>>>(e.d_tp1x, n1, Ul, x, t, part);
  H3D_LAUNCH_CHECK();
  if (!e.tp1_red.empty())
    reduce_kernel<A><<<wave_grid(e.tp1_red.size()), 256, 0, st>>>(
        e.d_tp1r, (int)e.tp1_red.size(), part, e.d_tdst, t, 1);
  H3D_LAUNCH_CHECK();
  if (n2)
    tphase2_kernel<S, A><<<wave_grid(n2), 256, 0, st>>>(e.d_tp2, n2, Vl, t, e.d_tdst, e.d_ylev, e.N);
  H3D_LAUNCH_CHECK();
}

}  // namespace

void matvec_device(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool transpose) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "matvec before assemble");
  if (prec < 0 || prec > 2) throw Error(H3D_ERR_INVALID, "unknown precision");
  if (prec != H3D_PREC_FP64 && e.fac[prec].bytes == 0) prepare_precision(e, prec);
  ensure_scratch(e, 1);
  hipStream_t st = e.stream;
  const int64_t N = e.N;
  const unsigned gN = (unsigned)((N + 255) / 256);
  for (int r = 0; r < nrhs; ++r) {
    const double* b = d_b + (int64_t)r * N;
    double* y = d_y + (int64_t)r * N;
    gather_kernel<<<gN, 256, 0, st>>>(b, e.d_perm, N, e.d_xp, prec == 0 ? nullptr : e.d_xf);
    H3D_LAUNCH_CHECK();
    H3D_HIP(hipMemsetAsync(e.d_ylev, 0, sizeof(double) * N * e.nL, st));
    const int n1 = (int)e.p1.size(), n2 = (int)e.p2.size();
    if (prec == H3D_PREC_FP64) {
      if (!transpose)
        run_forward<double, double>(e, e.d_xp, e.d_t, e.d_part, n1, n2);
      else
        run_transpose<double, double>(e, e.d_xp, e.d_t, e.d_part);
    } else if (prec == H3D_PREC_FP32) {
      float* t = (float*)e.d_t;
      float* part = (float*)e.d_part;
      if (!transpose)
        run_forward<float, float>(e, e.d_xf, t, part, n1, n2);
      else
        run_transpose<float, float>(e, e.d_xf, t, part);
    } else {
      float* t = (float*)e.d_t;
      float* part = (float*)e.d_part;
      if (!transpose)
        run_forward<uint16_t, float>(e, e.d_xf, t, part, n1, n2);
      else
        run_transpose<uint16_t, float>(e, e.d_xf, t, part);
    }
    // near field into the leaf level (symmetric kernels: the same for A^T)
    const int nn = (int)e.nitems.size();
    double* yL = e.d_ylev + (int64_t)(e.nL - 1) * N;
    const int64_t lo = level_offset(e.nL);
    KernelDev kd = e.kern;
    kd.diag_vec = e.d_diag;
    if (nn) {
      if (prec == H3D_PREC_FP64 && (kd.q == 7 || kd.q == 16))
        near64_kernel<InvR><<<wave_grid(nn), 256, 0, st>>>(e.d_near, nn, e.lev[e.nL].d_nbr,
                                                           e.d_bstart + lo, e.d_bcount + lo, e.d_px,
                                                           e.d_py, e.d_pz, e.d_perm, kd, e.d_xp, yL);
      else if (prec == H3D_PREC_FP64)
        near64_kernel<AnyQ><<<wave_grid(nn), 256, 0, st>>>(e.d_near, nn, e.lev[e.nL].d_nbr,
                                                           e.d_bstart + lo, e.d_bcount + lo, e.d_px,
                                                           e.d_py, e.d_pz, e.d_perm, kd, e.d_xp, yL);
      else
        near32_kernel<<<wave_grid(nn), 256, 0, st>>>(e.d_near, nn, e.lev[e.nL].d_nbr,
                                                     e.d_bstart + lo, e.d_bcount + lo, e.d_lx,
                                                     e.d_ly, e.d_lz, e.d_leafc, e.d_perm, kd,
                                                     e.d_xf, yL);
      H3D_LAUNCH_CHECK();
    }
    combine_kernel<<<gN, 256, 0, st>>>(e.d_ylev, e.nL, N, e.d_iperm, y);
    H3D_LAUNCH_CHECK();
  }
}

}  // namespace h3d
