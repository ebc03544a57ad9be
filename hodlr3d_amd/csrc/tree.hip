// Octree construction on the GPU: point bucketing, leaf-DFS permutation,
// neighbour slots and interaction lists.
//
// Reference behaviour followed:
//   box k at level j has children 8k+c, centres parent +- boxRadius[j]/2 with
//   the sign table of src/HODLR3DTree.cpp:1518-1556;
//   a point goes to the child selected by (z <= cz), (x <= cx), (y <= cy)
//   against the parent centre (src/HODLR3DTree.cpp:1558-1608), which keeps
//   each box's points in ascending original index; non-leaf boxes list their
//   children's points in child order (:1610-1619), so a stable sort of the
//   points by leaf number gives every box as one contiguous segment;
//   neighbour slot s <-> offset (m%3-1, (m%9)/3-1, m/9-1), m = s<13 ? s : s+1,
//   holds face and edge neighbours only, and the interaction list is the
//   corner sibling followed, for each parent neighbour slot in ascending
//   order, by that neighbour's corner-adjacent child and then its
//   non-adjacent children in child order (src/HODLR3DTree.cpp:110-1489).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "engine.hpp"
#include "kernels.hpp"

namespace h3d {

namespace {

struct Shifts {
  double s[8];
};

__host__ __device__ inline int cxb(int c) { return (c == 1 || c == 2 || c == 5 || c == 6); }
__host__ __device__ inline int cyb(int c) { return (c == 2 || c == 3 || c == 6 || c == 7); }
__host__ __device__ inline int czb(int c) { return c >= 4; }
__host__ __device__ inline int child_of(int xb, int yb, int zb) {
  return zb * 4 + (yb ? (xb ? 2 : 3) : (xb ? 1 : 0));
}

__global__ void leaf_key_kernel(const double* __restrict__ xyz, int64_t N, int nL, double c0x,
                                double c0y, double c0z, Shifts sh, uint32_t* __restrict__ key,
                                int32_t* __restrict__ val) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  double cx = c0x, cy = c0y, cz = c0z;
  uint32_t box = 0;
  for (int j = 0; j < nL; ++j) {
    int zb = (z <= cz) ? 0 : 1;
    int xb = (x <= cx) ? 0 : 1;
    int yb = (y <= cy) ? 0 : 1;
    box = 8 * box + child_of(xb, yb, zb);
    double s = sh.s[j];
    cx = xb ? cx + s : cx - s;
    cy = yb ? cy + s : cy - s;
    cz = zb ? cz + s : cz - s;
  }
  key[i] = box;
  val[i] = (int32_t)i;
}

__global__ void count_kernel(const uint32_t* __restrict__ key, int64_t N, int32_t* __restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) atomicAdd(&cnt[key[i]], 1);
}

__global__ void gather_points_kernel(const double* __restrict__ xyz, const int32_t* __restrict__ perm,
                                     const uint32_t* __restrict__ leafkey,
                                     const double* __restrict__ leafc, int64_t N,
                                     double* __restrict__ px, double* __restrict__ py,
                                     double* __restrict__ pz, float* __restrict__ lx,
                                     float* __restrict__ ly, float* __restrict__ lz,
                                     int32_t* __restrict__ iperm) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int32_t o = perm[s];
  double x = xyz[3 * (int64_t)o], y = xyz[3 * (int64_t)o + 1], z = xyz[3 * (int64_t)o + 2];
  px[s] = x;
  py[s] = y;
  pz[s] = z;
  uint32_t k = leafkey[s];
  lx[s] = (float)(x - leafc[3 * (int64_t)k]);
  ly[s] = (float)(y - leafc[3 * (int64_t)k + 1]);
  lz[s] = (float)(z - leafc[3 * (int64_t)k + 2]);
  iperm[o] = (int32_t)s;
}

__device__ inline void box_coords(int64_t k, int j, int& X, int& Y, int& Z) {
  X = Y = Z = 0;
  for (int t = 0; t < j; ++t) {
    int c = (int)((k >> (3 * (j - 1 - t))) & 7);
    X = 2 * X + cxb(c);
    Y = 2 * Y + cyb(c);
    Z = 2 * Z + czb(c);
  }
}
__device__ inline int64_t box_number(int X, int Y, int Z, int j) {
  int64_t k = 0;
  for (int t = j - 1; t >= 0; --t) k = 8 * k + child_of((X >> t) & 1, (Y >> t) & 1, (Z >> t) & 1);
  return k;
}
__device__ inline void slot_off(int s, int& dx, int& dy, int& dz) {
  int m = s < 13 ? s : s + 1;
  dx = m % 3 - 1;
  dy = (m % 9) / 3 - 1;
  dz = m / 9 - 1;
}

// mode 0: write counts; mode 1: fill lists at off[k].
__global__ void il_kernel(int j, int64_t nb, int mode, int32_t* __restrict__ cnt,
                          const int32_t* __restrict__ off, int32_t* __restrict__ out,
                          int32_t* __restrict__ nbr) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  int side = 1 << j;
  int X, Y, Z;
  box_coords(k, j, X, Y, Z);
  int c = (int)(k & 7);
  int64_t p = k >> 3;
  int n = 0;
  int32_t* dst = mode ? out + off[k] : nullptr;
  auto push = [&](int64_t b) {
    if (mode) dst[n] = (int32_t)b;
    ++n;
  };
  push(8 * p + child_of(1 - cxb(c), 1 - cyb(c), 1 - czb(c)));
  if (j >= 2) {
    int PX = X >> 1, PY = Y >> 1, PZ = Z >> 1, ps = side >> 1;
    for (int s = 0; s < 26; ++s) {
      int dx, dy, dz;
      slot_off(s, dx, dy, dz);
      if (abs(dx) + abs(dy) + abs(dz) > 2) continue;
      int QX = PX + dx, QY = PY + dy, QZ = PZ + dz;
      if (QX < 0 || QY < 0 || QZ < 0 || QX >= ps || QY >= ps || QZ >= ps) continue;
      int64_t corner = -1;
      for (int cc = 0; cc < 8; ++cc) {
        int ax = abs(2 * QX + cxb(cc) - X), ay = abs(2 * QY + cyb(cc) - Y), az = abs(2 * QZ + czb(cc) - Z);
        if (max(ax, max(ay, az)) == 1 && ax + ay + az == 3)
          corner = box_number(2 * QX + cxb(cc), 2 * QY + cyb(cc), 2 * QZ + czb(cc), j);
      }
      if (corner >= 0) push(corner);
      for (int cc = 0; cc < 8; ++cc) {
        int CX = 2 * QX + cxb(cc), CY = 2 * QY + cyb(cc), CZ = 2 * QZ + czb(cc);
        int ax = abs(CX - X), ay = abs(CY - Y), az = abs(CZ - Z);
        if (max(ax, max(ay, az)) >= 2) push(box_number(CX, CY, CZ, j));
      }
    }
  }
  if (mode == 0) {
    cnt[k] = n;
    for (int s = 0; s < 26; ++s) {
      int dx, dy, dz;
      slot_off(s, dx, dy, dz);
      int32_t b = -1;
      int AX = X + dx, AY = Y + dy, AZ = Z + dz;
      if (abs(dx) + abs(dy) + abs(dz) <= 2 && AX >= 0 && AY >= 0 && AZ >= 0 && AX < side &&
          AY < side && AZ < side)
        b = (int32_t)box_number(AX, AY, AZ, j);
      nbr[k * 26 + s] = b;
    }
  }
}

// Flags leaves holding two points with identical coordinates (their entry
// 1/r would be infinite).
__global__ void duplicate_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz, const int32_t* __restrict__ bstart,
                                 const int32_t* __restrict__ bcount, int64_t nleaf,
                                 int* __restrict__ flag) {
  int64_t k = blockIdx.x;
  if (k >= nleaf) return;
  int s0 = bstart[k], n = bcount[k];
  for (int a = threadIdx.x; a < n; a += blockDim.x)
    for (int b = a + 1; b < n; ++b)
      if (px[s0 + a] == px[s0 + b] && py[s0 + a] == py[s0 + b] && pz[s0 + a] == pz[s0 + b])
        atomicOr(flag, 1);
}

}  // namespace

void build_tree(Engine& e, const double* h_xyz) {
  auto t0 = std::chrono::steady_clock::now();
  const int nL = e.nL;
  const int64_t N = e.N;
  hipStream_t st = e.stream;
  // boxRadius (src/HODLR3DTree.cpp:22-26)
  e.radius.assign(nL + 1, 0.0);
  e.radius[0] = e.L;
  for (int j = 1; j <= nL; ++j) e.radius[j] = 0.5 * e.radius[j - 1];
  Shifts sh{};
  for (int j = 0; j < nL; ++j) sh.s[j] = 0.5 * e.radius[j];

  double* d_xyz = dalloc<double>(3 * N);
  H3D_HIP(hipMemcpyAsync(d_xyz, h_xyz, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
  uint32_t* d_key = dalloc<uint32_t>(N);
  uint32_t* d_key2 = dalloc<uint32_t>(N);
  int32_t* d_val = dalloc<int32_t>(N);
  e.d_perm = dalloc<int32_t>(N);
  int nb = (int)((N + 255) / 256);
  leaf_key_kernel<<<nb, 256, 0, st>>>(d_xyz, N, nL, e.ctr[0], e.ctr[1], e.ctr[2], sh, d_key,
                                      d_val);
  H3D_LAUNCH_CHECK();
  size_t tmp = 0;
  H3D_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, d_key, d_key2, d_val, e.d_perm,
                                              (int)N, 0, 3 * nL, st));
  void* d_tmp = dalloc<char>((int64_t)tmp + 16);
  H3D_HIP(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, d_key, d_key2, d_val, e.d_perm, (int)N,
                                              0, 3 * nL, st));
  dfree(d_tmp);
  const int64_t nleaf = int64_t(1) << (3 * nL);
  int32_t* d_cnt = dalloc<int32_t>(nleaf);
  H3D_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * nleaf, st));
  count_kernel<<<nb, 256, 0, st>>>(d_key2, N, d_cnt);
  H3D_LAUNCH_CHECK();
  // box segments for every level (host: O(#boxes) bookkeeping)
  const int64_t nbox_all = level_offset(nL + 1);
  e.bstart.assign(nbox_all, 0);
  e.bcount.assign(nbox_all, 0);
  H3D_HIP(hipMemcpyAsync(e.bcount.data() + level_offset(nL), d_cnt, sizeof(int32_t) * nleaf,
                         hipMemcpyDeviceToHost, st));
  H3D_HIP(hipStreamSynchronize(st));
  {
    int32_t pos = 0;
    int64_t o = level_offset(nL);
    for (int64_t k = 0; k < nleaf; ++k) {
      e.bstart[o + k] = pos;
      pos += e.bcount[o + k];
    }
    for (int j = nL - 1; j >= 0; --j) {
      int64_t oj = level_offset(j), oc = level_offset(j + 1);
      for (int64_t k = 0; k < (int64_t(1) << (3 * j)); ++k) {
        e.bstart[oj + k] = e.bstart[oc + 8 * k];
        int32_t c = 0;
        for (int q = 0; q < 8; ++q) c += e.bcount[oc + 8 * k + q];
        e.bcount[oj + k] = c;
      }
    }
  }
  e.d_bstart = dalloc<int32_t>(nbox_all);
  e.d_bcount = dalloc<int32_t>(nbox_all);
  H3D_HIP(hipMemcpyAsync(e.d_bstart, e.bstart.data(), sizeof(int32_t) * nbox_all,
                         hipMemcpyHostToDevice, st));
  H3D_HIP(hipMemcpyAsync(e.d_bcount, e.bcount.data(), sizeof(int32_t) * nbox_all,
                         hipMemcpyHostToDevice, st));
  // leaf centres by the reference recurrence (src/HODLR3DTree.cpp:1518-1556)
  {
    std::vector<double> c{e.ctr[0], e.ctr[1], e.ctr[2]}, nc;
    for (int j = 0; j < nL; ++j) {
      int64_t nbj = int64_t(1) << (3 * j);
      double shift = 0.5 * e.radius[j];
      nc.assign(nbj * 8 * 3, 0.0);
      for (int64_t k = 0; k < nbj; ++k)
        for (int q = 0; q < 8; ++q) {
          nc[3 * (8 * k + q) + 0] = cxb(q) ? c[3 * k] + shift : c[3 * k] - shift;
          nc[3 * (8 * k + q) + 1] = cyb(q) ? c[3 * k + 1] + shift : c[3 * k + 1] - shift;
          nc[3 * (8 * k + q) + 2] = czb(q) ? c[3 * k + 2] + shift : c[3 * k + 2] - shift;
        }
      c.swap(nc);
    }
    e.d_leafc = dalloc<double>(3 * nleaf);
    H3D_HIP(hipMemcpyAsync(e.d_leafc, c.data(), sizeof(double) * 3 * nleaf, hipMemcpyHostToDevice,
                           st));
    H3D_HIP(hipStreamSynchronize(st));
  }
  e.d_px = dalloc<double>(N);
  e.d_py = dalloc<double>(N);
  e.d_pz = dalloc<double>(N);
  e.d_lx = dalloc<float>(N);
  e.d_ly = dalloc<float>(N);
  e.d_lz = dalloc<float>(N);
  e.d_iperm = dalloc<int32_t>(N);
  gather_points_kernel<<<nb, 256, 0, st>>>(d_xyz, e.d_perm, d_key2, e.d_leafc, N, e.d_px, e.d_py,
                                           e.d_pz, e.d_lx, e.d_ly, e.d_lz, e.d_iperm);
  H3D_LAUNCH_CHECK();
  int* d_flag = dalloc<int>(1);
  H3D_HIP(hipMemsetAsync(d_flag, 0, sizeof(int), st));
  duplicate_kernel<<<(unsigned)nleaf, 64, 0, st>>>(e.d_px, e.d_py, e.d_pz,
                                                   e.d_bstart + level_offset(nL),
                                                   e.d_bcount + level_offset(nL), nleaf, d_flag);
  H3D_LAUNCH_CHECK();
  int flag = 0;
  H3D_HIP(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  e.perm.resize(N);
  H3D_HIP(hipMemcpyAsync(e.perm.data(), e.d_perm, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st));
  H3D_HIP(hipStreamSynchronize(st));
  dfree(d_flag);
  dfree(d_xyz);
  dfree(d_key);
  dfree(d_key2);
  dfree(d_val);
  dfree(d_cnt);
  // coincident points are fine for kernels finite at r = 0: Qchoice 0-6 and
  // 11 (src/kernel.cpp:76-143 evaluate phi(0)), a plugin whose entry of a
  // point with itself is finite, or a zero kernel
  if (flag) {
    const int q = e.kern.q;
    const h3d_point o{0.0, 0.0, 0.0};
    const bool finite0 = e.kern.scale == 0.0 || (q >= 0 && q <= 6) || q == 11 ||
                         (q == kQPlugin && std::isfinite(h3d_user_kernel::entry(o, o)));
    if (!finite0) throw Error(H3D_ERR_NONFINITE, "duplicate points (K(i,j) infinite at r = 0)");
  }

  // neighbour slots and interaction lists per level
  e.lev.assign(nL + 1, Level());
  for (int j = 1; j <= nL; ++j) {
    Level& lv = e.lev[j];
    lv.nboxes = int64_t(1) << (3 * j);
    int32_t* d_c = dalloc<int32_t>(lv.nboxes);
    lv.d_nbr = dalloc<int32_t>(lv.nboxes * 26);
    lv.d_iloff = dalloc<int32_t>(lv.nboxes + 1);
    unsigned g = (unsigned)((lv.nboxes + 127) / 128);
    il_kernel<<<g, 128, 0, st>>>(j, lv.nboxes, 0, d_c, nullptr, nullptr, lv.d_nbr);
    H3D_LAUNCH_CHECK();
    std::vector<int32_t> cnt(lv.nboxes);
    H3D_HIP(hipMemcpyAsync(cnt.data(), d_c, sizeof(int32_t) * lv.nboxes, hipMemcpyDeviceToHost, st));
    H3D_HIP(hipStreamSynchronize(st));
    lv.iloff.assign(lv.nboxes + 1, 0);
    for (int64_t k = 0; k < lv.nboxes; ++k) lv.iloff[k + 1] = lv.iloff[k] + cnt[k];
    H3D_HIP(hipMemcpyAsync(lv.d_iloff, lv.iloff.data(), sizeof(int32_t) * (lv.nboxes + 1),
                           hipMemcpyHostToDevice, st));
    int64_t nil = lv.iloff[lv.nboxes];
    lv.d_ilidx = dalloc<int32_t>(nil);
    il_kernel<<<g, 128, 0, st>>>(j, lv.nboxes, 1, nullptr, lv.d_iloff, lv.d_ilidx, nullptr);
    H3D_LAUNCH_CHECK();
    lv.ilidx.resize(nil);
    H3D_HIP(hipMemcpyAsync(lv.ilidx.data(), lv.d_ilidx, sizeof(int32_t) * nil,
                           hipMemcpyDeviceToHost, st));
    H3D_HIP(hipStreamSynchronize(st));
    dfree(d_c);
  }
  H3D_HIP(hipStreamSynchronize(st));
  e.t_tree_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace h3d
