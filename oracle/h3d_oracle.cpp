// h3d_oracle.cpp -- CPU restatement of the SAFRAN-LAB/HODLR3D hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
// entry point and bench.py's cpu_baseline leg may load it, and only to compare
// against or to time beside the HIP path.  Nothing in the product links it.
//
// It restates, in plain C++17 + OpenMP without Eigen, the algorithm of
//   src/HODLR3D.cpp:3-27             facade: nLevels, tree set-up, assemble, matvec
//   src/HODLR3DTree.cpp:14-32        box radii
//   src/HODLR3DTree.cpp:110-1489     neighbour slots and interaction lists
//   src/HODLR3DTree.cpp:1518-1556    box centres
//   src/HODLR3DTree.cpp:1558-1619    bucketing and non-leaf concatenation
//   src/HODLR3DTree.cpp:1621-1645    assembly loop
//   src/ACA.cpp:3-23, 127-238        partially pivoted ACA (CUR form)
//   src/HODLR3DTree.cpp:1647-1805    far field, near field, collect, reorder
//   src/HODLR3DTree.cpp:1757-1818    findMemory, getMaxRank
//   src/gmres.cpp:5-94               unrestarted GMRES
//   src/kernel.cpp:65-227            kernel catalogue
//   examples/testHODLR3Dsolve.cpp:32-76  Qchoice dispatch and i==j convention
//   src/Integral3D.cpp:9-43          self-cell integral for Qchoice 16
//
// Reductions.  The reference takes vector norms and dot products through Eigen,
// whose summation order depends on the (unpinned) Eigen version and SIMD width.
// The oracle fixes one order, "canonical": element s goes to lane s % 256, each
// lane sums its elements in increasing s, then lanes are combined by a pairwise
// tree (lane t += lane t+d for d = 128, 64, ..., 1).  The GPU ACA uses the same
// order, which makes its pivots and ranks bit-comparable with this file.  Mode
// ORC_RED_EIGEN_SSE2 restates Eigen 3.3/3.4's 2-wide packet reduction so that a
// test can count how often the choice of order changes an ACA decision.
//
// Floating point: build with -ffp-contract=off (no FMA), as the reference's
// default x86-64 gcc build has none; pow(x,2) is written x*x, which gcc folds
// it to.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

// The user kernel plugin (h3d_user_kernel, Qchoice 100), the same header the
// library is built with.
#ifdef H3D_KERNEL_PLUGIN
#include H3D_KERNEL_PLUGIN
#else
#include "../include/h3d_plugin_plummer.hpp"
#endif

namespace {

// ---------------------------------------------------------------- kernels --
// src/kernel.cpp:65-227 and the dispatch of examples/testHODLR3Dsolve.cpp:32-76.
struct KernelSpec {
  int q = 7;            // Qchoice
  double scale = 1.0;   // multiplier of the off-diagonal entry (h3 for Qchoice 16)
  double diag = 0.0;    // K(i,i) (0 for Qchoice 0..15, Kii for 16)
  const double* diag_vec = nullptr;  // optional per-point diagonal (original order)
};

inline double phi(int q, double R2) {
  double R = std::sqrt(R2);
  switch (q) {
    case 0: return std::log(1.0 + R / 1.0);
    case 1: return std::exp(-R / 1.0);
    case 2: return 1.0 / (1.0 + (R / 1.0) * (R / 1.0));
    case 3: return 1.0 + (R / 1.0) * (R / 1.0);
    case 4: return 1.0 / std::sqrt((R / 1.0) * (R / 1.0) + 1);
    case 5: return std::exp(-(R / 1.0) * (R / 1.0));
    case 6: return std::sqrt((R / 1.0) * (R / 1.0) + 1);
    case 7: return 1 / R;
    case 8: return 1.0 / (R2 * R2);
    case 9: return std::log(R);
    case 10: return R * R * std::log(R);
    case 11: return std::cyl_bessel_j(0.0, 1.0 * R);
    case 12: return std::cyl_neumann(0.0, 1.0 * R);
    case 13: return std::cos(1.0 * R) / R;
    case 14: return std::cyl_bessel_k(0.5, R) / std::sqrt(R);
    case 15: return std::exp(-50.0 * R) / R;
    case 16: return 1 / R;
    default: return std::numeric_limits<double>::quiet_NaN();
  }
}

// Entry between two points given by coordinates; i, j are original indices.
inline double entry(const KernelSpec& k, double xi, double yi, double zi, double xj, double yj,
                    double zj, int64_t i, int64_t j) {
  if (i == j) return k.diag_vec ? k.diag_vec[i] : k.diag;
  if (k.q == 100) return k.scale * h3d_user_kernel::entry({xi, yi, zi}, {xj, yj, zj});
  double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  double R2 = dx * dx + dy * dy + dz * dz;
  return k.scale * phi(k.q, R2);
}

// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n
// (Golub & Welsch give the same nodes; Newton from the Chebyshev guess is the
// textbook route, e.g. Press et al., "gauleg").
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    long double z = std::cos(M_PI * (i + 0.75L) / (n + 0.5L)), pp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs((double)(z - z1)) < 1e-19) break;
    }
    x[n - 1 - i] = (double)z;  // ascending order
    w[n - 1 - i] = (double)(2 / ((1 - z * z) * pp * pp));
  }
}

// src/Integral3D.cpp:9-43 (with `result` starting from 0).
double triple_integral_cube(double half) {
  const int n = 16;  // include/definitions.hpp:5
  std::vector<double> x, w;
  gauss_legendre(n, x, w);
  double c = 0.5 * (half - 0.0), Lc = 0.5 * (half + 0.0);
  double result = 0.0;
  for (int i = 0; i < n; i++) {
    double tx = c * x[i] + Lc;
    for (int j = 0; j < n; j++) {
      double ty = c * x[j] + Lc;
      for (int k = 0; k < n; k++) {
        double tz = c * x[k] + Lc;
        result += (w[i] * w[j] * w[k] * (1.0 / ((4 * 3.1415926535897932384) *
                                                 std::sqrt(tx * tx + ty * ty + tz * tz))));
      }
    }
  }
  result *= (c * c * c);
  result *= 8;
  return result;
}

// ------------------------------------------------------------- reductions --
enum { ORC_RED_CANON = 0, ORC_RED_EIGEN_SSE2 = 1 };
int g_red_mode = ORC_RED_CANON;

template <class F>
double canon_reduce(int64_t n, F term) {
  double acc[256];
  for (int t = 0; t < 256; ++t) acc[t] = 0.0;
  for (int64_t s = 0; s < n; ++s) {
    int t = (int)(s & 255);
    acc[t] = acc[t] + term(s);
  }
  for (int d = 128; d >= 1; d >>= 1)
    for (int t = 0; t < d; ++t) acc[t] = acc[t] + acc[t + d];
  return acc[0];
}

// Eigen's LinearVectorizedTraversal redux with 2-wide packets and 16-byte
// aligned storage: two packet accumulators over strides of 4, a trailing
// packet, a horizontal add, then scalar tail.
template <class F>
double eigen_sse2_reduce(int64_t n, F term) {
  if (n == 0) return 0.0;
  const int64_t ps = 2;
  int64_t aligned = (n / ps) * ps, aligned2 = (n / (2 * ps)) * (2 * ps);
  double res;
  if (aligned) {
    double a0 = term(0), a1 = term(1);
    if (aligned > ps) {
      double b0 = term(2), b1 = term(3);
      for (int64_t s = 4; s < aligned2; s += 4) {
        a0 = a0 + term(s);
        a1 = a1 + term(s + 1);
        b0 = b0 + term(s + 2);
        b1 = b1 + term(s + 3);
      }
      a0 = a0 + b0;
      a1 = a1 + b1;
      if (aligned > aligned2) {
        a0 = a0 + term(aligned2);
        a1 = a1 + term(aligned2 + 1);
      }
    }
    res = a0 + a1;
    for (int64_t s = aligned; s < n; ++s) res = res + term(s);
  } else {
    res = term(0);
    for (int64_t s = 1; s < n; ++s) res = res + term(s);
  }
  return res;
}

inline double vsum(int64_t n, const double* a, const double* b) {
  auto f = [&](int64_t s) { return a[s] * b[s]; };
  return g_red_mode == ORC_RED_CANON ? canon_reduce(n, f) : eigen_sse2_reduce(n, f);
}
inline double vnorm(int64_t n, const double* a) { return std::sqrt(vsum(n, a, a)); }

// ------------------------------------------------------------------ tree --
inline void slot_offset(int s, int& dx, int& dy, int& dz) {
  int m = s < 13 ? s : s + 1;
  dx = m % 3 - 1;
  dy = (m % 9) / 3 - 1;
  dz = m / 9 - 1;
}
// Child position bits (src/HODLR3DTree.cpp:1518-1556 sign table).
inline int cxb(int c) { return (c == 1 || c == 2 || c == 5 || c == 6) ? 1 : 0; }
inline int cyb(int c) { return (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0; }
inline int czb(int c) { return c >= 4 ? 1 : 0; }
inline int child_of_bits(int xb, int yb, int zb) {
  return zb * 4 + (yb ? (xb ? 2 : 3) : (xb ? 1 : 0));
}

struct Block {
  int32_t src = -1;   // source box ki
  int32_t rank = 0;
  int32_t status = 0; // 0 ok, 1 non-finite entry
  std::vector<int32_t> I, J;           // local pivots (row within CL_k, col within CL_ki)
  std::vector<double> L, R;            // r x r row-major
  std::vector<double> U, V;            // optional: u_l (n each), v_l (m each)
};

struct Handle {
  int64_t N = 0;
  int nLevels = 0;
  int leaf1d = 0;
  double Lhalf = 1.0;
  double ctr[3] = {0, 0, 0};
  int tol_pow = 7;
  double tol = 1e-7;
  KernelSpec ks;
  std::vector<double> diag_store;
  std::vector<double> x, y, z;            // original order
  std::vector<double> px, py, pz;         // permuted order
  std::vector<int32_t> perm;              // permuted position -> original index
  std::vector<double> radius;             // boxRadius[j]
  std::vector<std::vector<double>> cx, cy, cz;  // centres per level
  std::vector<std::vector<int32_t>> bstart, bcount;  // segment of perm per box
  std::vector<std::vector<int32_t>> nbr;  // [level][box*26+slot]
  std::vector<std::vector<int32_t>> iloff, ilidx;  // IL CSR per level
  std::vector<std::vector<Block>> blocks;  // per level, CSR-aligned with ilidx
  std::vector<std::vector<uint8_t>> assembled;  // per level, per target box
  bool keep_uv = false;
  double assTime = 0, matVecTime = 0;
};

int nlevels_formula(int64_t N, int leaf1d) {
  // src/HODLR3D.cpp:6-7
  double cubeRootN = std::pow((double)N, 1.0 / 3.0);
  return (int)std::ceil(3 * std::log(double(cubeRootN) / leaf1d) / std::log(8));
}

void build_tree(Handle& h) {
  const int nL = h.nLevels;
  // src/HODLR3DTree.cpp:22-26
  h.radius.assign(nL + 1, 0.0);
  h.radius[0] = h.Lhalf;
  for (int j = 1; j <= nL; ++j) h.radius[j] = 0.5 * h.radius[j - 1];
  // centres, src/HODLR3DTree.cpp:1518-1556
  h.cx.assign(nL + 1, {});
  h.cy.assign(nL + 1, {});
  h.cz.assign(nL + 1, {});
  h.cx[0] = {h.ctr[0]};
  h.cy[0] = {h.ctr[1]};
  h.cz[0] = {h.ctr[2]};
  for (int j = 0; j < nL; ++j) {
    int64_t nb = 1LL << (3 * j);
    double shift = 0.5 * h.radius[j];
    h.cx[j + 1].resize(nb * 8);
    h.cy[j + 1].resize(nb * 8);
    h.cz[j + 1].resize(nb * 8);
    for (int64_t k = 0; k < nb; ++k)
      for (int c = 0; c < 8; ++c) {
        h.cx[j + 1][8 * k + c] = cxb(c) ? h.cx[j][k] + shift : h.cx[j][k] - shift;
        h.cy[j + 1][8 * k + c] = cyb(c) ? h.cy[j][k] + shift : h.cy[j][k] - shift;
        h.cz[j + 1][8 * k + c] = czb(c) ? h.cz[j][k] + shift : h.cz[j][k] - shift;
      }
  }
  // bucketing, src/HODLR3DTree.cpp:1558-1608 (top-down, stable within a box)
  std::vector<std::vector<int32_t>> lists(1);
  lists[0].resize(h.N);
  for (int64_t i = 0; i < h.N; ++i) lists[0][i] = (int32_t)i;
  for (int j = 0; j < nL; ++j) {
    int64_t nb = 1LL << (3 * j);
    std::vector<std::vector<int32_t>> next(nb * 8);
    for (int64_t k = 0; k < nb; ++k) {
      double ccx = h.cx[j][k], ccy = h.cy[j][k], ccz = h.cz[j][k];
      for (int32_t idx : lists[k]) {
        int zb = (h.z[idx] <= ccz) ? 0 : 1;
        int xb = (h.x[idx] <= ccx) ? 0 : 1;
        int yb = (h.y[idx] <= ccy) ? 0 : 1;
        next[8 * k + child_of_bits(xb, yb, zb)].push_back(idx);
      }
    }
    lists.swap(next);
  }
  // leaf-DFS permutation; non-leaf lists are concatenations (:1610-1619), so
  // every box at every level is a contiguous segment of perm.
  h.perm.clear();
  h.perm.reserve(h.N);
  int64_t nleaf = 1LL << (3 * nL);
  for (int64_t k = 0; k < nleaf; ++k)
    for (int32_t idx : lists[k]) h.perm.push_back(idx);
  h.bstart.assign(nL + 1, {});
  h.bcount.assign(nL + 1, {});
  h.bstart[nL].resize(nleaf);
  h.bcount[nL].resize(nleaf);
  int32_t pos = 0;
  for (int64_t k = 0; k < nleaf; ++k) {
    h.bstart[nL][k] = pos;
    h.bcount[nL][k] = (int32_t)lists[k].size();
    pos += (int32_t)lists[k].size();
  }
  for (int j = nL - 1; j >= 0; --j) {
    int64_t nb = 1LL << (3 * j);
    h.bstart[j].resize(nb);
    h.bcount[j].resize(nb);
    for (int64_t k = 0; k < nb; ++k) {
      h.bstart[j][k] = h.bstart[j + 1][8 * k];
      int32_t c = 0;
      for (int q = 0; q < 8; ++q) c += h.bcount[j + 1][8 * k + q];
      h.bcount[j][k] = c;
    }
  }
  h.px.resize(h.N);
  h.py.resize(h.N);
  h.pz.resize(h.N);
  for (int64_t s = 0; s < h.N; ++s) {
    h.px[s] = h.x[h.perm[s]];
    h.py[s] = h.y[h.perm[s]];
    h.pz[s] = h.z[h.perm[s]];
  }
  // neighbours and interaction lists: the closed form of the reference's
  // per-child assignments (src/HODLR3DTree.cpp:110-1489).  Level j boxes have
  // integer coordinates in [0,2^j)^3; a child's neighbour slots hold the
  // same-level boxes at Chebyshev distance 1 and Manhattan distance <= 2; its
  // list is the corner sibling, then for each parent neighbour slot in
  // ascending order the corner-adjacent child of that neighbour followed by its
  // non-adjacent children in ascending child order.
  h.nbr.assign(nL + 1, {});
  h.iloff.assign(nL + 1, {});
  h.ilidx.assign(nL + 1, {});
  h.nbr[0].assign(26, -1);
  h.iloff[0] = {0, 0};
  for (int j = 1; j <= nL; ++j) {
    int64_t nb = 1LL << (3 * j);
    int side = 1 << j;
    std::vector<int32_t> ix(nb), iy(nb), iz(nb);
    std::vector<int32_t> at((size_t)side * side * side);
    for (int64_t k = 0; k < nb; ++k) {
      int X = 0, Y = 0, Z = 0;
      for (int t = 0; t < j; ++t) {
        int c = (int)((k >> (3 * (j - 1 - t))) & 7);
        X = 2 * X + cxb(c);
        Y = 2 * Y + cyb(c);
        Z = 2 * Z + czb(c);
      }
      ix[k] = X;
      iy[k] = Y;
      iz[k] = Z;
      at[((size_t)Z * side + Y) * side + X] = (int32_t)k;
    }
    auto box_at = [&](int X, int Y, int Z) -> int32_t {
      if (X < 0 || Y < 0 || Z < 0 || X >= side || Y >= side || Z >= side) return -1;
      return at[((size_t)Z * side + Y) * side + X];
    };
    h.nbr[j].assign(nb * 26, -1);
    for (int64_t k = 0; k < nb; ++k)
      for (int s = 0; s < 26; ++s) {
        int dx, dy, dz;
        slot_offset(s, dx, dy, dz);
        if (std::abs(dx) + std::abs(dy) + std::abs(dz) > 2) continue;
        h.nbr[j][k * 26 + s] = box_at(ix[k] + dx, iy[k] + dy, iz[k] + dz);
      }
    h.iloff[j].assign(nb + 1, 0);
    std::vector<int32_t>& il = h.ilidx[j];
    il.clear();
    for (int64_t k = 0; k < nb; ++k) {
      int64_t p = k / 8;
      int c = (int)(k % 8);
      il.push_back((int32_t)(8 * p + child_of_bits(1 - cxb(c), 1 - cyb(c), 1 - czb(c))));
      if (j >= 2) {
        int PX = ix[k] >> 1, PY = iy[k] >> 1, PZ = iz[k] >> 1;
        for (int s = 0; s < 26; ++s) {
          int dx, dy, dz;
          slot_offset(s, dx, dy, dz);
          if (std::abs(dx) + std::abs(dy) + std::abs(dz) > 2) continue;
          int QX = PX + dx, QY = PY + dy, QZ = PZ + dz;
          if (QX < 0 || QY < 0 || QZ < 0 || QX >= side / 2 || QY >= side / 2 || QZ >= side / 2)
            continue;
          int32_t corner = -1;
          std::vector<int32_t> far;
          for (int cc = 0; cc < 8; ++cc) {
            int CX = 2 * QX + cxb(cc), CY = 2 * QY + cyb(cc), CZ = 2 * QZ + czb(cc);
            int ax = std::abs(CX - ix[k]), ay = std::abs(CY - iy[k]), az = std::abs(CZ - iz[k]);
            int cheb = std::max(ax, std::max(ay, az));
            int man = ax + ay + az;
            int32_t b = box_at(CX, CY, CZ);
            if (cheb >= 2) far.push_back(b);
            else if (man == 3) corner = b;
          }
          if (corner >= 0) il.push_back(corner);
          for (int32_t b : far) il.push_back(b);
        }
      }
      h.iloff[j][k + 1] = (int32_t)il.size();
    }
  }
  h.blocks.assign(nL + 1, {});
  h.assembled.assign(nL + 1, {});
  for (int j = 1; j <= nL; ++j) {
    h.blocks[j].assign(h.ilidx[j].size(), Block());
    h.assembled[j].assign((size_t)1 << (3 * j), 0);
  }
}

// ----------------------------------------------------------------- ACA --
// src/ACA.cpp:127-238, LowRank::ACA_only_nodesCUR; rows = CL_k, cols = CL_ki.
void aca_block(const Handle& h, int rs, int n, int cs, int m, Block& B) {
  B.rank = 0;
  B.status = 0;
  B.I.clear();
  B.J.clear();
  B.L.clear();
  B.R.clear();
  B.U.clear();
  B.V.clear();
  if (n == 0 || m == 0) return;  // empty row box: the reference returns rank 0;
                                 // empty column box: undefined there, rank 0 here
  const KernelSpec& ks = h.ks;
  auto K = [&](int r, int c) {
    int64_t a = rs + r, b = cs + c;
    return entry(ks, h.px[a], h.py[a], h.pz[a], h.px[b], h.py[b], h.pz[b], h.perm[a], h.perm[b]);
  };
  const int mn = std::min(n, m);
  std::vector<uint8_t> rrem(n, 1), crem(m, 1);
  std::vector<double> row(m), col(n);
  std::vector<double>& U = B.U;  // u_l stored at l*n
  std::vector<double>& V = B.V;  // v_l stored at l*m
  auto argmax_rem = [](const std::vector<double>& v, const std::vector<uint8_t>& rem) {
    int idx = -1;
    double mx = 0;
    for (size_t s = 0; s < v.size(); ++s) {
      if (!rem[s]) continue;
      double a = std::fabs(v[s]);
      if (idx < 0) {
        idx = (int)s;
        mx = a;
      } else if (a > mx) {
        idx = (int)s;
        mx = a;
      }
    }
    return idx;
  };
  bool nonfinite = false;
  // first row (ACA.cpp:152-161)
  int i = -1, jj = -1;
  for (int l = 0; l < n; ++l) {
    for (int s = 0; s < m; ++s) row[s] = K(l, s);
    int c = argmax_rem(row, crem);
    if (std::fabs(row[c]) > 1e-16) {
      i = l;
      jj = c;
      break;
    }
  }
  if (i < 0) return;
  for (int s = 0; s < m; ++s) nonfinite |= !std::isfinite(row[s]);
  double piv = row[jj];
  for (int s = 0; s < n; ++s) col[s] = K(s, jj);
  for (int s = 0; s < n; ++s) nonfinite |= !std::isfinite(col[s]);
  V.insert(V.end(), row.begin(), row.end());
  for (int s = 0; s < n; ++s) U.push_back(col[s] / piv);
  B.I.push_back(i);
  B.J.push_back(jj);
  crem[jj] = 0;
  rrem[i] = 0;
  int r = 1;
  double normS = 0.0;
  double nu = vnorm(n, &U[0]), nv = vnorm(m, &V[0]);
  int inext = (r < n) ? argmax_rem(col, rrem) : -1;  // ACA.cpp:171 (guarded)
  while (nu * nv > h.tol * normS && r < mn) {
    i = inext;
    // residual row (ACA.cpp:181-187): l outer, element inner, no FMA
    for (int s = 0; s < m; ++s) row[s] = K(i, s);
    for (int l = 0; l < r; ++l) {
      double ul = U[(size_t)l * n + i];
      const double* vl = &V[(size_t)l * m];
      for (int s = 0; s < m; ++s) row[s] -= ul * vl[s];
    }
    jj = argmax_rem(row, crem);
    // an exactly zero pivot (residual row zero on every remaining column, e.g.
    // coincident points): the reference divides by it and its factors turn
    // NaN; stop here, as the small-norm test below does (ACA.cpp:196-200)
    if (row[jj] == 0.0) break;
    for (int s = 0; s < n; ++s) col[s] = K(s, jj);
    for (int l = 0; l < r; ++l) {
      double vl = V[(size_t)l * m + jj];
      const double* ul = &U[(size_t)l * n];
      for (int s = 0; s < n; ++s) col[s] -= vl * ul[s];
    }
    for (int s = 0; s < m; ++s) nonfinite |= !std::isfinite(row[s]);
    for (int s = 0; s < n; ++s) nonfinite |= !std::isfinite(col[s]);
    piv = row[jj];
    size_t uoff = U.size(), voff = V.size();
    for (int s = 0; s < n; ++s) U.push_back(col[s] / piv);
    V.insert(V.end(), row.begin(), row.end());
    double nu2 = vnorm(n, &U[uoff]), nv2 = vnorm(m, &V[voff]);
    if (nu2 < 1e-16 || nv2 < 1e-16) {  // ACA.cpp:196-200
      U.resize(uoff);
      V.resize(voff);
      break;
    }
    B.I.push_back(i);
    B.J.push_back(jj);
    ++r;
    crem[jj] = 0;
    rrem[i] = 0;
    nu = nu2;
    nv = nv2;
    // normS update (ACA.cpp:206-212); the sum includes the new pair itself
    double ns2 = normS * normS + (nu * nv) * (nu * nv);
    const double* un = &U[uoff];
    const double* vn = &V[voff];
    for (int l = 0; l < r; ++l)
      ns2 += 2 * std::fabs(vsum(n, &U[(size_t)l * n], un) * vsum(m, &V[(size_t)l * m], vn));
    normS = std::sqrt(ns2);
    inext = (r < n) ? argmax_rem(col, rrem) : -1;  // ACA.cpp:213 on the raw residual column
  }
  B.rank = r;
  B.status = nonfinite ? 1 : 0;
  // L, R (ACA.cpp:215-236), row-major r x r
  B.L.assign((size_t)r * r, 0.0);
  B.R.assign((size_t)r * r, 0.0);
  for (int a = 0; a < r; ++a) {
    B.L[(size_t)a * r + a] = 1.0;
    for (int b = 0; b < a; ++b) B.L[(size_t)a * r + b] = U[(size_t)b * n + B.I[a]];
    B.R[(size_t)a * r + a] = V[(size_t)a * m + B.J[a]];
    for (int b = 0; b < a; ++b) B.R[(size_t)b * r + a] = V[(size_t)b * m + B.J[a]];
  }
  if (!h.keep_uv) {
    std::vector<double>().swap(B.U);
    std::vector<double>().swap(B.V);
  }
}

void assemble_level(Handle& h, int j, const std::vector<int32_t>& targets) {
  // src/HODLR3DTree.cpp:1621-1631: boxes in parallel, IL entries in order
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < (int64_t)targets.size(); ++t) {
    int32_t k = targets[t];
    for (int32_t b = h.iloff[j][k]; b < h.iloff[j][k + 1]; ++b) {
      int32_t ki = h.ilidx[j][b];
      Block& B = h.blocks[j][b];
      B.src = ki;
      aca_block(h, h.bstart[j][k], h.bcount[j][k], h.bstart[j][ki], h.bcount[j][ki], B);
    }
    h.assembled[j][k] = 1;
  }
}

// -------------------------------------------------------------- matvec --
// src/HODLR3DTree.cpp:1660-1706 (far) with entry regeneration, CUR form.
void far_block_apply(const Handle& h, int j, int k, const Block& B, const double* qp, double* pot,
                     double& tgen, double& tmv) {
  const int r = B.rank;
  if (r == 0) return;
  const int rs = h.bstart[j][k], n = h.bcount[j][k];
  const int cs = h.bstart[j][B.src], m = h.bcount[j][B.src];
  const KernelSpec& ks = h.ks;
  double t0 = omp_get_wtime();
  std::vector<double> Ac((size_t)n * r), Ar((size_t)r * m);
  for (int a = 0; a < n; ++a)
    for (int l = 0; l < r; ++l) {
      int64_t p = rs + a, q = cs + B.J[l];
      Ac[(size_t)a * r + l] = entry(ks, h.px[p], h.py[p], h.pz[p], h.px[q], h.py[q], h.pz[q],
                                    h.perm[p], h.perm[q]);
    }
  for (int l = 0; l < r; ++l)
    for (int b = 0; b < m; ++b) {
      int64_t p = rs + B.I[l], q = cs + b;
      Ar[(size_t)l * m + b] = entry(ks, h.px[p], h.py[p], h.pz[p], h.px[q], h.py[q], h.pz[q],
                                    h.perm[p], h.perm[q]);
    }
  double t1 = omp_get_wtime();
  std::vector<double> t(r);
  for (int l = 0; l < r; ++l) {
    double s = 0;
    for (int b = 0; b < m; ++b) s += Ar[(size_t)l * m + b] * qp[cs + b];
    t[l] = s;
  }
  // forward solve with unit-lower L (column oriented)
  for (int a = 0; a < r; ++a) {
    t[a] = t[a] / B.L[(size_t)a * r + a];
    for (int c = a + 1; c < r; ++c) t[c] -= B.L[(size_t)c * r + a] * t[a];
  }
  // back solve with upper R (column oriented)
  for (int a = r - 1; a >= 0; --a) {
    t[a] = t[a] / B.R[(size_t)a * r + a];
    for (int c = 0; c < a; ++c) t[c] -= B.R[(size_t)c * r + a] * t[a];
  }
  for (int a = 0; a < n; ++a) {
    double s = 0;
    for (int l = 0; l < r; ++l) s += Ac[(size_t)a * r + l] * t[l];
    pot[rs + a] += s;
  }
  double t2 = omp_get_wtime();
  tgen += t1 - t0;
  tmv += t2 - t1;
}

// Same block, applied from the ACA vectors: sum_l u_l (v_l . q).
void far_block_apply_uv(const Handle& h, int j, int k, const Block& B, const double* qp,
                        double* pot) {
  const int r = B.rank;
  if (r == 0) return;
  const int rs = h.bstart[j][k], n = h.bcount[j][k];
  const int cs = h.bstart[j][B.src], m = h.bcount[j][B.src];
  for (int l = 0; l < r; ++l) {
    double s = 0;
    for (int b = 0; b < m; ++b) s += B.V[(size_t)l * m + b] * qp[cs + b];
    for (int a = 0; a < n; ++a) pot[rs + a] += B.U[(size_t)l * n + a] * s;
  }
}

void near_apply(const Handle& h, const double* qp, double* pot, const uint8_t* leafmask,
                double& tgen, double& tmv) {
  // src/HODLR3DTree.cpp:1708-1755: slots 0..25, then self
  const int nL = h.nLevels;
  const int64_t nleaf = (int64_t)1 << (3 * nL);
  const KernelSpec& ks = h.ks;
  double gmax = 0, mmax = 0;
#pragma omp parallel
  {
    double tg = 0, tm = 0;
    std::vector<double> blk;
#pragma omp for schedule(dynamic, 16) nowait
    for (int64_t k = 0; k < nleaf; ++k) {
      if (leafmask && !leafmask[k]) continue;
      const int rs = h.bstart[nL][k], n = h.bcount[nL][k];
      for (int s = 0; s <= 26; ++s) {
        int32_t nn = (s < 26) ? h.nbr[nL][k * 26 + s] : (int32_t)k;
        if (nn < 0) continue;
        const int cs = h.bstart[nL][nn], m = h.bcount[nL][nn];
        double a0 = omp_get_wtime();
        blk.resize((size_t)n * m);
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < m; ++b) {
            int64_t p = rs + a, q = cs + b;
            blk[(size_t)a * m + b] = entry(ks, h.px[p], h.py[p], h.pz[p], h.px[q], h.py[q],
                                           h.pz[q], h.perm[p], h.perm[q]);
          }
        double a1 = omp_get_wtime();
        for (int a = 0; a < n; ++a) {
          double acc = 0;
          for (int b = 0; b < m; ++b) acc += blk[(size_t)a * m + b] * qp[cs + b];
          pot[rs + a] += acc;
        }
        double a2 = omp_get_wtime();
        tg += a1 - a0;
        tm += a2 - a1;
      }
    }
#pragma omp critical
    {
      gmax = std::max(gmax, tg);
      mmax = std::max(mmax, tm);
    }
  }
  tgen += gmax;
  tmv += mmax;
}

// Full matvec in original order (src/HODLR3D.cpp:21-27).  use_uv selects the
// ACA-vector form; leafmask restricts the near field.
void matvec(Handle& h, const double* q, double* out, bool use_uv, const uint8_t* leafmask) {
  const int nL = h.nLevels;
  const int64_t N = h.N;
  if (use_uv && !h.keep_uv) {  // the ACA vectors were not kept: no vector form
    for (int64_t s = 0; s < N; ++s) out[s] = std::numeric_limits<double>::quiet_NaN();
    return;
  }
  std::vector<double> qp(N), total(N, 0.0), pot(N);
  for (int64_t s = 0; s < N; ++s) qp[s] = q[h.perm[s]];  // assignCharges (:1647-1658)
  for (int j = 1; j <= nL; ++j) {
    std::fill(pot.begin(), pot.end(), 0.0);
    int64_t nb = (int64_t)1 << (3 * j);
    double gmax = 0, mmax = 0;
#pragma omp parallel
    {
      double tg = 0, tm = 0;
#pragma omp for schedule(dynamic, 4) nowait
      for (int64_t k = 0; k < nb; ++k) {
        if (!h.assembled[j][k]) continue;
        for (int32_t b = h.iloff[j][k]; b < h.iloff[j][k + 1]; ++b) {
          if (use_uv)
            far_block_apply_uv(h, j, (int)k, h.blocks[j][b], qp.data(), pot.data());
          else
            far_block_apply(h, j, (int)k, h.blocks[j][b], qp.data(), pot.data(), tg, tm);
        }
      }
#pragma omp critical
      {
        gmax = std::max(gmax, tg);
        mmax = std::max(mmax, tm);
      }
    }
    h.assTime += gmax;
    h.matVecTime += mmax;
    if (j == nL) near_apply(h, qp.data(), pot.data(), leafmask, h.assTime, h.matVecTime);
    for (int64_t s = 0; s < N; ++s) total[s] = total[s] + pot[s];  // collectPotential
  }
  for (int64_t s = 0; s < N; ++s) out[h.perm[s]] = total[s];  // reorder (:1795-1805)
}

// ------------------------------------------------------------- GMRES --
// src/gmres.cpp:5-94.  The final (K+1)x(K+1) system is upper triangular after
// the Givens rotations; it is solved by back substitution.
int gmres(Handle& h, const double* b, int maxit, double thr, double* x, double* err_out,
          int* its_out, double* hist) {
  const int64_t N = h.N;
  const int m = maxit;
  std::vector<double> xv(N, 0.0), temp(N), r(N);
  matvec(h, xv.data(), temp.data(), false, nullptr);
  for (int64_t s = 0; s < N; ++s) r[s] = b[s] - temp[s];
  double b_norm = vnorm(N, b);
  double error = vnorm(N, r.data()) / b_norm;
  std::vector<double> sn(m, 0.0), cs(m, 0.0), beta(m + 1, 0.0);
  int nh = 0;
  hist[nh++] = error;
  double r_norm = vnorm(N, r.data());
  std::vector<double> Q((size_t)N * (m + 1), 0.0);
  std::vector<double> H((size_t)(m + 1) * m, 0.0);  // column k at k*(m+1)
  for (int64_t s = 0; s < N; ++s) Q[s] = r[s] / r_norm;
  beta[0] = r_norm * 1.0;
  int K = m - 1;
  std::vector<double> q(N), hv(m + 1);
  for (int k = 0; k < m; ++k) {
    // arnoldi (:29-38), modified Gram-Schmidt
    matvec(h, &Q[(size_t)k * N], q.data(), false, nullptr);
    for (int i = 0; i <= k; ++i) {
      hv[i] = vsum(N, q.data(), &Q[(size_t)i * N]);
      for (int64_t s = 0; s < N; ++s) q[s] = q[s] - hv[i] * Q[(size_t)i * N + s];
    }
    hv[k + 1] = vnorm(N, q.data());
    for (int64_t s = 0; s < N; ++s) Q[(size_t)(k + 1) * N + s] = q[s] / hv[k + 1];
    // applyGivensRotation (:14-26)
    for (int i = 0; i < k; ++i) {
      double tmp = cs[i] * hv[i] + sn[i] * hv[i + 1];
      hv[i + 1] = -sn[i] * hv[i] + cs[i] * hv[i + 1];
      hv[i] = tmp;
    }
    double t = std::sqrt(hv[k] * hv[k] + hv[k + 1] * hv[k + 1]);
    double cs_k = hv[k] / t, sn_k = hv[k + 1] / t;
    hv[k] = cs_k * hv[k] + sn_k * hv[k + 1];
    hv[k + 1] = 0.0;
    for (int i = 0; i <= k + 1; ++i) H[(size_t)k * (m + 1) + i] = hv[i];
    cs[k] = cs_k;
    sn[k] = sn_k;
    beta[k + 1] = -sn[k] * beta[k];
    beta[k] = cs[k] * beta[k];
    error = std::fabs(beta[k + 1]) / b_norm;
    hist[nh++] = error;
    if (error <= thr) {
      K = k;
      break;
    }
  }
  std::vector<double> y(K + 1);
  for (int a = K; a >= 0; --a) {
    double s = beta[a];
    for (int c = a + 1; c <= K; ++c) s -= H[(size_t)c * (m + 1) + a] * y[c];
    y[a] = s / H[(size_t)a * (m + 1) + a];
  }
  for (int64_t s = 0; s < N; ++s) {
    double acc = 0;
    for (int c = 0; c <= K; ++c) acc += Q[(size_t)c * N + s] * y[c];
    x[s] = xv[s] + acc;
  }
  *err_out = error;
  *its_out = K + 1;
  return nh;
}

}  // namespace

// ================================================================ C API ==
extern "C" {

int orc_nlevels(int64_t N, int leaf1d) { return nlevels_formula(N, leaf1d); }

void orc_set_reduction(int mode) { g_red_mode = mode; }

double orc_ie_cube_kii(int64_t N) {
  // include/kernel.hpp:45-60: h = 1/cbrt(N), integral over [0,h/2]^3 (x8) + 1
  double hh = 1.0 / std::cbrt(double(N));
  return triple_integral_cube(hh * 0.5) + 1;
}

void orc_gauss_legendre(int n, double* x, double* w) {
  std::vector<double> xv, wv;
  gauss_legendre(n, xv, wv);
  for (int i = 0; i < n; ++i) {
    x[i] = xv[i];
    w[i] = wv[i];
  }
}

// Create a handle; xyz is N x 3 (pts3D layout).  nlevels < 0 uses the
// reference formula.  Returns nullptr on invalid input.
void* orc_create(const double* xyz, int64_t N, int leaf1d, double L, const double* center,
                 int nlevels, int tol_pow, int qchoice, double scale, double diag,
                 const double* diag_vec, int keep_uv) {
  if (N <= 0 || leaf1d <= 0) return nullptr;
  Handle* h = new Handle();
  h->N = N;
  h->leaf1d = leaf1d;
  h->Lhalf = L;
  if (center) {
    h->ctr[0] = center[0];
    h->ctr[1] = center[1];
    h->ctr[2] = center[2];
  }
  h->nLevels = nlevels >= 0 ? nlevels : nlevels_formula(N, leaf1d);
  if (h->nLevels < 1 || h->nLevels > 7) {
    delete h;
    return nullptr;
  }
  h->tol_pow = tol_pow;
  h->tol = std::pow(10, -1.0 * tol_pow);  // src/ACA.cpp:5
  h->ks.q = qchoice;
  h->ks.scale = scale;
  h->ks.diag = diag;
  if (diag_vec) {
    h->diag_store.assign(diag_vec, diag_vec + N);
    h->ks.diag_vec = h->diag_store.data();
  }
  h->keep_uv = keep_uv != 0;
  h->x.resize(N);
  h->y.resize(N);
  h->z.resize(N);
  for (int64_t i = 0; i < N; ++i) {
    h->x[i] = xyz[3 * i];
    h->y[i] = xyz[3 * i + 1];
    h->z[i] = xyz[3 * i + 2];
  }
  build_tree(*h);
  return h;
}

void orc_destroy(void* p) { delete (Handle*)p; }

int orc_get_nlevels(void* p) { return ((Handle*)p)->nLevels; }

void orc_get_perm(void* p, int32_t* perm) {
  Handle* h = (Handle*)p;
  std::memcpy(perm, h->perm.data(), sizeof(int32_t) * h->N);
}

void orc_get_boxes(void* p, int level, int32_t* start, int32_t* count, double* centers) {
  Handle* h = (Handle*)p;
  int64_t nb = (int64_t)1 << (3 * level);
  for (int64_t k = 0; k < nb; ++k) {
    start[k] = h->bstart[level][k];
    count[k] = h->bcount[level][k];
    if (centers) {
      centers[3 * k] = h->cx[level][k];
      centers[3 * k + 1] = h->cy[level][k];
      centers[3 * k + 2] = h->cz[level][k];
    }
  }
}

void orc_get_neighbors(void* p, int level, int32_t* nbr26) {
  Handle* h = (Handle*)p;
  std::memcpy(nbr26, h->nbr[level].data(), sizeof(int32_t) * h->nbr[level].size());
}

int64_t orc_il_size(void* p, int level) { return (int64_t)((Handle*)p)->ilidx[level].size(); }

void orc_get_il(void* p, int level, int32_t* off, int32_t* idx) {
  Handle* h = (Handle*)p;
  std::memcpy(off, h->iloff[level].data(), sizeof(int32_t) * h->iloff[level].size());
  std::memcpy(idx, h->ilidx[level].data(), sizeof(int32_t) * h->ilidx[level].size());
}

// Assemble every block (src/HODLR3DTree.cpp:1621-1631).  Returns seconds.
double orc_assemble(void* p) {
  Handle* h = (Handle*)p;
  double t0 = omp_get_wtime();
  for (int j = 1; j <= h->nLevels; ++j) {
    std::vector<int32_t> t((size_t)1 << (3 * j));
    for (size_t k = 0; k < t.size(); ++k) t[k] = (int32_t)k;
    assemble_level(*h, j, t);
  }
  return omp_get_wtime() - t0;
}

// Assemble the IL blocks of the listed target boxes at one level.
double orc_assemble_boxes(void* p, int level, int64_t nbox, const int32_t* boxes) {
  Handle* h = (Handle*)p;
  double t0 = omp_get_wtime();
  std::vector<int32_t> t(boxes, boxes + nbox);
  assemble_level(*h, level, t);
  return omp_get_wtime() - t0;
}

// Per-block results of one level (CSR order of the IL): rank, status and the
// offset of each block's pivots / r x r factors in the flat arrays.
void orc_get_ranks(void* p, int level, int32_t* rank, int32_t* status) {
  Handle* h = (Handle*)p;
  for (size_t b = 0; b < h->blocks[level].size(); ++b) {
    rank[b] = h->blocks[level][b].rank;
    status[b] = h->blocks[level][b].status;
  }
}

// Local pivots (row within target box, col within source box), r each.
void orc_get_block(void* p, int level, int64_t b, int32_t* I, int32_t* J, double* L, double* R,
                   double* U, double* V) {
  Handle* h = (Handle*)p;
  const Block& B = h->blocks[level][b];
  int r = B.rank;
  for (int a = 0; a < r; ++a) {
    I[a] = B.I[a];
    J[a] = B.J[a];
  }
  if (L) std::memcpy(L, B.L.data(), sizeof(double) * r * r);
  if (R) std::memcpy(R, B.R.data(), sizeof(double) * r * r);
  if (U && !B.U.empty()) std::memcpy(U, B.U.data(), sizeof(double) * B.U.size());
  if (V && !B.V.empty()) std::memcpy(V, B.V.data(), sizeof(double) * B.V.size());
}

// Run ACA on explicit row / column lists of permuted positions (for tests).
int orc_aca_positions(void* p, int rs, int n, int cs, int m, int32_t* I, int32_t* J, double* U,
                      double* V) {
  Handle* h = (Handle*)p;
  Block B;
  bool keep = h->keep_uv;
  h->keep_uv = true;
  aca_block(*h, rs, n, cs, m, B);
  h->keep_uv = keep;
  for (int a = 0; a < B.rank; ++a) {
    I[a] = B.I[a];
    J[a] = B.J[a];
  }
  if (U && !B.U.empty()) std::memcpy(U, B.U.data(), sizeof(double) * B.U.size());
  if (V && !B.V.empty()) std::memcpy(V, B.V.data(), sizeof(double) * B.V.size());
  return B.rank;
}

// y = A q in original order.  mode 0: CUR with regeneration (reference form);
// mode 1: ACA vectors.  leafmask (may be null) restricts the near field.
void orc_matvec(void* p, const double* q, double* y, int mode, const uint8_t* leafmask) {
  Handle* h = (Handle*)p;
  matvec(*h, q, y, mode == 1, leafmask);
}

void orc_timers(void* p, double* assTime, double* matVecTime) {
  Handle* h = (Handle*)p;
  *assTime = h->assTime;
  *matVecTime = h->matVecTime;
}

// src/HODLR3DTree.cpp:1757-1778 (entries, not bytes) and :1807-1818.
double orc_find_memory(void* p) {
  Handle* h = (Handle*)p;
  double sum = 0;
  for (int j = 1; j <= h->nLevels; ++j) {
    int64_t nb = (int64_t)1 << (3 * j);
    for (int64_t k = 0; k < nb; ++k)
      for (int32_t b = h->iloff[j][k]; b < h->iloff[j][k + 1]; ++b) {
        int32_t ki = h->ilidx[j][b];
        double r = h->blocks[j][b].rank;
        sum += r * h->bcount[j][k];
        sum += r * r;
        sum += h->bcount[j][ki] * r;
      }
  }
  int nL = h->nLevels;
  int64_t nleaf = (int64_t)1 << (3 * nL);
  for (int64_t k = 0; k < nleaf; ++k) {
    double n = h->bcount[nL][k];
    sum += n * n;
    for (int s = 0; s < 26; ++s) {
      int32_t nn = h->nbr[nL][k * 26 + s];
      if (nn >= 0) sum += n * h->bcount[nL][nn];
    }
  }
  return sum;
}

int orc_max_rank(void* p) {
  Handle* h = (Handle*)p;
  int mr = 0;
  for (int j = 1; j <= h->nLevels; ++j)
    for (const Block& B : h->blocks[j]) mr = std::max(mr, B.rank);
  return mr;
}

int orc_gmres(void* p, const double* b, int maxit, double thr, double* x, double* err, int* its,
              double* hist) {
  return gmres(*(Handle*)p, b, maxit, thr, x, err, its, hist);
}

// Direct FP64 sum y_i = sum_j K(i,j) q_j over the listed rows (original
// indices), including the diagonal, with Neumaier compensated summation.
void orc_direct_rows(const double* xyz, int64_t N, int qchoice, double scale, double diag,
                     const double* diag_vec, const double* q, int64_t nrows, const int64_t* rows,
                     double* out) {
  KernelSpec ks;
  ks.q = qchoice;
  ks.scale = scale;
  ks.diag = diag;
  ks.diag_vec = diag_vec;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < nrows; ++t) {
    int64_t i = rows[t];
    double s = 0, c = 0;
    for (int64_t j = 0; j < N; ++j) {
      double v = entry(ks, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], xyz[3 * j],
                       xyz[3 * j + 1], xyz[3 * j + 2], i, j) * q[j];
      double tt = s + v;
      if (std::fabs(s) >= std::fabs(v))
        c += (s - tt) + v;
      else
        c += (v - tt) + s;
      s = tt;
    }
    out[t] = s + c;
  }
}

// Single entry (for kernel catalogue tests).
double orc_entry(int qchoice, double scale, double diag, const double* a, const double* b,
                 int64_t i, int64_t j) {
  KernelSpec ks;
  ks.q = qchoice;
  ks.scale = scale;
  ks.diag = diag;
  return entry(ks, a[0], a[1], a[2], b[0], b[1], b[2], i, j);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// Wall seconds of the reference-semantics far-field apply (entry regeneration
// + BLAS-2, src/HODLR3DTree.cpp:1660-1706) over the assembled target boxes of
// one level; q in original order.  Used to time a bounded sample.
double orc_time_far_level(void* p, int level, const double* q) {
  Handle* h = (Handle*)p;
  std::vector<double> qp(h->N), pot(h->N, 0.0);
  for (int64_t s = 0; s < h->N; ++s) qp[s] = q[h->perm[s]];
  std::vector<int32_t> boxes;
  for (size_t k = 0; k < h->assembled[level].size(); ++k)
    if (h->assembled[level][k]) boxes.push_back((int32_t)k);
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < (int64_t)boxes.size(); ++t) {
    double tg = 0, tm = 0;
    int32_t k = boxes[t];
    for (int32_t b = h->iloff[level][k]; b < h->iloff[level][k + 1]; ++b)
      far_block_apply(*h, level, k, h->blocks[level][b], qp.data(), pot.data(), tg, tm);
  }
  return omp_get_wtime() - t0;
}

// Wall seconds of the near field (:1708-1755) over the leaves in leafmask.
double orc_time_near(void* p, const double* q, const uint8_t* leafmask) {
  Handle* h = (Handle*)p;
  std::vector<double> qp(h->N), pot(h->N, 0.0);
  for (int64_t s = 0; s < h->N; ++s) qp[s] = q[h->perm[s]];
  double tg = 0, tm = 0;
  double t0 = omp_get_wtime();
  near_apply(*h, qp.data(), pot.data(), leafmask, tg, tm);
  return omp_get_wtime() - t0;
}

}  // extern "C"
