// Recompressed mixed-precision copy of the far field.
//
// For every low-rank block B = U V^T (ACA vectors, fp64 copy):
//   gram_kernel     Gu = U^T U, Gv = V^T V (compensated sums)
//   eig_kernel      Gu = Lu Lu^T, Gv = Lv Lv^T (Cholesky), one-sided Jacobi on
//                   M = Lu^T Lv: M J = Y with orthogonal columns, sorted by
//                   lam_k = ||Y_k||^2 descending; Tu = Lv J, Tv = Lv^-T J
//   project_kernel  U' = U Tu, V' = V Tv, so B = U' V'^T exactly (J orthogonal),
//                   V' has orthonormal columns and ||U'_k||^2 = lam_k
// A column k of a block then costs lam_k in ||.||_F^2 if dropped and about
// lam_k * eps_c^2 if stored in a format of relative precision eps_c.  One
// global budget tol^2 ||A||_F^2 (||A||_F^2 = sum of every block's lam plus the
// near field's sum of squares) is spent greedily on the columns with the
// smallest lam_k / (n + m) (error per stored element): the cheapest ones are
// dropped, the next ones stored in fp16, the rest in fp32.  For iid right-hand
// sides E||(A - A~) x|| / ||A x|| = ||A - A~||_F / ||A||_F, so `tol` is the
// expected relative matvec error added to the ACA error.
//
// Blocks whose Gu or Gv is numerically singular (or of rank above kMaxEigRank)
// are kept "raw": every ACA column in fp32, unscaled.
//
// The SVD step runs twice (once to collect every lam for the thresholds,
// once to write the copy) so that only O(rank) numbers per block are kept
// between the passes; both runs are deterministic and give the same bits.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"

namespace h3d {

namespace {

constexpr int kMaxEigRank = 142;          // larger blocks stay raw
// Expected relative Frobenius error of a stored column pair u v^T (both in
// the format): round-to-nearest leaves an error uniform in +-ulp/2, and over
// a binade E[ulp(x)^2 / x^2] = 2^-2p * 0.54, so ||du|| / ||u|| ~ 0.425 * 2^-p
// per side and sqrt(2) * 0.425 * 2^-p ~ 0.6 * 2^-p for the pair (p = 11 for
// fp16, 24 for fp32); the constants carry a 15% margin.
constexpr double kFp16Err = 3.4e-4;
constexpr double kFp32Err = 4.2e-8;
constexpr double kFp8Err = 4.3e-2;         // e4m3: p = 4
constexpr double kFp24Err = 1.05e-5;       // fp32 rounded to 16 significant bits: p = 16
// fp8 columns are stored 8x larger than the other classes' (rms entry 8, not
// 1): the format's normal range, [2^-6, 448], then holds entries from 1/500
// to 56 times a column's rms (larger ones are clamped to 448)
constexpr double kFp8Gain = 8.0;
constexpr int kClasses = kMixClasses;
constexpr int kClassElem[kClasses] = {4, 2, 1, 3};  // fp32, fp16, fp8, fp24
// a block's kept columns (sorted, largest first) by class: fp32, fp24, fp16, fp8
constexpr int kColumnOrder[kClasses] = {0, 3, 1, 2};
constexpr double kCholFail = 1e-12;       // pivot / original diagonal below this: raw block
constexpr int kMaxSweeps = 24;
constexpr int kBinsPerOctave = 8;
// int8 columns (kMixInt8): rho^2 = (max / rms)^2 of a stored column side,
// modelled as kI8Rho2 sqrt(rows).  Exact SVDs of 1/r blocks between
// well-separated jittered boxes of 64-2197 points give 1.10-1.19 sqrt(rows)
// for the columns of this tier; the constant carries a 15% margin.
constexpr double kI8Rho2 = 1.37;
constexpr int kSizeBuckets = 64;  // quarter octaves of sqrt(n) + sqrt(m): up to 2^16
constexpr double kColShare = 1.0 / 4096;  // largest share of the error budget one column may take
constexpr int64_t kMixItemBytes = 131072;           // factor bytes of a mixed phase-1 item
constexpr int64_t kVChunkElems = int64_t(1) << 28;  // elements of a V chunk (whole source boxes)

// Per-block work description (pass 1 and pass 2).
struct CDesc {
  int64_t xu;     // fp64 element offset of U_b's first column in its chunk
  int64_t xv;     // fp64 element offset of V_b in its chunk
  int64_t w;      // workspace offset: Gu/Tu (r*r), Gv/Tv (r*r), lam (r), status, floor
  int64_t j;      // offset of the block's stored J (sorted columns, r*r), -1: none
  int64_t lamo;   // offset of the block's lam in the batch's lam array (second pass)
  int32_t n, m, r, ldu, ldv, chunk;
};

// Per-block output description (pass 2).
struct PDesc {
  int64_t u_out[kClasses];  // class element offset of the block's first U column (in its U chunk)
  int64_t v_out[kClasses];  // class element offset of the block's V (in its V chunk)
  int32_t vch[kClasses];    // V chunk of the block per class
  int32_t rc[kClasses];     // columns kept per class (consecutive runs of the sorted
                            // columns, in kColumnOrder)
  int32_t raw;
  int32_t ldu_o[kClasses];  // U column stride per class
  int32_t ldv_o[kClasses];  // V column stride per class
  int32_t t8;               // class 2's t index of the block's first column (int8 maxima)
  double su, sv;            // power-of-two scales of the stored columns
};

// One project work item: rows [row0, row0 + 256) of one side of one block.
struct PItem {
  int32_t blk;   // index into the batch's CDesc / PDesc
  int32_t side;  // 0: U, 1: V
  int32_t row0;
  int32_t pad;
};

// Runs f(chunk, i0, i1) over [0, n) split into kHostChunks fixed chunks, on up
// to hardware_concurrency host threads.  The chunks do not depend on the thread
// count, so per-chunk partial results combined in chunk order are the same on
// every machine.
constexpr int kHostChunks = 64;
template <class F>
void host_chunks(int64_t n, F f) {
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto run = [&](int t) {
    for (int c = t; c < kHostChunks; c += nt) f(c, n * c / kHostChunks, n * (c + 1) / kHostChunks);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(run, t);
  run(0);
  for (auto& x : th) x.join();
}

// ------------------------------------------------------------------ Gram
// One workgroup per (block, side): G = X^T X for the n x r (or m x r) column
// block X, row tiles of 32 staged in LDS, each thread owning up to EMAX of
// the r(r+1)/2 entries (l <= l'); the tile sums are accumulated with Kahan
// compensation, so long columns add no rounding beyond a 32-term sum.
// sides = 2: both (workgroup 2 i + side); sides = 1: the V side only.
template <int EMAX, class S>
__global__ __launch_bounds__(256) void gram_kernel(const CDesc* __restrict__ ds,
                                                   void* const* __restrict__ base64,
                                                   const int32_t* __restrict__ list, int nlist,
                                                   double* __restrict__ ws, int sides) {
  constexpr int TR = 32;
  extern __shared__ double tile[];  // r x (TR + 1)
  const int item = sides == 2 ? (int)(blockIdx.x >> 1) : (int)blockIdx.x;
  if (item >= nlist) return;
  const int side = sides == 2 ? (int)(blockIdx.x & 1) : 1;
  const CDesc d = ds[list[item]];
  const int r = d.r;
  const int rows = side ? d.m : d.n;
  const int ld = side ? d.ldv : d.ldu;
  const S* X = (const S*)base64[d.chunk] + (side ? d.xv : d.xu);
  const int E = r * (r + 1) / 2;
  int el[EMAX], er[EMAX];
  double acc[EMAX], cmp[EMAX];  // sums with Kahan compensation over the row tiles
#pragma unroll
  for (int q = 0; q < EMAX; ++q) {
    cmp[q] = 0.0;
    const int e = threadIdx.x + 256 * q;
    int b = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (b * (b + 1) / 2 > e) --b;
    while ((b + 1) * (b + 2) / 2 <= e) ++b;
    el[q] = e - b * (b + 1) / 2;  // l <= l' = b
    er[q] = b;
    acc[q] = 0.0;
  }
  for (int r0 = 0; r0 < rows; r0 += TR) {
    const int nr = min(TR, rows - r0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < TR * r; idx += 256) {
      const int l = idx / TR, i = idx - l * TR;
      tile[l * (TR + 1) + i] = i < nr ? (double)X[(int64_t)l * ld + r0 + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EMAX; ++q) {
      if (threadIdx.x + 256 * q < E) {
        const double* a = tile + el[q] * (TR + 1);
        const double* b = tile + er[q] * (TR + 1);
        double s = 0.0;
        for (int i = 0; i < TR; ++i) s += a[i] * b[i];
        const double y = s - cmp[q];
        const double t = acc[q] + y;
        cmp[q] = (t - acc[q]) - y;
        acc[q] = t;
      }
    }
  }
  double* G = ws + d.w + (side ? (int64_t)r * r : 0);
#pragma unroll
  for (int q = 0; q < EMAX; ++q) {
    if (threadIdx.x + 256 * q < E) {
      G[el[q] + er[q] * r] = acc[q];
      G[er[q] + el[q] * r] = acc[q];
    }
  }
}

// ------------------------------------------------------------------ SVD
// One workgroup per block.  With Gv = Lv Lv^T and Gu = Lu Lu^T (Cholesky),
// U = Qu Lu^T and V = Qv Lv^T for orthonormal Qu, Qv, so B = Qu M Qv^T with the
// r x r matrix M = Lu^T Lv.  One-sided (Hestenes) Jacobi rotates M's columns
// until they are orthogonal, M J = Y; then B = (U Lv J)(V Lv^-T J)^T with the
// second factor orthonormal and ||U Lv J_k|| = ||Y_k||.  Working on M rather
// than on S = M^T M keeps the small singular values to relative accuracy
// (M's columns are graded like the ACA vectors), which a two-sided eigen
// solver on S would only resolve down to about sqrt(r eps) of the largest.
// L (= Lv), A (Lu, then J) and M live in LDS, or in global scratch (GLOBAL,
// 3 * gstride doubles per work item) for ranks whose matrices do not fit;
// per-pair rotation data is always in LDS.  All matrices row-major with the
// odd row stride r | 1 (a column read by consecutive lanes then hits distinct
// LDS banks; with an even stride they all fell on a few).
__device__ inline void wg_sync() { __syncthreads(); }

// In-place right-looking Cholesky of the r x r matrix A (lower triangle used);
// sets *flag when a pivot falls below kCholFail times the original diagonal.
// Called by every thread of the workgroup.
// Element (i, j) of an r x r work matrix with stride ld: row-major in LDS
// (odd ld: a column read by consecutive lanes hits distinct banks), column-
// major in global scratch (a column is contiguous, so the column reads and
// rotations coalesce).
template <bool CM>
__device__ inline int ix(int i, int j, int ld) {
  return CM ? j * ld + i : i * ld + j;
}

template <bool CM>
__device__ inline void chol_inplace(double* A, int r, int ld, double* dg, int* flag, int tid, int nt) {
  for (int i = tid; i < r; i += nt) dg[i] = A[ix<CM>(i, i, ld)];
  wg_sync();
  for (int k = 0; k < r; ++k) {
    const double dkk = A[ix<CM>(k, k, ld)];
    if (!(dkk > kCholFail * dg[k])) {  // also catches NaN
      if (tid == 0) *flag = 1;
      break;
    }
    const double lkk = sqrt(dkk);
    wg_sync();
    for (int i = k + 1 + tid; i < r; i += nt) A[ix<CM>(i, k, ld)] /= lkk;
    if (tid == 0) A[ix<CM>(k, k, ld)] = lkk;
    wg_sync();
    const int m = r - k - 1;
    for (int idx = tid; idx < m * m; idx += nt) {
      // consecutive threads take consecutive i (a column) in the column-major layout
      const int i = k + 1 + (CM ? idx % m : idx / m), j = k + 1 + (CM ? idx / m : idx % m);
      if (j <= i) A[ix<CM>(i, j, ld)] -= A[ix<CM>(i, k, ld)] * A[ix<CM>(j, k, ld)];
    }
    wg_sync();
  }
  wg_sync();
}

// VEC = false (the first pass, which needs the singular values only): J is not
// accumulated and no projections are formed; M's rotations, and so the values,
// are the same bit for bit as with VEC = true.  STOREJ (with VEC): the sorted
// columns of J go to jbuf + d.j instead of the projections, for finish_kernel.
template <bool GLOBAL, bool VEC, bool STOREJ = false>
__global__ __launch_bounds__(256) void eig_kernel(const CDesc* __restrict__ ds,
                                                  const int32_t* __restrict__ list, int nlist,
                                                  double* __restrict__ ws,
                                                  double* __restrict__ gscratch,
                                                  int64_t gstride, double* __restrict__ jbuf) {
  extern __shared__ double sm[];
  const int item = blockIdx.x;
  if (item >= nlist) return;
  const CDesc d = ds[list[item]];
  const int r = d.r;
  const int rr = r + (r & 1);
  // LDS: odd row stride (distinct banks); global: column-major with the same stride
  const int ld = r | 1;
  auto at = [&](int i, int j) { return ix<GLOBAL>(i, j, ld); };
  double* L = GLOBAL ? gscratch + 3 * gstride * item : sm;
  double* A = L + r * ld;
  double* M = A + r * ld;
  double* cs = GLOBAL ? sm : M + r * ld;  // rr/2
  double* sn = cs + rr / 2;      // rr/2
  double* dg = sn + rr / 2;      // r: original diagonal of the matrix being factored
  double* lam = dg + r;          // r
  int* pp = (int*)(lam + r);     // rr/2
  int* pq = pp + rr / 2;         // rr/2
  int* perm = pq + rr / 2;       // r
  __shared__ int flag;
  double* W = ws + d.w;
  const int tid = threadIdx.x, nt = blockDim.x;

  for (int idx = tid; idx < r * r; idx += nt) {
    const int i = idx / r, j = idx % r;
    L[at(i, j)] = W[(int64_t)r * r + idx];  // Gv (symmetric: row/column major alike)
    A[at(i, j)] = W[idx];                   // Gu
  }
  if (tid == 0) flag = 0;
  wg_sync();
  chol_inplace<GLOBAL>(L, r, ld, dg, &flag, tid, nt);
  if (!flag) chol_inplace<GLOBAL>(A, r, ld, dg, &flag, tid, nt);
  if (flag) {
    if (tid == 0) W[2 * (int64_t)r * r + r] = 1.0;  // raw
    return;
  }
  // M = Lu^T Lv: M[i][j] = sum_{k >= max(i,j)} Lu[k][i] Lv[k][j]
  for (int idx = tid; idx < r * r; idx += nt) {
    const int i = GLOBAL ? idx % r : idx / r, j = GLOBAL ? idx / r : idx % r;
    double s = 0.0;
    for (int k = max(i, j); k < r; ++k) s += A[at(k, i)] * L[at(k, j)];
    M[at(i, j)] = s;
  }
  wg_sync();
  if (VEC)
    for (int idx = tid; idx < r * r; idx += nt) A[at(idx / r, idx % r)] = (idx / r == idx % r) ? 1.0 : 0.0;  // J
  wg_sync();
  // one-sided Jacobi on M's columns, round-robin pairing (index 0 fixed).  A
  // pair's column dot products are split over tpp lanes of one wave (row i to
  // lane i mod tpp, then an xor butterfly: a fixed order), so that the rounds
  // do not wait on r / 2 threads each running an r-long loop.
  const int np = rr / 2;
  int tpp = 1;
  while (tpp < 64 && 2 * tpp * np <= nt) tpp <<= 1;
  for (int sweep = 0; sweep < kMaxSweeps && r > 1; ++sweep) {
    if (tid == 0) flag = 0;
    wg_sync();
    for (int round = 0; round < rr - 1; ++round) {
      for (int t0 = 0; t0 < np; t0 += nt / tpp) {
        const int t = t0 + tid / tpp, sub = tid % tpp;
        const bool act = t < np;
        auto pos = [&](int i) { return i == 0 ? 0 : 1 + (i - 1 + round) % (rr - 1); };
        int p = 0, q = r;
        if (act) {
          p = pos(t);
          q = pos(rr - 1 - t);
          if (p > q) {
            const int x = p;
            p = q;
            q = x;
          }
        }
        double a = 0.0, b = 0.0, g = 0.0;
        if (act && q < r)
          for (int i = sub; i < r; i += tpp) {
            const double mp = M[at(i, p)], mq = M[at(i, q)];
            a += mp * mp;
            b += mq * mq;
            g += mp * mq;
          }
        for (int dd = tpp >> 1; dd >= 1; dd >>= 1) {  // every lane of the wave takes part
          a += __shfl_xor(a, dd);
          b += __shfl_xor(b, dd);
          g += __shfl_xor(g, dd);
        }
        if (act && sub == 0) {
          double c = 1.0, s = 0.0;
          if (q < r && g != 0.0 && fabs(g) > 1e-15 * sqrt(a * b)) {
            const double th = (b - a) / (2.0 * g);
            const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
            flag = 1;
          }
          cs[t] = c;
          sn[t] = s;
          pp[t] = p;
          pq[t] = q < r ? q : -1;
        }
      }
      wg_sync();
      // columns p, q of M and J: (x_p, x_q) <- (c x_p - s x_q, s x_p + c x_q)
      for (int idx = tid; idx < np * r; idx += nt) {
        const int t = idx / r, i = idx % r;
        const int p = pp[t], q = pq[t];
        if (q < 0 || sn[t] == 0.0) continue;
        const double c = cs[t], s = sn[t];
        const double a = M[at(i, p)], b = M[at(i, q)];
        M[at(i, p)] = c * a - s * b;
        M[at(i, q)] = s * a + c * b;
        if (VEC) {
          const double ja = A[at(i, p)], jb = A[at(i, q)];
          A[at(i, p)] = c * ja - s * jb;
          A[at(i, q)] = s * ja + c * jb;
        }
      }
      wg_sync();
    }
    const int any = flag;
    wg_sync();
    if (!any) break;
  }
  // squared singular values = squared column norms of M J, sorted descending
  for (int k = tid; k < r; k += nt) {
    double s = 0.0;
    for (int i = 0; i < r; ++i) s += M[at(i, k)] * M[at(i, k)];
    lam[k] = s;
  }
  wg_sync();
  for (int i = tid; i < r; i += nt) {
    int rank = 0;
    for (int j = 0; j < r; ++j) rank += (lam[j] > lam[i]) || (lam[j] == lam[i] && j < i);
    perm[rank] = i;
  }
  wg_sync();
  double tr = 0.0;
  for (int i = 0; i < r; ++i) tr += lam[i];
  // rounding floor of the computed values (a few eps * r of the total): the
  // stored columns are scaled by sqrt(max(lam, floor)), so that a value below
  // it cannot blow up a stored fp16 column
  const double noise = 4.0 * r * 2.220446049250313e-16 * tr;
  // Tu[:, k] = Lv J[:, perm k], Tv[:, k] = Lv^-T J[:, perm k] (column-major, ld r)
  double* out_lam = W + 2 * (int64_t)r * r;
  for (int k = tid; k < r; k += nt) {
    const int c = perm[k];
    out_lam[k] = lam[c];
    if (!VEC) continue;
    if (STOREJ) {
      double* jo = jbuf + d.j + (int64_t)k * r;
      for (int i = 0; i < r; ++i) jo[i] = A[at(i, c)];
      continue;
    }
    double* tu = W + (int64_t)k * r;
    double* tv = W + (int64_t)r * r + (int64_t)k * r;
    for (int i = 0; i < r; ++i) {
      double s = 0.0;
      for (int j = 0; j <= i; ++j) s += L[at(i, j)] * A[at(j, c)];
      tu[i] = s;
    }
    for (int i = r - 1; i >= 0; --i) {
      double s = A[at(i, c)];
      for (int j = i + 1; j < r; ++j) s -= L[at(j, i)] * tv[j];
      tv[i] = s / L[at(i, i)];
    }
  }
  if (tid == 0) {
    W[2 * (int64_t)r * r + r] = 0.0;
    W[2 * (int64_t)r * r + r + 1] = noise;
  }
}

// The same eigen step for ranks up to G (16 or 32) with every matrix in
// registers: one wave holds 64 / G blocks, one per group of G lanes, and lane
// c of a group holds column c of the group's matrices (rows beyond r, and the
// columns c >= r, are zero).  The Choleskys broadcast the pivot column k with
// shuffles inside the group; M = Lu^T Lv reads Lu from the group's LDS slice,
// which then holds Lv for the projections of the second pass.  The Jacobi
// rounds pair column c with column c ^ k for k = 1 .. G - 1, which for a power
// of two G meets every pair once per sweep; both lanes of a pair form the same
// dot products from the same values in the same order, so they agree on the
// rotation, and each updates its own column.  A sweep with no rotation leaves
// the columns unchanged, so a group whose block has converged sweeps on
// without effect while the other groups of its wave finish.  VEC = false
// computes the values only; its rotations of M are those of VEC = true bit for
// bit.  The workgroup is one wave.
template <int G, bool VEC, bool STOREJ = false>
__global__ __launch_bounds__(64, 4) void eig_wave_kernel(const CDesc* __restrict__ ds,
                                                      const int32_t* __restrict__ list,
                                                      int nlist, double* __restrict__ ws,
                                                      double* __restrict__ jbuf) {
  constexpr int NG = 64 / G;
  __shared__ double slice[NG][G * G];  // per group, column-major: Lu, then Lv
  const int lane = threadIdx.x;
  const int grp = lane / G, c = lane % G, gb = grp * G;
  const int item = blockIdx.x * NG + grp;
  const bool valid = item < nlist;
  CDesc d{};
  if (valid) d = ds[list[item]];
  const int r = valid ? d.r : 0;
  double* W = ws + d.w;
  double* L = slice[grp];

  // column c of the symmetric r x r matrix at W + off (zero outside it;
  // loads at clamped in-bounds indices, then masked)
  const int cc = c < r ? c : (r > 0 ? r - 1 : 0);
  auto load_sym = [&](int64_t off, double (&a)[G]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int ic = i < r ? i : (r > 0 ? r - 1 : 0);
      const double x = W[off + (int64_t)cc * r + ic];
      a[i] = (i < r && c < r) ? x : 0.0;
    }
  };
  // in-place Cholesky (lower triangle, column c in lane c); true when a pivot
  // falls below kCholFail times its original diagonal (the group agrees)
  auto chol = [&](double (&a)[G]) __attribute__((always_inline)) {
    double dg = 0.0;
#pragma unroll
    for (int i = 0; i < G; ++i) dg = i == c ? a[i] : dg;
    bool fail = false;
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const double dkk = __shfl(a[k], gb + k);
      const double dgk = __shfl(dg, gb + k);
      if (k < r && !(dkk > kCholFail * dgk)) fail = true;
      const bool act = k < r && !fail;
      const double lkk = act ? sqrt(dkk) : 1.0;
      double lk[G];
      double lck = 0.0;
#pragma unroll
      for (int i = k + 1; i < G; ++i) {
        lk[i] = __shfl(a[i], gb + k) / lkk;
        lck = i == c ? lk[i] : lck;
      }
      if (act) {
        if (c == k) {
          a[k] = lkk;
#pragma unroll
          for (int i = k + 1; i < G; ++i) a[i] = lk[i];
        } else if (c > k) {
#pragma unroll
          for (int i = k + 1; i < G; ++i)
            if (i >= c) a[i] -= lk[i] * lck;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < G; ++i) a[i] = i < c ? 0.0 : a[i];
    return fail;
  };

  bool raw;
  {
    double u[G];
    load_sym(0, u);              // Gu
    raw = chol(u);               // Lu, to the group's slice
#pragma unroll
    for (int i = 0; i < G; ++i) L[i + c * G] = u[i];
  }
  double v[G], m[G];
  load_sym((int64_t)r * r, v);  // Gv
  raw = chol(v) || raw;          // Lv
  __syncthreads();
  // M[i][c] = sum_k Lu[k][i] Lv[k][c] (k increasing; the zeros of the
  // triangles add nothing)
#pragma unroll
  for (int i = 0; i < G; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < G; ++k) s += L[k + i * G] * v[k];
    m[i] = raw ? 0.0 : s;
  }
  __syncthreads();
  double j[VEC ? G : 1];
  if constexpr (VEC) {
#pragma unroll
    for (int i = 0; i < G; ++i) L[i + c * G] = v[i];  // Lv, for the projections
#pragma unroll
    for (int i = 0; i < G; ++i) j[i] = i == c ? 1.0 : 0.0;
  }
  // one-sided Jacobi on the columns of M
  for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
    bool rot = false;
    for (int k = 1; k < G; ++k) {
      const int p = c ^ k;
      const bool lower = c < p;
      double so = 0.0, sp = 0.0, g = 0.0;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const double y = __shfl_xor(m[i], k);
        so += m[i] * m[i];
        sp += y * y;
        g += m[i] * y;
      }
      const double a = lower ? so : sp, b = lower ? sp : so;
      double cs = 1.0, sn = 0.0;
      if ((c > p ? c : p) < r && g != 0.0 && fabs(g) > 1e-15 * sqrt(a * b)) {
        const double th = (b - a) / (2.0 * g);
        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
        cs = 1.0 / sqrt(1.0 + tt * tt);
        sn = tt * cs;
        rot = true;
      }
      // (x_lo, x_hi) <- (c x_lo - s x_hi, s x_lo + c x_hi); the shuffle reads
      // the partner's element before either lane updates it
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const double y = __shfl_xor(m[i], k);
        m[i] = lower ? cs * m[i] - sn * y : sn * y + cs * m[i];
      }
      if constexpr (VEC) {
#pragma unroll
        for (int i = 0; i < G; ++i) {
          const double y = __shfl_xor(j[i], k);
          j[i] = lower ? cs * j[i] - sn * y : sn * y + cs * j[i];
        }
      }
    }
    if (!__any(rot)) break;
  }
  // squared singular values = squared column norms of M J, sorted descending
  double lam = 0.0;
#pragma unroll
  for (int i = 0; i < G; ++i) lam += m[i] * m[i];
  int rank = 0;
  double tr = 0.0;
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const double lq = __shfl(lam, gb + q);
    if (q < r) {
      rank += (lq > lam) || (lq == lam && q < c);
      tr += lq;
    }
  }
  if (!valid) return;
  double* out_lam = W + 2 * (int64_t)r * r;
  if (raw) {
    if (c == 0) out_lam[r] = 1.0;
    return;
  }
  if (c < r) out_lam[rank] = lam;
  if (c == 0) {
    out_lam[r] = 0.0;
    out_lam[r + 1] = 4.0 * r * 2.220446049250313e-16 * tr;
  }
  if constexpr (VEC && STOREJ) {
    double* jo = jbuf + d.j + (int64_t)rank * r;
#pragma unroll
    for (int i = 0; i < G; ++i)
      if (c < r && i < r) jo[i] = j[i];
  } else if constexpr (VEC) {
    __builtin_amdgcn_wave_barrier();
    // Tu[:, rank] = Lv J[:, c], Tv[:, rank] = Lv^-T J[:, c] (column-major, ld r)
    double* tu = W + (int64_t)rank * r;
    double* tv = W + (int64_t)r * r + (int64_t)rank * r;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q <= i; ++q) s += L[i + q * G] * j[q];
      if (c < r && i < r) tu[i] = s;
    }
    double t[G];
#pragma unroll
    for (int i = G - 1; i >= 0; --i) {
      double s = j[i];
#pragma unroll
      for (int q = i + 1; q < G; ++q) s -= L[q + i * G] * t[q];
      t[i] = i < r ? s / L[i + i * G] : 0.0;
      if (c < r && i < r) tv[i] = t[i];
    }
  }
}

// Second pass of the single-Jacobi build: Tu[:, k] = Lv J[:, k] and
// Tv[:, k] = Lv^-T J[:, k] from the stored sorted columns of J and the
// Cholesky factor of Gv (recomputed from the same Gram matrix, so the same
// bits as in the first pass), plus the block's lam and floor for the
// projection.  Lv in LDS, or in global scratch (gstride doubles per item).
template <bool GLOBAL>
__global__ __launch_bounds__(256) void finish_kernel(const CDesc* __restrict__ ds,
                                                     const int32_t* __restrict__ list, int nlist,
                                                     double* __restrict__ ws,
                                                     const double* __restrict__ jbuf,
                                                     const double* __restrict__ lam_in,
                                                     const double* __restrict__ floor_in,
                                                     double* __restrict__ gscratch,
                                                     int64_t gstride) {
  extern __shared__ double sm[];
  const int item = blockIdx.x;
  if (item >= nlist) return;
  const int li = list[item];
  const CDesc d = ds[li];
  const int r = d.r;
  const int ld = r | 1;
  auto at = [&](int i, int j) { return ix<GLOBAL>(i, j, ld); };
  double* L = GLOBAL ? gscratch + gstride * item : sm;
  double* dg = GLOBAL ? sm : L + r * ld;  // r
  __shared__ int flag;
  double* W = ws + d.w;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int idx = tid; idx < r * r; idx += nt) L[at(idx / r, idx % r)] = W[(int64_t)r * r + idx];
  if (tid == 0) flag = 0;
  wg_sync();
  chol_inplace<GLOBAL>(L, r, ld, dg, &flag, tid, nt);
  double* out_lam = W + 2 * (int64_t)r * r;
  for (int k = tid; k < r; k += nt) out_lam[k] = lam_in[d.lamo + k];
  if (tid == 0) {
    out_lam[r] = 0.0;
    out_lam[r + 1] = floor_in[li];
  }
  const double* J = jbuf + d.j;
  for (int k = tid; k < r; k += nt) {
    const double* jk = J + (int64_t)k * r;
    double* tu = W + (int64_t)k * r;
    double* tv = W + (int64_t)r * r + (int64_t)k * r;
    for (int i = 0; i < r; ++i) {
      double s = 0.0;
      for (int q = 0; q <= i; ++q) s += L[at(i, q)] * jk[q];
      tu[i] = s;
    }
    for (int i = r - 1; i >= 0; --i) {
      double s = jk[i];
      for (int q = i + 1; q < r; ++q) s -= L[at(q, i)] * tv[q];
      tv[i] = s / L[at(i, i)];
    }
  }
}

// ------------------------------------------------------------------ project
// Per class, the base pointers of its U chunks and V chunks.
struct ClassBases {
  void* const* u[kClasses];
  void* const* v[kClasses];
};

// Thread per row: out_k = sum_l X[row, l] T[l, k] for the kept columns, scaled
// and stored in the column's class (the sorted columns: rc[0] in fp32, then
// rc[3] in fp24, rc[1] in fp16, rc[2] in fp8 or int8).  Raw blocks copy their
// columns unscaled into fp32.  MODE 0 writes class 2 in e4m3; MODE 1 only
// takes the largest |entry| of each class-2 column side (a wave maximum, then
// an atomic maximum of its fp32 bits, which order as the values do for x >=
// 0) into mx8[side * T2 + t8 + column]; MODE 2 writes class 2 in int8,
// round(127 v / max).  Both int8 passes compute the same values.
template <class S, int MODE>
__global__ __launch_bounds__(256) void project_kernel(
    const PItem* __restrict__ items, const CDesc* __restrict__ ds, const PDesc* __restrict__ ps,
    void* const* __restrict__ base64, const double* __restrict__ ws, ClassBases cb,
    uint32_t* __restrict__ mx8, int64_t T2) {
  constexpr int KT = 8;
  const PItem it = items[blockIdx.x];
  const CDesc d = ds[it.blk];
  const PDesc p = ps[it.blk];
  const int side = it.side;
  const int rows = side ? d.m : d.n;
  const int row = it.row0 + threadIdx.x;
  if constexpr (MODE == 1) {
    // every lane of the wave takes part in the maxima (lanes past the rows add 0)
    if (p.raw || p.rc[2] == 0) return;
    const bool act = row < rows;
    const int r = d.r;
    const int ld = side ? d.ldv : d.ldu;
    const S* X = (const S*)base64[d.chunk] + (side ? d.xv : d.xu) + (act ? row : 0);
    const double* W = ws + d.w;
    const double* T = W + (side ? (int64_t)r * r : 0);
    const double* lam = W + 2 * (int64_t)r * r;
    const double floor_ = lam[r + 1];
    const int k8 = p.rc[0] + p.rc[3] + p.rc[1], keep = k8 + p.rc[2];
    const double sc = side ? p.sv : p.su;
    uint32_t* out = mx8 + (side ? T2 : 0) + p.t8 - k8;
    for (int k0 = k8; k0 < keep; k0 += KT) {
      double acc[KT];
#pragma unroll
      for (int q = 0; q < KT; ++q) acc[q] = 0.0;
      if (act)
        for (int l = 0; l < r; ++l) {
          const double x = (double)X[(int64_t)l * ld];
#pragma unroll
          for (int q = 0; q < KT; ++q)
            if (k0 + q < keep) acc[q] += x * T[l + (int64_t)(k0 + q) * r];
        }
#pragma unroll
      for (int q = 0; q < KT; ++q) {
        const int k = k0 + q;
        if (k >= keep) break;
        const double v = side ? acc[q] * sc : acc[q] * (sc / sqrt(fmax(lam[k], floor_)));
        float a = act ? fabsf((float)v) : 0.0f;
        for (int o = 32; o >= 1; o >>= 1) a = fmaxf(a, __shfl_xor(a, o));
        if ((threadIdx.x & 63) == 0 && a > 0.0f) atomicMax(out + k, __float_as_uint(a));
      }
    }
    return;
  }
  if (row >= rows) return;
  const int r = d.r;
  const int ld = side ? d.ldv : d.ldu;  // the source copy's stride
  const int* ldo = side ? p.ldv_o : p.ldu_o;
  const S* X = (const S*)base64[d.chunk] + (side ? d.xv : d.xu) + row;
  const double* W = ws + d.w;
  const double* T = W + (side ? (int64_t)r * r : 0);
  const double* lam = W + 2 * (int64_t)r * r;
  const double floor_ = lam[r + 1];
  auto out = [&](int c) -> uint8_t* {
    if (p.rc[c] == 0) return nullptr;
    return (uint8_t*)(side ? cb.v[c][p.vch[c]] : cb.u[c][d.chunk]) +
           ((side ? p.v_out[c] : p.u_out[c]) + row) * kClassElem[c];
  };
  float* o32 = (float*)out(0);
  _Float16* o16 = (_Float16*)out(1);
  uint8_t* o8 = out(2);
  uint8_t* o24 = out(3);
  if (p.raw) {
    for (int k = 0; k < p.rc[0]; ++k) o32[(int64_t)k * ldo[0]] = (float)X[(int64_t)k * ld];  // S -> fp32
    return;
  }
  // first sorted column of each class
  const int k24 = p.rc[0], k16 = k24 + p.rc[3], k8 = k16 + p.rc[1], keep = k8 + p.rc[2];
  const double sc = side ? p.sv : p.su;
  for (int k0 = 0; k0 < keep; k0 += KT) {
    double acc[KT];
#pragma unroll
    for (int q = 0; q < KT; ++q) acc[q] = 0.0;
    for (int l = 0; l < r; ++l) {
      const double x = (double)X[(int64_t)l * ld];
#pragma unroll
      for (int q = 0; q < KT; ++q)
        if (k0 + q < keep) acc[q] += x * T[l + (int64_t)(k0 + q) * r];
    }
#pragma unroll
    for (int q = 0; q < KT; ++q) {
      const int k = k0 + q;
      if (k >= keep) break;
      // U side: unit-norm column times su; V side: V' columns are orthonormal
      const double v = side ? acc[q] * sc : acc[q] * (sc / sqrt(fmax(lam[k], floor_)));
      if (k < k24) {
        o32[(int64_t)k * ldo[0]] = (float)v;
      } else if (k < k16) {
        // fp32 rounded to 16 significant bits (half away from zero), top 3 bytes kept
        const uint32_t u = __float_as_uint((float)v) + 0x80u;
        uint8_t* o = o24 + (int64_t)(k - k24) * ldo[3] * 3;
        o[0] = (uint8_t)(u >> 8);
        o[1] = (uint8_t)(u >> 16);
        o[2] = (uint8_t)(u >> 24);
      } else if (k < k8) {
        o16[(int64_t)(k - k16) * ldo[1]] = (_Float16)v;
      } else if constexpr (MODE == 2) {
        const float m = __uint_as_float(mx8[(side ? T2 : 0) + p.t8 + (k - k8)]);
        const float f = m > 0.0f ? rintf((float)v * (127.0f / m)) : 0.0f;
        o8[(int64_t)(k - k8) * ldo[2]] = (uint8_t)(int8_t)fminf(127.0f, fmaxf(-127.0f, f));
      } else {
        float f = (float)(v * kFp8Gain);
        if (fabsf(f) > 448.0f) f = copysignf(448.0f, f);  // NaN passes
        o8[(int64_t)(k - k8) * ldo[2]] =
            (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(f, 0.0f, 0, false) & 0xff);
      }
    }
  }
}

// Sum over near pairs (i != j) of K(i, j)^2, one partial per near item.
template <class KF>
__global__ __launch_bounds__(256) void near_fro2_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    KernelDev kd, double* __restrict__ out) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  const bool act = lane < it.nr;
  const int ip = rs + it.i0 + (act ? lane : 0);
  double acc = 0.0;
  if (act) {
    for (int sl = 0; sl <= 26; ++sl) {
      const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
      if (nn < 0) continue;
      const int cs = bstart[nn], m = bcount[nn];
      for (int jp = cs; jp < cs + m; ++jp) {
        if (jp == ip) continue;
        const double v = kd.scale * kpair<KF>(kd.q, px[ip], py[ip], pz[ip], px[jp], py[jp], pz[jp]);
        acc += v * v;
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) out[w] = acc;
}

// Packs every block's lam (r values), status and noise floor:
// out[off[i] + k] = lam_k, status[i] = raw flag, flo[i] = floor.
__global__ void pack_lam_kernel(const CDesc* __restrict__ ds, int nblk,
                                const int64_t* __restrict__ off, const double* __restrict__ ws,
                                double* __restrict__ out, uint8_t* __restrict__ status,
                                double* __restrict__ flo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblk) return;
  const CDesc d = ds[i];
  const double* W = ws + d.w + 2 * (int64_t)d.r * d.r;
  for (int k = 0; k < d.r; ++k) out[off[i] + k] = W[k];
  status[i] = W[d.r] != 0.0;
  flo[i] = W[d.r + 1];
}

// Runs f(c) for every class on a host thread of its own; the first exception
// of any of them is rethrown after all have ended.
template <class F>
void run_per_class(F&& f) {
  std::exception_ptr err[kClasses];
  std::vector<std::thread> th;
  for (int c = 0; c < kClasses; ++c)
    th.emplace_back([&, c] {
      try {
        f(c);
      } catch (...) {
        err[c] = std::current_exception();
      }
    });
  for (auto& t : th) t.join();
  for (auto& x : err)
    if (x) std::rethrow_exception(x);
}

inline double pow2_near_sqrt(int n) {
  return std::ldexp(1.0, (int)std::lround(0.5 * std::log2((double)std::max(n, 1))));
}

struct BlockRef {
  int j;       // level
  int32_t b;   // block (CSR index within the level)
  int32_t k;   // target box
};

}  // namespace

void FactorClass::free_all() {
  if (arena) (void)hipFree(arena);
  arena = nullptr;
  udata.clear();
  vdata.clear();
  nU.clear();
  nV.clear();
  dfree(d_ubase);
  dfree(d_vbase);
  dfree(d_blk);
  dfree(d_btoff);
  dfree(d_mp1);
  dfree(d_mp1r);
  dfree(d_tmap);
  dfree(d_scale32);
  dfree(d_tp2v);
  tp2v.clear();
  plan.free_device();
  plan = Plan();
  mp1.clear();
  mp1r.clear();
  tmap.clear();
  lf.clear();
  T = tbase = mpartial = 0;
  u_elems = v_elems = rank_sum = bytes = 0;
}

void MixedCopy::free_all() {
  for (FactorClass& c : cls) c.free_all();
  dfree(d_scale);
  ready = false;
}

void release_mixed(Engine& e) {
  H3D_HIP(hipStreamSynchronize(e.stream));
  release_mixed_groups(e);  // its tables point into the classes
  e.mix.free_all();
}

void build_mixed(Engine& e, double tol) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "compress before assemble");
  if (!(tol > 0.0) || !std::isfinite(tol)) throw Error(H3D_ERR_INVALID, "tol must be > 0");
  // source: the fp64 ACA copy, else the fp32 one (same element offsets; its
  // rounding, 2^-24 of each entry, is far below any useful budget)
  const int src = e.fac_bytes[0] ? 0 : (e.fac_bytes[1] ? 1 : -1);
  if (src < 0)
    throw Error(H3D_ERR_STATE, "the recompressed copy is built from the fp64 or the fp32 factors");
  auto t0 = std::chrono::steady_clock::now();
  release_mixed(e);
  MixedCopy& mx = e.mix;
  mx.tol = tol;
  hipStream_t st = e.stream;
  const int nL = e.nL;

  int dev = 0, lds_max = 65536;
  H3D_HIP(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  // gfx950 lets one workgroup take all 160 KB of a CU's LDS (with the
  // attribute set below): ranks up to ~80 keep their matrices there
  // (less a margin for the kernel's own static LDS)
  lds_max = std::min(lds_max, 163840) - 1024;
  // LDS of the eigen step: matrices (unless global) + rotation data
  auto eig_lds = [](int r, bool global) {
    const int rr = r + (r & 1);
    return (size_t)((global ? 0 : 3 * r * (r | 1)) + rr + 2 * r) * sizeof(double) +
           (size_t)(rr + r) * sizeof(int);
  };
  int rmax = 1;  // largest rank whose matrices fit in LDS
  while (rmax + 1 <= kMaxEigRank && eig_lds(rmax + 1, false) <= (size_t)lds_max) ++rmax;

  // every block in chunk order, with its work description
  std::vector<BlockRef> refs;
  std::vector<CDesc> cd;
  std::vector<int32_t> chunk_first(e.chunks.size() + 1, 0);
  {
    int cur = -1;
    for (int j = 1; j <= nL; ++j) {
      const Level& lv = e.lev[j];
      const int64_t lo = level_offset(j);
      for (int64_t k = 0; k < lv.nboxes; ++k) {
        const int n = e.bcount[lo + k];
        for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
          const int r = lv.rank[b];
          if (r == 0) continue;
          const int m = e.bcount[lo + lv.ilidx[b]];
          CDesc c{};
          c.n = n;
          c.m = m;
          c.r = r;
          c.ldu = aca_ldu(n);
          c.ldv = m + (m & 1);
          c.chunk = lv.chunk[k];
          c.xu = lv.uoff[k] + (int64_t)lv.col[b] * c.ldu;
          c.xv = e.chunks[c.chunk].nU + lv.voff[b];
          while (cur < c.chunk) chunk_first[++cur] = (int32_t)cd.size();
          refs.push_back({j, b, (int32_t)k});
          cd.push_back(c);
        }
      }
    }
    while (cur < (int)e.chunks.size()) chunk_first[++cur] = (int32_t)cd.size();
  }
  const int64_t nb = (int64_t)cd.size();
  // base pointers of the source chunks (whole chunk: U region then V region)
  std::vector<void*> b64(e.chunks.size());
  for (size_t c = 0; c < e.chunks.size(); ++c) b64[c] = e.chunks[c].data[src];
  void** d_b64 = upload(b64, st);

  // per block workspace size (doubles) and batches bounded by a budget
  auto wsize = [](int r) { return 2 * (int64_t)r * r + r + 2; };
  size_t freeb = 0, totb = 0;
  H3D_HIP(hipMemGetInfo(&freeb, &totb));
  const int64_t ws_budget =
      std::max<int64_t>(int64_t(64) << 20, std::min<int64_t>((int64_t)(freeb / 4), int64_t(8) << 30)) / 8;

  std::vector<double> lam_all;      // every column's lam (sorted per block), pass 1
  std::vector<int64_t> lam_off(nb + 1, 0);
  for (int64_t i = 0; i < nb; ++i) lam_off[i + 1] = lam_off[i] + cd[i].r;
  lam_all.assign(lam_off[nb], 0.0);
  std::vector<uint8_t> raw(nb, 0);
  std::vector<double> lam_floor(nb, 0.0);

  // Runs gram + the eigen step on blocks [i0, i1) (ws offsets assigned here)
  // in one of four modes: kValues (the singular values only), kVectors (and
  // the projections Tu, Tv), kStoreJ (values, and the sorted columns of J to
  // d_j), kFinish (Tu, Tv from the stored J, with lam and floor from the
  // first pass).  `sel` (optional) picks the blocks that need the step.
  enum { kValues, kVectors, kStoreJ, kFinish };
  double* d_ws = nullptr;
  int64_t ws_cap = 0;
  CDesc* d_cd = nullptr;
  int32_t* d_list = nullptr;
  double* d_j = nullptr;  // stored J of every block (single-Jacobi build), or null
  for (const void* f : {(const void*)eig_kernel<false, true>, (const void*)eig_kernel<false, false>,
                        (const void*)eig_kernel<false, true, true>, (const void*)finish_kernel<false>})
    H3D_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  auto run_eig = [&](int64_t i0, int64_t i1, int mode, const std::vector<uint8_t>* sel) {
    int64_t need = 0;
    for (int64_t i = i0; i < i1; ++i) {
      cd[i].w = need;
      cd[i].lamo = lam_off[i] - lam_off[i0];
      need += wsize(cd[i].r);
    }
    if (need > ws_cap) {
      dfree(d_ws);
      d_ws = dalloc<double>(need);
      ws_cap = need;
    }
    dfree(d_cd);
    std::vector<CDesc> sub(cd.begin() + i0, cd.begin() + i1);
    d_cd = upload(sub, st);
    // groups by rank class: gram EMAX; ranks above kMaxEigRank stay raw
    std::vector<int32_t> g4, g16, g40, eg;
    for (int64_t i = i0; i < i1; ++i) {
      const int r = cd[i].r;
      const int32_t li = (int32_t)(i - i0);
      if (r > kMaxEigRank || (sel && !(*sel)[i])) continue;
      const int E = r * (r + 1) / 2;
      (E <= 4 * 256 ? g4 : (E <= 16 * 256 ? g16 : g40)).push_back(li);
      eg.push_back(li);
    }
    std::vector<int32_t> all;
    all.insert(all.end(), g4.begin(), g4.end());
    all.insert(all.end(), g16.begin(), g16.end());
    all.insert(all.end(), g40.begin(), g40.end());
    dfree(d_list);
    d_list = upload(all, st);
    const int sides = mode == kFinish ? 1 : 2;  // the finish needs Gv only
    int64_t off = 0;
    auto launch_g = [&](std::vector<int32_t>& g, int which) {
      if (g.empty()) return;
      int rm = 1;
      for (int32_t li : g) rm = std::max(rm, cd[i0 + li].r);
      const size_t lds = sizeof(double) * 33 * (size_t)rm;
      for (int64_t s = 0; s < (int64_t)g.size(); s += 1000000) {
        const int cnt = (int)std::min<int64_t>(1000000, (int64_t)g.size() - s);
        const unsigned grid = (unsigned)sides * (unsigned)cnt;
        if (which == 0)
          (src ? gram_kernel<4, float> : gram_kernel<4, double>)<<<grid, 256, lds, st>>>(
              d_cd, d_b64, d_list + off + s, cnt, d_ws, sides);
        else if (which == 1)
          (src ? gram_kernel<16, float> : gram_kernel<16, double>)<<<grid, 256, lds, st>>>(
              d_cd, d_b64, d_list + off + s, cnt, d_ws, sides);
        else
          (src ? gram_kernel<40, float> : gram_kernel<40, double>)<<<grid, 256, lds, st>>>(
              d_cd, d_b64, d_list + off + s, cnt, d_ws, sides);
        H3D_LAUNCH_CHECK();
      }
      off += (int64_t)g.size();
    };
    launch_g(g4, 0);
    launch_g(g16, 1);
    launch_g(g40, 2);
    // the first pass's lam and floor of these blocks, for the finish
    double* d_lam = nullptr;
    double* d_flo = nullptr;
    if (mode == kFinish) {
      std::vector<double> lam(lam_all.begin() + lam_off[i0], lam_all.begin() + lam_off[i1]);
      std::vector<double> flo(lam_floor.begin() + i0, lam_floor.begin() + i1);
      d_lam = upload(lam, st);
      d_flo = upload(flo, st);
    }
    // eig: ranks up to 16 / 32 in registers (eig_wave_kernel, 4 / 2 blocks per
    // wave); larger ranks four waves per block, launched in classes of LDS
    // footprint (each launch sizes its LDS by its largest rank, and that sets
    // how many workgroups share a CU); ranks beyond rmax with their matrices
    // in global scratch.  The finish: one wave per block up to rank 32, four
    // above, Lv in LDS (global scratch where it does not fit).
    constexpr int kEigClasses = 7;  // 0, 1: registers; 2-5: LDS up to the caps; 6: global
    const size_t caps[4] = {40 << 10, 64 << 10, 96 << 10, (size_t)lds_max};
    auto finish_lds = [](int r, bool global) {
      return (size_t)((global ? 0 : r * (r | 1)) + r) * sizeof(double);
    };
    std::vector<int32_t> grp[kEigClasses];
    for (int32_t li : eg) {
      const int r = cd[i0 + li].r;
      int q = 6;
      if (mode == kFinish) q = r <= 32 ? 0 : (finish_lds(r, false) <= (size_t)lds_max ? 1 : 6);
      else if (r <= 16) q = 0;
      else if (r <= 32) q = 1;
      else if (r <= rmax)
        for (int c = 0; c < 4; ++c)
          if (eig_lds(r, false) <= caps[c]) {
            q = 2 + c;
            break;
          }
      grp[q].push_back(li);
    }
    const bool vec = mode != kValues, store = mode == kStoreJ;
    for (int q = 0; q < kEigClasses; ++q) {
      std::vector<int32_t>& g = grp[q];
      if (g.empty()) continue;
      const bool glob = q == kEigClasses - 1;
      int rm = 1;
      for (int32_t li : g) rm = std::max(rm, cd[i0 + li].r);
      int32_t* d_g = upload(g, st);
      const int64_t gstride = mode == kFinish ? (int64_t)rm * (rm | 1) : (int64_t)rm * (rm | 1);
      const int64_t mats = mode == kFinish ? 1 : 3;  // work matrices per item in global scratch
      const int64_t per = glob ? std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t(1) << 28) / (mats * gstride)))
                               : 1000000;
      double* gs = glob ? dalloc<double>(mats * gstride * std::min<int64_t>(per, (int64_t)g.size())) : nullptr;
      for (int64_t s = 0; s < (int64_t)g.size(); s += per) {
        const int cnt = (int)std::min<int64_t>(per, (int64_t)g.size() - s);
        if (mode == kFinish) {
          const int nt = q == 0 ? 64 : 256;
          if (glob)
            finish_kernel<true><<<cnt, nt, finish_lds(rm, true), st>>>(d_cd, d_g + s, cnt, d_ws, d_j,
                                                                        d_lam, d_flo, gs, gstride);
          else
            finish_kernel<false><<<cnt, nt, finish_lds(rm, false), st>>>(
                d_cd, d_g + s, cnt, d_ws, d_j, d_lam, d_flo, nullptr, 0);
        } else if (q == 0) {
          auto k = store ? eig_wave_kernel<16, true, true>
                         : (vec ? eig_wave_kernel<16, true> : eig_wave_kernel<16, false>);
          k<<<(cnt + 3) / 4, 64, 0, st>>>(d_cd, d_g + s, cnt, d_ws, d_j);
        } else if (q == 1) {
          auto k = store ? eig_wave_kernel<32, true, true>
                         : (vec ? eig_wave_kernel<32, true> : eig_wave_kernel<32, false>);
          k<<<(cnt + 1) / 2, 64, 0, st>>>(d_cd, d_g + s, cnt, d_ws, d_j);
        } else if (!glob) {
          auto k = store ? eig_kernel<false, true, true>
                         : (vec ? eig_kernel<false, true> : eig_kernel<false, false>);
          k<<<cnt, 256, eig_lds(rm, false), st>>>(d_cd, d_g + s, cnt, d_ws, nullptr, 0, d_j);
        } else {
          auto k = store ? eig_kernel<true, true, true>
                         : (vec ? eig_kernel<true, true> : eig_kernel<true, false>);
          k<<<cnt, 256, eig_lds(rm, true), st>>>(d_cd, d_g + s, cnt, d_ws, gs, gstride, d_j);
        }
        H3D_LAUNCH_CHECK();
      }
      H3D_HIP(hipStreamSynchronize(st));
      dfree(gs);
      dfree(d_g);
    }
    dfree(d_lam);
    dfree(d_flo);
  };

  // scratch freed on every exit, exceptions included
  struct Guard {
    double*& ws;
    CDesc*& cd;
    int32_t*& list;
    double*& j;
    void**& b64;
    ~Guard() {
      dfree(ws);
      dfree(cd);
      dfree(list);
      dfree(j);
      dfree(b64);
    }
  } guard{d_ws, d_cd, d_list, d_j, d_b64};

  // Single Jacobi: the first pass keeps every block's sorted J (r^2 doubles
  // each) so that the second pass forms the projections without repeating the
  // sweeps.  Only where it fits beside the copy being built; otherwise (and if
  // the copy's own allocation later needs the room) the second pass runs the
  // identical sweeps again with vectors.
  {
    int64_t jneed = 0;
    for (int64_t i = 0; i < nb; ++i) {
      cd[i].j = cd[i].r <= kMaxEigRank ? jneed : -1;
      if (cd[i].r <= kMaxEigRank) jneed += (int64_t)cd[i].r * cd[i].r;
    }
    size_t fb = 0, tb = 0;
    H3D_HIP(hipMemGetInfo(&fb, &tb));
    // room left for the batch workspace and the copy's classes (a third of
    // the source copy's bytes, well above what a useful budget keeps)
    int64_t src_bytes = 0;
    for (const Chunk& ch : e.chunks) src_bytes += (ch.nU + ch.nV) * elem_size(src);
    const int64_t reserve = 8 * ws_budget + src_bytes / 3 + (int64_t(2) << 30);
    if (jneed > 0 && (int64_t)fb - 8 * jneed > reserve) {
      try {
        d_j = dalloc<double>(jneed);
      } catch (const Error& er) {
        if (er.code != H3D_ERR_OOM) throw;
        d_j = nullptr;
      }
    }
    mx.single_jacobi = d_j != nullptr;
  }

  // ---------------- pass 1: eigenvalues of every block
  {
    int64_t i0 = 0;
    while (i0 < nb) {
      int64_t i1 = i0, need = 0;
      while (i1 < nb && (i1 == i0 || need + wsize(cd[i1].r) <= ws_budget)) need += wsize(cd[i1++].r);
      run_eig(i0, i1, d_j ? kStoreJ : kValues, nullptr);
      const int cnt = (int)(i1 - i0);
      std::vector<int64_t> off(cnt);
      for (int64_t i = i0; i < i1; ++i) off[i - i0] = lam_off[i] - lam_off[i0];
      const int64_t nlam = lam_off[i1] - lam_off[i0];
      int64_t* d_off = upload(off, st);
      double* d_lam = dalloc<double>(std::max<int64_t>(nlam, 1));
      uint8_t* d_stat = dalloc<uint8_t>(cnt);
      double* d_flo = dalloc<double>(cnt);
      pack_lam_kernel<<<(cnt + 255) / 256, 256, 0, st>>>(d_cd, cnt, d_off, d_ws, d_lam, d_stat,
                                                         d_flo);
      H3D_LAUNCH_CHECK();
      std::vector<uint8_t> stat(cnt);
      H3D_HIP(hipMemcpyAsync(lam_floor.data() + i0, d_flo, sizeof(double) * cnt,
                             hipMemcpyDeviceToHost, st));
      H3D_HIP(hipMemcpyAsync(lam_all.data() + lam_off[i0], d_lam, sizeof(double) * nlam,
                             hipMemcpyDeviceToHost, st));
      H3D_HIP(hipMemcpyAsync(stat.data(), d_stat, cnt, hipMemcpyDeviceToHost, st));
      H3D_HIP(hipStreamSynchronize(st));
      dfree(d_off);
      dfree(d_lam);
      dfree(d_stat);
      dfree(d_flo);
      for (int64_t i = i0; i < i1; ++i)
        raw[i] = cd[i].r > kMaxEigRank ? 1 : stat[i - i0];
      i0 = i1;
    }
  }
  // ||A||_F^2: far blocks (exact for the recompressed ones) + near field
  {
    mx.fro2_far = 0;
    std::vector<double> part(kHostChunks, 0.0);
    host_chunks(nb, [&](int c, int64_t i0, int64_t i1) {
      double acc = 0.0;
      for (int64_t i = i0; i < i1; ++i)
        if (!raw[i])
          for (int k = 0; k < cd[i].r; ++k) acc += lam_all[lam_off[i] + k];
      part[c] = acc;
    });
    for (double v : part) mx.fro2_far += v;
    const int nn = (int)e.nitems.size();
    mx.fro2_near = 0;
    if (nn) {
      double* d_out = dalloc<double>(nn);
      const int64_t lo = level_offset(nL);
      const unsigned g = (unsigned)((nn + 3) / 4);
      if (e.kern.q == 7 || e.kern.q == 16)
        near_fro2_kernel<InvR><<<g, 256, 0, st>>>(e.d_near, nn, e.lev[nL].d_nbr, e.d_bstart + lo,
                                                  e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, e.kern,
                                                  d_out);
      else if (e.kern.q == kQPlugin)
        near_fro2_kernel<Plugin><<<g, 256, 0, st>>>(e.d_near, nn, e.lev[nL].d_nbr,
                                                    e.d_bstart + lo, e.d_bcount + lo, e.d_px,
                                                    e.d_py, e.d_pz, e.kern, d_out);
      else
        near_fro2_kernel<AnyQ><<<g, 256, 0, st>>>(e.d_near, nn, e.lev[nL].d_nbr, e.d_bstart + lo,
                                                  e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, e.kern,
                                                  d_out);
      H3D_LAUNCH_CHECK();
      std::vector<double> h(nn);
      H3D_HIP(hipMemcpyAsync(h.data(), d_out, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
      H3D_HIP(hipStreamSynchronize(st));
      dfree(d_out);
      for (double v : h) mx.fro2_near += v;
    }
  }
  // Each kept column goes to one tier t of (drop, fp8, fp16, fp24, fp32) with
  // bytes b_t per stored entry and squared relative error a_t (a_drop = 1).
  // Spending the budget with the least bytes is a Lagrangian choice: a column
  // of error per element lam / (n + m) takes the tier minimising
  // b_t + mu a_t lam / (n + m), so the tiers are ranges of that ratio with
  // fixed ratios between their thresholds, th = (b_t' - b_t) / (a_t - a_t')
  // for consecutive tiers t < t' (the formats turned off are skipped); one
  // search for the lowest threshold spends the budget.
  const int kBinLo = -1100 * kBinsPerOctave, kBins = 2200 * kBinsPerOctave;
  auto bin_of = [&](double lam, int nm) {
    if (!(lam > 0)) return 0;
    const double v = std::log2(lam / nm) * kBinsPerOctave;
    return (int)std::min<double>(kBins - 1, std::max<double>(0, std::floor(v) - kBinLo));
  };
  constexpr int kTiers = 5;
  const bool int8 = (mx.formats & kMixInt8) != 0;
  // The 8-bit tier's error: e4m3's fixed relative error, or for int8 (each
  // column scaled by its largest entry) rho^2 / (12 127^2) per side, rho =
  // max / rms of the column, modelled as kI8Rho2 sqrt(rows) per side; the
  // model depends on the block's box sizes, so the blocks fall into size
  // buckets (a quarter octave of sqrt(n) + sqrt(m), the error taken at the
  // bucket's upper edge) with thresholds of their own.
  const int nbk = int8 ? kSizeBuckets : 1;
  auto bucket_of = [&](int n, int m) {
    if (!int8) return 0;
    const double sz = std::sqrt((double)n) + std::sqrt((double)m);
    return std::min(kSizeBuckets - 1, std::max(0, (int)std::floor(std::log2(sz) * 4.0)));
  };
  std::vector<std::array<double, kTiers>> ta(nbk);
  for (int q = 0; q < nbk; ++q) {
    const double a8 = int8 ? kI8Rho2 * std::exp2((q + 1) / 4.0) / (12.0 * 127.0 * 127.0)
                           : kFp8Err * kFp8Err;
    ta[q] = {1.0, a8, kFp16Err * kFp16Err, kFp24Err * kFp24Err, kFp32Err * kFp32Err};
  }
  const double tb[kTiers] = {0.0, 1.0, 2.0, 3.0, 4.0};
  const int tier_class[kTiers] = {-1, 2, 1, 3, 0};
  int av[kTiers], nav = 0;  // available tiers, increasing
  av[nav++] = 0;
  if (mx.formats & (kMixFp8 | kMixInt8)) av[nav++] = 1;
  if (mx.formats & kMixFp16) av[nav++] = 2;
  if (mx.formats & kMixFp24) av[nav++] = 3;
  av[nav++] = 4;
  // per bucket, bin offsets of the thresholds above the lowest one (the
  // lowest, drop against the first stored tier, is the same for every bucket
  // to within 1 - a_t)
  std::vector<std::array<int, kTiers>> toff(nbk);
  {
    const double th0 = (tb[av[1]] - tb[av[0]]) / (ta[0][av[0]] - ta[0][av[1]]);
    for (int q = 0; q < nbk; ++q) {
      toff[q].fill(0);
      for (int i = 1; i + 1 < nav; ++i) {
        const double th = (tb[av[i + 1]] - tb[av[i]]) / (ta[q][av[i]] - ta[q][av[i + 1]]);
        toff[q][i] = (int)std::lround(std::log2(th / th0) * kBinsPerOctave);
      }
    }
  }
  // No single column may carry more than kColShare of the budget (dropped, or
  // as rounding): the added error is then a sum over thousands of independent
  // terms and the error for one right-hand side stays close to its
  // expectation.  A column above a tier's cap takes the next available tier.
  const double budget = tol * tol * (mx.fro2_far + mx.fro2_near);
  const double cap = kColShare * budget;
  auto min_tier = [&](int q, double l) {  // index into av
    int i = 0;
    while (i + 1 < nav && ta[q][av[i]] * l > cap) ++i;
    return i;
  };
  // index into av of the tier of a column of bucket q in bin b with the lowest threshold at bin bd
  auto bin_tier = [&](int q, int b, int bd) {
    int i = 0;
    while (i + 1 < nav && b >= bd + toff[q][i]) ++i;
    return i;
  };
  int bd = 0;
  {
    // the bins and buckets the columns occupy (a first pass; per chunk,
    // combined in order)
    int blo = kBins, bhi = 0, qlo = nbk, qhi = 0;
    {
      std::vector<std::array<int, 4>> cr(kHostChunks, {kBins, 0, nbk, 0});
      host_chunks(nb, [&](int c, int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
          if (raw[i] || cd[i].r == 0) continue;
          const int nm = cd[i].n + cd[i].m;
          const int q = bucket_of(cd[i].n, cd[i].m);
          cr[c][2] = std::min(cr[c][2], q);
          cr[c][3] = std::max(cr[c][3], q);
          for (int k = 0; k < cd[i].r; ++k) {
            const int b = bin_of(lam_all[lam_off[i] + k], nm);
            cr[c][0] = std::min(cr[c][0], b);
            cr[c][1] = std::max(cr[c][1], b);
          }
        }
      });
      for (int c = 0; c < kHostChunks; ++c) {
        blo = std::min(blo, cr[c][0]), bhi = std::max(bhi, cr[c][1]);
        qlo = std::min(qlo, cr[c][2]), qhi = std::max(qhi, cr[c][3]);
      }
      if (blo > bhi) blo = bhi = 0;
      if (qlo > qhi) qlo = qhi = 0;
    }
    const int nbin = bhi - blo + 1;
    // H[hix(q, m, b)]: per bucket and bin, lam of the columns whose lowest
    // allowed tier is av[m] (buckets qlo..qhi, bins blo..bhi)
    const int64_t hsz = (int64_t)(qhi - qlo + 1) * nav * (nbin + 1);
    auto hix = [&](int q, int m, int b) {
      return ((int64_t)(q - qlo) * nav + m) * (nbin + 1) + (b - blo);
    };
    std::vector<double> H(hsz, 0.0);
    {  // per-chunk histograms, added in chunk order
      std::vector<std::vector<double>> ch_(kHostChunks);
      host_chunks(nb, [&](int c, int64_t i0, int64_t i1) {
        ch_[c].assign(hsz, 0.0);
        for (int64_t i = i0; i < i1; ++i) {
          if (raw[i]) continue;
          const int nm = cd[i].n + cd[i].m;
          const int q = bucket_of(cd[i].n, cd[i].m);
          for (int k = 0; k < cd[i].r; ++k) {
            const double l = lam_all[lam_off[i] + k];
            ch_[c][hix(q, min_tier(q, l), bin_of(l, nm))] += l;
          }
        }
      });
      for (int c = 0; c < kHostChunks; ++c)
        for (int64_t x = 0; x < hsz; ++x) H[x] += ch_[c][x];
    }
    // prefix sums per (bucket, lowest tier): P[hix(q, m, blo + j)] = sum of bins < blo + j
    std::vector<double> P(hsz, 0.0);
    for (int q = qlo; q <= qhi; ++q)
      for (int m = 0; m < nav; ++m) {
        const int64_t o = hix(q, m, blo);
        for (int j = 0; j < nbin; ++j) P[o + j + 1] = P[o + j] + H[o + j];
      }
    auto pref = [&](int q, int m, int b) {  // sum over bins < b (any b)
      const int j = std::min(nbin, std::max(0, b - blo));
      return P[hix(q, m, blo) + j];
    };
    // squared error per tier (index into av) with the lowest threshold at bd
    auto parts = [&](int bdv, double* out) {
      for (int i = 0; i < nav; ++i) out[i] = 0.0;
      for (int q = qlo; q <= qhi; ++q)
        for (int m = 0; m < nav; ++m) {
          int lo = 0;
          for (int i = 0; i < nav; ++i) {
            const int hi = i + 1 < nav ? std::min(kBins, std::max(lo, bdv + toff[q][i])) : kBins;
            const int t = std::max(m, i);
            out[t] += ta[q][av[t]] * (pref(q, m, hi) - pref(q, m, lo));
            lo = hi;
          }
        }
    };
    auto err = [&](int bdv) {
      double q[kTiers];
      parts(bdv, q);
      double sum = 0.0;
      for (int i = 0; i < nav; ++i) sum += q[i];
      return sum;
    };
    // largest bd with err(bd) <= budget (err is non-decreasing in bd)
    int lo = 0, hi = kBins;
    if (err(0) > budget) {
      bd = 0;
    } else {
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (err(mid) <= budget) lo = mid;
        else hi = mid - 1;
      }
      bd = lo;
    }
    mx.theta_drop = std::pow(2.0, (double)(bd + kBinLo) / kBinsPerOctave);
    double q[kTiers];
    parts(bd, q);
    double by[kTiers] = {0, 0, 0, 0, 0};
    for (int i = 0; i < nav; ++i) by[av[i]] = q[i];
    mx.err2_drop = by[0];
    mx.err2_fp8 = by[1];
    mx.err2_fp16 = by[2];
    mx.err2_fp24 = by[3];
    mx.err2_fp32 = by[4];
  }
  // per block class ranks: the tier only rises with lam, so each class is a
  // run of the block's sorted columns (fp32, then fp24, fp16, fp8)
  std::vector<std::array<int32_t, kClasses>> rk(nb);
  auto kept = [&](int64_t i) { return rk[i][0] + rk[i][1] + rk[i][2] + rk[i][3]; };
  mx.aca_rank_sum = 0;
  mx.dropped = 0;
  mx.raw_blocks = 0;
  mx.raw_rank = 0;
  host_chunks(nb, [&](int, int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const int r = cd[i].r;
      rk[i].fill(0);
      if (raw[i]) {
        rk[i][0] = r;
        continue;
      }
      const int nm = cd[i].n + cd[i].m;
      const int q = bucket_of(cd[i].n, cd[i].m);
      for (int k = 0; k < r; ++k) {
        const double l = lam_all[lam_off[i] + k];
        const int t = av[std::max(min_tier(q, l), bin_tier(q, bin_of(l, nm), bd))];
        if (t > 0) rk[i][tier_class[t]]++;
      }
    }
  });
  for (int64_t i = 0; i < nb; ++i) {
    const int r = cd[i].r;
    mx.aca_rank_sum += r;
    if (raw[i]) {
      mx.raw_blocks++;
      mx.raw_rank += r;
      continue;
    }
    mx.dropped += r - kept(i);
  }

  // ---------------- class layouts
  // U: per target box as in the ACA copy (same chunk ids), the class's columns
  // of its blocks in interaction-list order.  V: per source box, the class's
  // columns of every block with that source in by-source order (V_s), in V
  // chunks of whole source boxes.
  std::vector<int64_t> ci(e.nblk_all, -1);  // global block id -> index in cd
  for (int64_t i = 0; i < nb; ++i) ci[e.lev[refs[i].j].bbase + refs[i].b] = i;
  for (int c = 0; c < kClasses; ++c) {
    FactorClass& fc = mx.cls[c];
    fc.elem = kClassElem[c];
    fc.lf.assign(nL + 1, LevelFactors());
    fc.nU.assign(e.chunks.size(), 0);
    fc.udata.assign(e.chunks.size(), nullptr);
    fc.nV.assign(1, 0);
    fc.tmap.clear();
  }
  std::vector<PDesc> pd(nb);
  std::vector<int32_t> vch[kClasses];
  for (int c = 0; c < kClasses; ++c) vch[c].assign(e.nblk_all, 0);
  // per level: the target box of each block and the blocks per source box
  std::vector<std::vector<int32_t>> nsrc_l(nL + 1), tgt_l(nL + 1);
  for (int j = 1; j <= nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    nsrc_l[j].assign(lv.nboxes, 0);
    tgt_l[j].assign(lv.ilidx.size(), 0);
    for (int32_t ki : lv.ilidx) nsrc_l[j][ki]++;
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
        tgt_l[j][b] = (int32_t)k;
        const int64_t i = ci[lv.bbase + b];
        if (i < 0) continue;
        const int m = e.bcount[lo + lv.ilidx[b]];
        for (int c = 0; c < kClasses; ++c) {
          pd[i].rc[c] = rk[i][c];
          pd[i].ldv_o[c] = class_ldv(m, kClassElem[c]);
        }
        pd[i].raw = raw[i];
        pd[i].su = raw[i] ? 1.0 : pow2_near_sqrt(n);
        pd[i].sv = raw[i] ? 1.0 : pow2_near_sqrt(m);
      }
    }
  }
  // each class on a host thread of its own (its layout depends only on its
  // own ranks; the threads write disjoint fields)
  auto layout_class = [&](int c) {
    FactorClass& fc = mx.cls[c];
    int32_t T = 0, tsrc_cur = 0;
    for (int j = 1; j <= nL; ++j) {
      const Level& lv = e.lev[j];
      const int64_t lo = level_offset(j);
      const int64_t nblk = (int64_t)lv.ilidx.size();
      LevelFactors& f = fc.lf[j];
      f.rank.assign(nblk, 0);
      f.col.assign(nblk, 0);
      f.voff.assign(nblk, 0);
      f.Rt.assign(lv.nboxes, 0);
      f.toff.assign(lv.nboxes, 0);
      f.uoff.assign(lv.nboxes, 0);
      for (int64_t k = 0; k < lv.nboxes; ++k) {
        const int n = e.bcount[lo + k];
        const int ch = lv.chunk[k];
        int32_t R = 0;
        for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
          const int64_t i = ci[lv.bbase + b];
          if (i < 0) continue;
          const int m = e.bcount[lo + lv.ilidx[b]];
          const int r = rk[i][c];
          f.rank[b] = r;
          f.col[b] = R;
          R += r;
          fc.rank_sum += r;
          fc.u_elems += (int64_t)r * n;
          fc.v_elems += (int64_t)r * m;
        }
        const int ldu = class_ldu(n, fc.elem);
        f.Rt[k] = R;
        f.uoff[k] = class_ustart(fc.nU[ch], n, fc.elem);
        fc.nU[ch] = f.uoff[k] + (int64_t)ldu * R;
        f.toff[k] = T;
        T += R;
        for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
          const int64_t i = ci[lv.bbase + b];
          if (i < 0) continue;
          pd[i].u_out[c] = f.uoff[k] + (int64_t)f.col[b] * ldu;
          pd[i].ldu_o[c] = ldu;
          if (c == 2) pd[i].t8 = f.toff[k] + f.col[b];
        }
      }
      // V by source box; the by-source list of level j occupies
      // e.by_src[bbase, bbase + nblk), grouped by source box in order
      const std::vector<int32_t>& nsrc = nsrc_l[j];
      const std::vector<int32_t>& tgt = tgt_l[j];
      fc.tmap.resize(T, 0);
      int64_t pos = lv.bbase;
      for (int64_t sb = 0; sb < lv.nboxes; ++sb) {
        const int m = e.bcount[lo + sb];
        const int ld = class_ldv(m, fc.elem);
        int64_t C = 0;
        for (int32_t q = 0; q < nsrc[sb]; ++q) C += f.rank[e.by_src[pos + q] - lv.bbase];
        if (fc.nV.back() > 0 && fc.nV.back() + C * ld > kVChunkElems) fc.nV.push_back(0);
        const int32_t vc = (int32_t)fc.nV.size() - 1;
        const int64_t vs = fc.nV.back();  // V_s column 0
        const int32_t tsb = tsrc_cur;     // source-order t of V_s column 0
        tsrc_cur += (int32_t)C;
        int32_t col = 0;
        for (int32_t q = 0; q < nsrc[sb]; ++q) {
          const int32_t gid = e.by_src[pos + q];
          const int32_t b = gid - (int32_t)lv.bbase;
          const int r = f.rank[b];
          f.voff[b] = vs + (int64_t)col * ld;
          vch[c][gid] = vc;
          const int32_t tt = f.toff[tgt[b]] + f.col[b];
          for (int l = 0; l < r; ++l) fc.tmap[tt + l] = tsb + col + l;
          const int64_t i = ci[gid];
          if (i >= 0) {
            pd[i].v_out[c] = f.voff[b];
            pd[i].vch[c] = vc;
          }
          col += r;
        }
        fc.nV.back() += C * ld;
        pos += nsrc[sb];
        // transposed phase 2 over V_s: rows in pieces of about 1 KB (a
        // multiple of 4 rows for fp24), t in source order
        if (m > 0 && C > 0) {
          const int p2r = fc.elem == 3 ? 256 : 1024 / fc.elem;
          for (int i0 = 0; i0 < m; i0 += p2r) {
            P2Item it{};
            it.uoff = vs;
            it.chunk = vc;
            it.ld = ld;
            it.i0 = i0;
            it.nr = std::min(p2r, m - i0);
            it.R = (int32_t)C;
            it.toff = tsb;
            it.yoff = e.bstart[lo + sb];
            it.level = j;
            fc.tp2v.push_back(it);
          }
        }
        // phase-1 items: rows in ranges of <= kMixRows, columns in runs of
        // about kMixItemBytes of factors
        if (m == 0 || C == 0) continue;
        const int kMixRows = mix_item_rows(fc.elem);
        const int nsplit = (m + kMixRows - 1) / kMixRows;
        const int rows0 = std::min(m, kMixRows);
        const int ncol = (int)std::max<int64_t>(
            16, std::min<int64_t>(4096, kMixItemBytes / ((int64_t)rows0 * fc.elem)));
        for (int32_t c0 = 0; c0 < C; c0 += ncol) {
          const int32_t c1 = (int32_t)std::min<int64_t>(C, (int64_t)c0 + ncol);
          int32_t pbase = -1;
          if (nsplit > 1) {
            pbase = fc.mpartial;
            fc.mpartial += nsplit * (c1 - c0);
            fc.mp1r.push_back({pbase, nsplit, c1 - c0, tsb + c0});
          }
          for (int q = 0; q < nsplit; ++q) {
            MixP1Item it{};
            it.voff = vs;
            it.vchunk = vc;
            it.ld = ld;
            it.s0 = q * kMixRows;
            it.s1 = std::min(m, (q + 1) * kMixRows);
            it.c0 = c0;
            it.c1 = c1;
            it.tsb = tsb;
            it.xoff = e.bstart[lo + sb];
            it.pslot = nsplit > 1 ? pbase + q * (c1 - c0) : -1;
            fc.mp1.push_back(it);
          }
        }
      }
    }
    fc.T = T;
  };
  run_per_class(layout_class);
  {
    int32_t tb = 0;
    for (int c = 0; c < kClasses; ++c) {
      mx.cls[c].tbase = tb;
      tb += mx.cls[c].T;
    }
  }
  // allocate and zero the class chunks: one allocation per class (a few large
  // allocations rather than one per chunk), each chunk at a 256-byte boundary
  for (int c = 0; c < kClasses; ++c) {
    FactorClass& fc = mx.cls[c];
    fc.vdata.assign(fc.nV.size(), nullptr);
    auto span = [&](int64_t n) { return ((n + kChunkTail) * fc.elem + 255) / 256 * 256; };
    int64_t total = 0;
    for (int64_t n : fc.nU) total += span(n);
    for (int64_t n : fc.nV) total += span(n);
    try {
      dmalloc(&fc.arena, (size_t)std::max<int64_t>(total, 256));
    } catch (const Error& er) {
      if (er.code != H3D_ERR_OOM || !d_j) throw;
      // the stored J gives way: the second pass repeats the sweeps instead
      dfree(d_j);
      mx.single_jacobi = false;
      dmalloc(&fc.arena, (size_t)std::max<int64_t>(total, 256));
    }
    H3D_HIP(hipMemsetAsync(fc.arena, 0, (size_t)std::max<int64_t>(total, 256), st));
    int64_t off = 0;
    auto carve = [&](void*& p, int64_t n) {
      p = (char*)fc.arena + off;
      off += span(n);
      fc.bytes += n * fc.elem;
    };
    for (size_t ch = 0; ch < fc.nU.size(); ++ch) carve(fc.udata[ch], fc.nU[ch]);
    for (size_t vc = 0; vc < fc.nV.size(); ++vc) carve(fc.vdata[vc], fc.nV[vc]);
    fc.d_ubase = upload(fc.udata, st);
    fc.d_vbase = upload(fc.vdata, st);
  }
  // per column factor c_k = nu_k nv_k / (su sv) = sqrt(lam_k) / (su sv) in t
  // order (fp8 columns: / kFp8Gain^2; raw columns: 1), block tables and the
  // classes' work lists: formed on a host thread while the GPU runs pass 2
  // (they need the layouts only), uploaded after it
  std::vector<double> scale;
  std::vector<float> scale32[kClasses];
  std::vector<BlockDev> blk[kClasses];
  std::vector<int32_t> btoff[kClasses];
  std::exception_ptr tables_err;
  auto tables = [&] {
    scale.assign((size_t)mx.cls[0].T + mx.cls[1].T + mx.cls[2].T + mx.cls[3].T, 1.0);
    for (int c = 0; c < kClasses; ++c) {
      blk[c].assign(e.nblk_all, BlockDev{});
      btoff[c].assign(e.nblk_all, 0);
    }
    for (int j = 1; j <= nL; ++j) {
      const Level& lv = e.lev[j];
      const int64_t lo = level_offset(j);
      for (int64_t k = 0; k < lv.nboxes; ++k)
        for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
          const int32_t ki = lv.ilidx[b];
          for (int c = 0; c < kClasses; ++c) {
            const FactorClass& fc = mx.cls[c];
            const LevelFactors& f = fc.lf[j];
            BlockDev d{};
            d.voff = f.voff[b];
            d.col = f.col[b];
            d.r = f.rank[b];
            d.m = e.bcount[lo + ki];
            d.src = e.bstart[lo + ki];
            d.chunk = vch[c][lv.bbase + b];
            blk[c][lv.bbase + b] = d;
            btoff[c][lv.bbase + b] = f.toff[k] + f.col[b];
          }
          const int64_t i = ci[lv.bbase + b];
          if (i < 0 || raw[i]) continue;
          const double inv = 1.0 / (pd[i].su * pd[i].sv);
          const double fl = lam_floor[i];
          int k0 = 0;  // first sorted column of the class
          for (int c : kColumnOrder) {
            const LevelFactors& f = mx.cls[c].lf[j];
            // e4m3 columns carry kFp8Gain per side; int8 columns their
            // maxima / 127, applied after the projection
            const double g = c == 2 && !int8 ? inv / (kFp8Gain * kFp8Gain) : inv;
            for (int kk = 0; kk < rk[i][c]; ++kk)
              scale[(size_t)mx.cls[c].tbase + f.toff[k] + f.col[b] + kk] =
                  std::sqrt(std::max(lam_all[lam_off[i] + k0 + kk], fl)) * g;
            k0 += rk[i][c];
          }
        }
    }
    // classes 1 and 2 sum in fp32: their factors converted once (the value
    // the gather would convert to)
    for (int c = 1; c < kClasses; ++c) {
      if (class_acc64(c)) continue;
      scale32[c].resize(mx.cls[c].T);
      for (int32_t q = 0; q < mx.cls[c].T; ++q)
        scale32[c][q] = (float)scale[(size_t)mx.cls[c].tbase + q];
    }
    for (int c = 0; c < kClasses; ++c) {
      FactorClass& fc = mx.cls[c];
      std::vector<const LevelFactors*> lfs(nL + 1, nullptr);
      for (int j = 1; j <= nL; ++j) lfs[j] = &fc.lf[j];
      // about 1 KB of rows per phase-2 item (a multiple of 4 rows for fp24, so
      // that its 4-entry vectors start 4-byte aligned); phase 1 runs on the
      // source-grouped items, so the plan keeps no phase-1 list
      build_plan_host(e, lfs, fc.plan, fc.elem == 3 ? 256 : 1024 / fc.elem, fc.elem);
      fc.plan.p1.clear();
      fc.plan.p1r.clear();
    }
  };
  std::thread tables_thread([&] {
    try {
      tables();
    } catch (...) {
      tables_err = std::current_exception();
    }
  });
  // joined on every exit from here (pass 2 may throw)
  struct Joiner {
    std::thread& t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } joiner{tables_thread};
  // ---------------- pass 2: the projections (from the stored J, or from the
  // sweeps again), then project into the classes; blocks that keep no column
  // and raw blocks need no eigen step
  std::vector<uint8_t> work(nb, 0);
  for (int64_t i = 0; i < nb; ++i) work[i] = !raw[i] && kept(i) > 0;
  // int8: each class-2 column side's largest |entry| (fp32 bits), U sides then V sides
  const int64_t T2 = int8 ? mx.cls[2].T : 0;
  uint32_t* d_mx8 = nullptr;
  struct Mx8Free {
    uint32_t*& p;
    ~Mx8Free() { dfree(p); }
  } mx8_free{d_mx8};
  if (T2 > 0) {
    d_mx8 = dalloc<uint32_t>(2 * T2);
    H3D_HIP(hipMemsetAsync(d_mx8, 0, sizeof(uint32_t) * 2 * T2, st));
  }
  {
    int64_t i0 = 0;
    while (i0 < nb) {
      int64_t i1 = i0, need = 0;
      while (i1 < nb && (i1 == i0 || need + wsize(cd[i1].r) <= ws_budget)) need += wsize(cd[i1++].r);
      run_eig(i0, i1, d_j ? kFinish : kVectors, &work);
      std::vector<PDesc> sub(pd.begin() + i0, pd.begin() + i1);
      // blocks of rank > kMaxEigRank have no eigen step: raw in both passes
      PDesc* d_pd = upload(sub, st);
      std::vector<PItem> items;
      for (int64_t i = i0; i < i1; ++i) {
        if (kept(i) == 0) continue;
        for (int side = 0; side < 2; ++side) {
          const int rows = side ? cd[i].m : cd[i].n;
          for (int r0 = 0; r0 < rows; r0 += 256) items.push_back({(int32_t)(i - i0), side, r0, 0});
        }
      }
      PItem* d_it = upload(items, st);
      ClassBases cb{};
      for (int c = 0; c < kClasses; ++c) {
        cb.u[c] = mx.cls[c].d_ubase;
        cb.v[c] = mx.cls[c].d_vbase;
      }
      for (int64_t s = 0; s < (int64_t)items.size(); s += 1 << 30) {
        const unsigned g = (unsigned)std::min<int64_t>(1 << 30, (int64_t)items.size() - s);
        if (T2 > 0) {
          (src ? project_kernel<float, 1> : project_kernel<double, 1>)<<<g, 256, 0, st>>>(
              d_it + s, d_cd, d_pd, d_b64, d_ws, cb, d_mx8, T2);
          H3D_LAUNCH_CHECK();
          (src ? project_kernel<float, 2> : project_kernel<double, 2>)<<<g, 256, 0, st>>>(
              d_it + s, d_cd, d_pd, d_b64, d_ws, cb, d_mx8, T2);
        } else {
          (src ? project_kernel<float, 0> : project_kernel<double, 0>)<<<g, 256, 0, st>>>(
              d_it + s, d_cd, d_pd, d_b64, d_ws, cb, nullptr, 0);
        }
        H3D_LAUNCH_CHECK();
      }
      H3D_HIP(hipStreamSynchronize(st));
      dfree(d_pd);
      dfree(d_it);
      i0 = i1;
    }
  }
  dfree(d_ws);
  dfree(d_cd);
  dfree(d_list);
  dfree(d_b64);
  tables_thread.join();
  if (tables_err) std::rethrow_exception(tables_err);
  mx.err2_fp8_model = mx.err2_fp8;
  if (T2 > 0) {
    // int8 factors: each column's factor times (max_u / 127) (max_v / 127),
    // and the realized rounding error of the class, sum over its columns of
    // lam (rho_u^2 + rho_v^2) / (12 127^2) with rho^2 = max^2 rows / s^2
    std::vector<uint32_t> hm(2 * T2);
    H3D_HIP(hipMemcpyAsync(hm.data(), d_mx8, sizeof(uint32_t) * 2 * T2, hipMemcpyDeviceToHost, st));
    H3D_HIP(hipStreamSynchronize(st));
    auto mxf = [&](int64_t q) {
      float f;
      std::memcpy(&f, &hm[q], 4);
      return (double)f;
    };
    const int64_t tb2 = mx.cls[2].tbase;
    for (int64_t q = 0; q < T2; ++q) {
      scale[tb2 + q] *= (mxf(q) / 127.0) * (mxf(T2 + q) / 127.0);
      scale32[2][q] = (float)scale[tb2 + q];
    }
    std::vector<double> ce(kHostChunks, 0.0);
    host_chunks(nb, [&](int c, int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        if (raw[i] || rk[i][2] == 0) continue;
        const int k8 = rk[i][0] + rk[i][3] + rk[i][1];
        const double su2 = pd[i].su * pd[i].su, sv2 = pd[i].sv * pd[i].sv;
        for (int k = 0; k < rk[i][2]; ++k) {
          const int64_t q = pd[i].t8 + k;
          const double l = std::max(lam_all[lam_off[i] + k8 + k], lam_floor[i]);
          const double mu = mxf(q), mv = mxf(T2 + q);
          ce[c] += l * (mu * mu * cd[i].n / su2 + mv * mv * cd[i].m / sv2) / (12.0 * 127.0 * 127.0);
        }
      }
    });
    mx.err2_fp8 = 0.0;
    for (double v : ce) mx.err2_fp8 += v;
  }
  {
    mx.d_scale = upload(scale, st);
    for (int c = 1; c < kClasses; ++c)
      if (!class_acc64(c)) mx.cls[c].d_scale32 = upload(scale32[c], st);
    for (int c = 0; c < kClasses; ++c) {
      FactorClass& fc = mx.cls[c];
      fc.d_blk = upload(blk[c], st);
      fc.d_btoff = upload(btoff[c], st);
      upload_plan(e, fc.plan);
      fc.d_mp1 = upload(fc.mp1, st);
      fc.d_mp1r = upload(fc.mp1r, st);
      fc.d_tp2v = upload(fc.tp2v, st);
      fc.d_tmap = upload(fc.tmap, st);
    }
    H3D_HIP(hipStreamSynchronize(st));
  }
  mx.built_formats = mx.formats;
  mx.int8 = int8;
  mx.ready = true;
  mx.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace h3d
