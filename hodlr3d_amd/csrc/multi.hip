// Products with many right-hand sides: the RHS are taken W (16 or 32) at a
// time and every factor element is read once per group and applied to all W
// of them on the matrix cores (the caller side of MatVecProduct,
// src/HODLR3D.cpp:21-27, called once per column in the reference).
//
//   gather    x by position in two layouts: quads (x[p..p+3] of one RHS
//             contiguous, the MFMA B operand of phase 1) and rows (the W
//             values of one point contiguous, read by the near field)
//   phase 1   T_b = V_b^T X_s   per block, 16x16x4 MFMA: A = 16 columns x 4
//             rows of V (one 4-row vector per lane), B = 4 rows x 16 RHS
//   phase 2   Y_k = U_k T_k     per <= 64 rows of a target box, 16x16x4 MFMA:
//             A = 16 row quads x 4 columns of U, B = 4 columns x 16 RHS of T
//   near      y_N = K_near X    lanes own target rows, W sums each, source
//             points and their W values through wave-uniform (scalar) loads;
//             own buffer, on the handle's second stream under phases 1 and 2
//   combine   y[:, j] = alpha * (sum_levels + y_N) + d * b[:, j] (an LDS transpose)
//
// The work lists are the single-RHS plan's (phase-1 items, partial sums and
// phase-2 items), so the order of the block loop is the same; fp64 factors
// and the fp64-accumulation modes use the fp64 MFMA, fp32/bf16 factors with
// fp32 accumulation the fp32 MFMA.
#include <algorithm>
#include <type_traits>

#include "engine.hpp"
#include "kernels.hpp"
#include "stream.hpp"

namespace h3d {
namespace {

constexpr int kGRows = 256;     // rows of one phase-1 item (x staged in LDS)

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16) on one wave: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16], and holds D[row(l / 16, a)][l % 16] in element a:
// row = 4 (l / 16) + a for the fp32 instruction, 4 a + l / 16 for the fp64
// one (checked element by element on gfx950).
template <class A>
struct Mfma;
template <>
struct Mfma<float> {
  using acc = f32x4;
  __device__ static inline acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int q, int a) { return 4 * q + a; }
};
template <>
struct Mfma<double> {
  using acc = f64x4;
  __device__ static inline acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int q, int a) { return 4 * a + q; }
};

// Four consecutive stored elements (4-byte aligned) converted to A.
template <class S, class A>
__device__ inline void load4(const S* p, A (&o)[4]) {
  if constexpr (sizeof(S) == 8) {
    const u32x4a v0 = ntload<u32x4a>(p), v1 = ntload<u32x4a>(p + 2);
    o[0] = (A)__hiloint2double((int)v0.y, (int)v0.x);
    o[1] = (A)__hiloint2double((int)v0.w, (int)v0.z);
    o[2] = (A)__hiloint2double((int)v1.y, (int)v1.x);
    o[3] = (A)__hiloint2double((int)v1.w, (int)v1.z);
  } else if constexpr (sizeof(S) == 4) {
    const u32x4a v = ntload<u32x4a>(p);
    o[0] = (A)__uint_as_float(v.x);
    o[1] = (A)__uint_as_float(v.y);
    o[2] = (A)__uint_as_float(v.z);
    o[3] = (A)__uint_as_float(v.w);
  } else {  // bf16: element 0 in the low half of each word
    const u32x2a v = ntload<u32x2a>(p);
    o[0] = (A)__uint_as_float(v.x << 16);
    o[1] = (A)__uint_as_float(v.x & 0xffff0000u);
    o[2] = (A)__uint_as_float(v.y << 16);
    o[3] = (A)__uint_as_float(v.y & 0xffff0000u);
  }
}

// Four consecutive x values of one RHS in the quad layout (16 or 32 bytes).
template <class A>
__device__ inline void load_quad(const A* p, A (&o)[4]) {
  if constexpr (sizeof(A) == 8) {
    const f64x4 v = *(const f64x4*)p;
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
    const f32x4 v = *(const f32x4*)p;
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
}

// Quad layout: x(p, j) at xq[((p / 4) W + j) 4 + p % 4], Npad positions (the
// rows past N are zero); row layout: x(p, j) at xr[p W + j].  Columns j >= cnt
// are zero.
template <class A, int W>
__global__ void gather_w_kernel(const double* __restrict__ b, const int32_t* __restrict__ perm,
                                int64_t N, int64_t Npad, int cnt, A* __restrict__ xq,
                                A* __restrict__ xr) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= Npad * W) return;
  const int j = (int)((id >> 2) % W);
  const int64_t p = 4 * (id / (4 * W)) + (id & 3);
  const A v = (p < N && j < cnt) ? (A)b[(int64_t)j * N + perm[p]] : (A)0;
  xq[id] = v;
  xr[p * W + j] = v;
}

// Phase 1, grouped by source box: one workgroup per item = a row range
// [s0, s1) (at most kGRows rows) of one source box and a slice of the blocks
// with that source (at most Engine::g_cols columns, tuned).  x of the range (as quads,
// rows outside the range zeroed) is staged once in LDS and shared by the four
// waves, which take the slice's blocks in turn: T_b(c, j) = sum over the range
// of V_b(c, p) x(p, j), 32 columns (two 16-column tiles) per pass, UF 16-row
// steps per unit of work (gdesc describes the blocks of rank >= 1 by source box).
// In step s lane (c = l % 16, q = l / 16) takes rows
// P + 16 s + 4q .. + 3 of column c.  The factor loads of the next unit (next
// steps, next pass or next block) are issued before the MFMAs of the current
// one (two register buffers).  A block whose source has more than kGRows rows
// writes partial sums (slot pbw[b] + range * r + c), summed in range order
// afterwards.
template <class S, class A, int W>
__global__ __launch_bounds__(256) void p1g_kernel(const GItem* __restrict__ items, int nitems,
                                                  const GDesc* __restrict__ gdesc,
                                                  void* const* __restrict__ vbase,
                                                  const A* __restrict__ xq, A* __restrict__ tw,
                                                  A* __restrict__ partw) {
  using M = Mfma<A>;
  using Acc = typename M::acc;
  constexpr int H = W / 16;
  constexpr int CT = 2;
  constexpr int UF = sizeof(A) == 8 ? 2 : 4;
  constexpr int QMAX = 4 * ((kGRows + 15) / 16);  // quads of the steps of one range
  __shared__ A xl[QMAX * W * 4];
  const int item = blockIdx.x;
  if (item >= nitems) return;  // whole workgroup
  const GItem it = items[item];
  const int a0 = it.src + it.s0;
  const int nrows = it.s1 - it.s0;
  const int nsteps = (nrows + 15) >> 4;
  {
    // x of range row i (position a0 + i) at xl[((i / 4) W + j) 4 + i % 4]; zero past the range
    const int ne = 16 * nsteps * W;
    for (int e = threadIdx.x; e < ne; e += 256) {
      const int i = (e / (4 * W)) * 4 + (e & 3), j = (e >> 2) % W;
      const int p = a0 + i;
      xl[e] = i < nrows ? xq[((int64_t)(p >> 2) * W + j) * 4 + (p & 3)] : (A)0;
    }
  }
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int lc = lane & 15, kq = lane >> 4;
  const int range = it.s0 / kGRows;

  // unit of work: block gdesc[q], column pass cb, steps [s0, s0 + UF)
  struct Unit {
    int q, r, cb, s0;
    const S* V;
    int64_t ldv;
    A* out;
  };
  // the wave's next block descriptor is loaded one block ahead
  GDesc pre = it.q0 + wv < it.q1 ? gdesc[it.q0 + wv] : GDesc{};
  auto open_block = [&](Unit& u) -> bool {  // first unit of block q (rank >= 1)
    if (u.q >= it.q1) return false;
    const GDesc d = pre;
    if (u.q + kWavesPerBlock < it.q1) pre = gdesc[u.q + kWavesPerBlock];
    u.r = d.r;
    u.cb = 0;
    u.s0 = 0;
    u.V = (const S*)vbase[d.chunk] + d.voff;
    u.ldv = d.m + (d.m & 1);
    u.out = d.pb < 0 ? tw + (int64_t)d.tb * W
                     : partw + ((int64_t)d.pb + (int64_t)range * d.r) * W;
    return true;
  };
  auto advance = [&](Unit& u) -> bool {
    u.s0 += UF;
    if (u.s0 < nsteps) return true;
    u.s0 = 0;
    u.cb += 16 * CT;
    if (u.cb < u.r) return true;
    u.q += kWavesPerBlock;
    return open_block(u);
  };
  // Branch-free loads: one 4-row vector per lane and tile at an in-bounds
  // address (columns clamped to r - 1; a vector starting inside the range ends
  // at most 3 rows past it, in the column padding, the next column or the
  // chunk's zero tail); every factor value read is finite.
  // The loaded vectors are consumed only by the MFMAs (compute), so the wait
  // for a unit's loads falls after the previous unit's MFMAs.
  auto load = [&](const Unit& u, A (&av)[UF][CT][4]) {
#pragma unroll
    for (int st = 0; st < UF; ++st) {
      const int i = 16 * (u.s0 + st) + 4 * kq;
      const int il = i < nrows ? i : 0;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int c = u.cb + 16 * ct + lc;
        const int cl = c < u.r ? c : u.r - 1;
        load4<S, A>(u.V + cl * u.ldv + it.s0 + il, av[st][ct]);
      }
    }
  };
  // Both column tiles are multiplied even when the second lies past the rank
  // (its loads are clamped, its output rows never stored): a wave-uniform
  // branch between the last MFMA and the accumulator reads of the store let
  // the compiler place those reads 2 wait states after the MFMA issue (it
  // needs 10 for the fp32 16x16x4 product), so the store took stale sums.
  auto compute = [&](const Unit& u, const A (&av)[UF][CT][4], Acc (&acc)[CT][H]) {
#pragma unroll
    for (int st = 0; st < UF; ++st) {
      if (u.s0 + st >= nsteps) break;
      A xb[H][4];
#pragma unroll
      for (int h = 0; h < H; ++h)
        load_quad<A>(xl + ((4 * (u.s0 + st) + kq) * W + 16 * h + lc) * 4, xb[h]);
      // no masks: rows past the range meet zero x (staged so), and the
      // output rows of clamped columns (c >= r) are never stored
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int h = 0; h < H; ++h) acc[ct][h] = M::mac(av[st][ct][e], xb[h][e], acc[ct][h]);
    }
  };
  auto store = [&](const Unit& u, Acc (&acc)[CT][H]) {
    const bool two = u.cb + 16 < u.r;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      if (ct == 0 || two) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int c = u.cb + 16 * ct + M::row(kq, a);
          if (c < u.r) {
#pragma unroll
            for (int h = 0; h < H; ++h) u.out[(int64_t)c * W + 16 * h + lc] = acc[ct][h][a];
          }
        }
      }
#pragma unroll
      for (int h = 0; h < H; ++h) acc[ct][h] = Acc{};
    }
  };

  Unit cur;
  cur.q = it.q0 + wv;
  if (!open_block(cur)) return;
  A av0[UF][CT][4], av1[UF][CT][4];
  Acc acc[CT][H];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int h = 0; h < H; ++h) acc[ct][h] = Acc{};
  load(cur, av0);
  for (;;) {
    Unit nxt = cur;
    bool more = advance(nxt);
    if (more) load(nxt, av1);
    compute(cur, av0, acc);
    if (!more || nxt.q != cur.q || nxt.cb != cur.cb) store(cur, acc);
    if (!more) break;
    cur = nxt;
    nxt = cur;
    more = advance(nxt);
    if (more) load(nxt, av0);
    compute(cur, av1, acc);
    if (!more || nxt.q != cur.q || nxt.cb != cur.cb) store(cur, acc);
    if (!more) break;
    cur = nxt;
  }
}

// Row-split partial sums, one thread per (column, RHS): column q of the table
// holds {first partial entry, number of ranges, entries per range, t index};
// the ranges are summed in order, eight loads issued ahead of their adds (a
// plain loop waits for each load).
template <class A, int W>
__global__ void reduce_w_kernel(const int4* __restrict__ cols, int64_t ncols,
                                const A* __restrict__ partw, A* __restrict__ tw) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= ncols * W) return;
  const int j = (int)(id % W);
  const int4 c = cols[id / W];
  const A* p = partw + (int64_t)c.x * W + j;
  const int64_t stride = (int64_t)c.z * W;
  A acc = 0;
  int q = 0;
  for (; q + 8 <= c.y; q += 8) {
    A v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(q + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; q < c.y; ++q) acc += p[(int64_t)q * stride];
  tw[(int64_t)c.w * W + j] = acc;
}

// Phase 2: y(i, j) = sum_c U_k(i, c) T_k(c, j) for <= 64 rows at a time; lane
// (g = l % 16, q = l / 16) takes rows i + 4g .. i + 4g + 3 of column c + q,
// UF 4-column steps per unit, the next unit's loads issued before the MFMAs
// of the current one.  The u-th MFMA of a step produces rows i + 4g' + u (g'
// the MFMA output row).
template <class S, class A, int W>
__global__ __launch_bounds__(256) void p2w_kernel(const P2Item* __restrict__ items, int nitems,
                                                  void* const* __restrict__ ubase,
                                                  const A* __restrict__ tw,
                                                  double* __restrict__ ylevw, int64_t N) {
  using M = Mfma<A>;
  using Acc = typename M::acc;
  constexpr int H = W / 16;
  constexpr int UF = sizeof(A) == 8 ? 2 : 4;
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const int lc = lane & 15, kq = lane >> 4;
  const P2Item it = items[w];
  const S* U = (const S*)ubase[it.chunk] + it.uoff;
  double* yb = ylevw + ((int64_t)(it.level - 1) * N + it.yoff) * W;
  for (int r0 = 0; r0 < it.nr; r0 += 64) {
    const int lim = it.i0 + min(it.nr, r0 + 64);  // rows < lim (box-relative)
    const int i = it.i0 + r0 + 4 * lc;
    bool ok[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) ok[u] = i + u < lim;
    // branch-free loads at in-bounds addresses (column clamped to R - 1, rows
    // of lanes past the chunk moved to its first row; a vector starting inside
    // the box ends at most 3 rows past it)
    const int il = ok[0] ? i : it.i0 + r0;
    // consumed only where used, so the wait for them follows the previous
    // unit's MFMAs
    auto load = [&](int c0, A (&av)[UF][4], A (&tb)[UF][H]) {
#pragma unroll
      for (int st = 0; st < UF; ++st) {
        const int c = c0 + 4 * st + kq;
        const int cl = c < it.R ? c : it.R - 1;
#pragma unroll
        for (int h = 0; h < H; ++h) tb[st][h] = tw[(int64_t)(it.toff + cl) * W + 16 * h + lc];
        load4<S, A>(U + (int64_t)cl * it.ld + il, av[st]);
      }
    };
    Acc acc[4][H];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int h = 0; h < H; ++h) acc[u][h] = Acc{};
    auto compute = [&](int c0, const A (&av)[UF][4], const A (&tb)[UF][H]) {
#pragma unroll
      for (int st = 0; st < UF; ++st) {
        if (c0 + 4 * st >= it.R) break;  // wave-uniform
        // only t is masked (clamped columns); rows past the box give output
        // rows that are never stored
        const bool cok = c0 + 4 * st + kq < it.R;
        A bm[H];
#pragma unroll
        for (int h = 0; h < H; ++h) bm[h] = cok ? tb[st][h] : (A)0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int h = 0; h < H; ++h) acc[u][h] = M::mac(av[st][u], bm[h], acc[u][h]);
      }
    };
    constexpr int STEP = 4 * UF;
    A av0[UF][4], av1[UF][4], tb0[UF][H], tb1[UF][H];
    if (it.R > 0) load(0, av0, tb0);
    for (int c0 = 0; c0 < it.R; c0 += 2 * STEP) {
      if (c0 + STEP < it.R) load(c0 + STEP, av1, tb1);
      compute(c0, av0, tb0);
      if (c0 + STEP >= it.R) break;
      if (c0 + 2 * STEP < it.R) load(c0 + 2 * STEP, av0, tb0);
      compute(c0 + STEP, av1, tb1);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = it.i0 + r0 + 4 * M::row(kq, a) + u;
        if (row < lim) {
#pragma unroll
          for (int h = 0; h < H; ++h) yb[(int64_t)row * W + 16 * h + lc] = (double)acc[u][h][a];
        }
      }
  }
}

// Near field, fp64 coordinates and sums (exact entries, or 1/r refined from
// the fp32 reciprocal square root): lanes own target rows; the source point
// and its W values are wave-uniform.
template <class KF, int W>
__global__ __launch_bounds__(256) void near_w64_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    KernelDev kd, const double* __restrict__ xr, double* __restrict__ yNw) {
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  const bool act = lane < it.nr;
  const int ip = rs + it.i0 + (act ? lane : 0);
  const double xi = px[ip], yi = py[ip], zi = pz[ip];
  double acc[W];
#pragma unroll
  for (int j = 0; j < W; ++j) acc[j] = 0.0;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    for (int s = cs; s < cs + m; ++s) {
      const double k = s == ip ? 0.0 : kpair<KF>(kd.q, xi, yi, zi, px[s], py[s], pz[s]);
      const double* xs = xr + (int64_t)s * W;
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] += k * xs[j];
    }
  }
  if (!act) return;
  double* yp = yNw + (int64_t)ip * W;
#pragma unroll
  for (int j = 0; j < W; ++j) yp[j] = kd.scale * acc[j];
}

// Near field with fp32 entries and sums: coordinates as fp32 offsets from the
// leaf centre plus the fp32 difference of the two centres (as the single-RHS
// fp32 near field).
template <class KF, int W>
__global__ __launch_bounds__(256) void near_w32_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const float* __restrict__ lx, const float* __restrict__ ly, const float* __restrict__ lz,
    const double* __restrict__ leafc, KernelDev kd, const float* __restrict__ xr,
    double* __restrict__ yNw) {
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  const bool act = lane < it.nr;
  const int ip = rs + it.i0 + (act ? lane : 0);
  const float xi = lx[ip], yi = ly[ip], zi = lz[ip];
  const double cxk = leafc[3 * (int64_t)it.leaf], cyk = leafc[3 * (int64_t)it.leaf + 1],
               czk = leafc[3 * (int64_t)it.leaf + 2];
  float acc[W];
#pragma unroll
  for (int j = 0; j < W; ++j) acc[j] = 0.0f;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    const float ax = xi + (float)(cxk - leafc[3 * (int64_t)nn]);
    const float ay = yi + (float)(cyk - leafc[3 * (int64_t)nn + 1]);
    const float az = zi + (float)(czk - leafc[3 * (int64_t)nn + 2]);
    for (int s = cs; s < cs + m; ++s) {
      const float dx = ax - lx[s], dy = ay - ly[s], dz = az - lz[s];
      const float R2 = dx * dx + dy * dy + dz * dz;
      const float k = s == ip ? 0.0f : KF::f(kd.q, R2);
      const float* xs = xr + (int64_t)s * W;
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] += k * xs[j];
    }
  }
  if (!act) return;
  double* yp = yNw + (int64_t)ip * W;
#pragma unroll
  for (int j = 0; j < W; ++j) yp[j] = kd.scale * (double)acc[j];
}

// y[j N + i] = alpha * (sum over levels of y_level(iperm[i], j) + y_N(iperm[i], j)) + d_i b[j N + i]
// for j < cnt: 64 points per block, transposed through LDS so that both the
// level sums (W per point) and the outputs (64 points per RHS) are contiguous.
template <int W>
__global__ __launch_bounds__(256) void combine_w_kernel(const double* __restrict__ ylevw, int nL,
                                                        int64_t N, const double* __restrict__ ynw,
                                                        const int32_t* __restrict__ iperm,
                                                        const double* __restrict__ b, int cnt,
                                                        double alpha, double diag,
                                                        const double* __restrict__ diag_vec,
                                                        double* __restrict__ y) {
  __shared__ double tile[W][65];
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  for (int q = threadIdx.x; q < 64 * W; q += 256) {
    const int ii = q / W, j = q % W;
    const int64_t i = i0 + ii;
    double acc = 0.0;
    if (i < N && j < cnt) {
      const int64_t s = iperm[i];
      for (int lv = 0; lv < nL; ++lv) acc = acc + ylevw[((int64_t)lv * N + s) * W + j];
      if (ynw) acc = acc + ynw[s * W + j];
    }
    tile[j][ii] = acc;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 64 * W; q += 256) {
    const int j = q / 64, ii = q % 64;
    const int64_t i = i0 + ii;
    if (i < N && j < cnt) {
      const double d = diag_vec ? diag_vec[i] : diag;
      y[(int64_t)j * N + i] = alpha * tile[j][ii] + d * b[(int64_t)j * N + i];
    }
  }
}

inline int64_t padded_points(int64_t N) { return ((N + 3) & ~int64_t(3)) + 16; }

// Phase-1 items grouped by source box, the partial-sum slot of every block
// whose source has more than kGRows rows (-1 otherwise), and one reduce entry
// per column of those blocks {first slot, ranges, slot stride, t index}.
void build_group_tables(Engine& e) {
  std::vector<GItem> items;
  std::vector<int32_t> gblk;  // blocks of rank >= 1, grouped by source box
  std::vector<int32_t> pbw(std::max<int64_t>(e.nblk_all, 1), -1);
  std::vector<int4> cols;
  std::vector<int32_t> btoff(std::max<int64_t>(e.nblk_all, 1), 0);
  if (e.nblk_all)
    H3D_HIP(hipMemcpy(btoff.data(), e.d_btoff, sizeof(int32_t) * e.nblk_all, hipMemcpyDeviceToHost));
  int64_t np = 0;
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    std::vector<int64_t> first(lv.nboxes + 1, 0);  // by-source offsets within the level
    for (int32_t k : lv.ilidx) ++first[k + 1];
    for (int64_t k = 0; k < lv.nboxes; ++k) first[k + 1] += first[k];
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int m = e.bcount[lo + k];
      const int64_t qb = lv.bbase + first[k], qe = lv.bbase + first[k + 1];
      if (m == 0 || qb == qe) continue;
      const int nsplit = (m + kGRows - 1) / kGRows;
      if (nsplit > 1)
        for (int64_t q = qb; q < qe; ++q) {
          const int32_t b = e.by_src[q];
          const int r = lv.rank[b - lv.bbase];
          pbw[b] = (int32_t)np;
          for (int c = 0; c < r; ++c) cols.push_back(make_int4((int)np + c, nsplit, r, btoff[b] + c));
          np += (int64_t)nsplit * r;
          if (np > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
        }
      const int64_t g0 = (int64_t)gblk.size();
      for (int64_t q = qb; q < qe; ++q)
        if (lv.rank[e.by_src[q] - lv.bbase] > 0) gblk.push_back(e.by_src[q]);
      const int64_t g1 = (int64_t)gblk.size();
      for (int s0 = 0; s0 < m; s0 += kGRows) {
        int64_t q = g0;
        while (q < g1) {
          GItem it{(int32_t)q, 0, e.bstart[lo + k], s0, std::min(m, s0 + kGRows), 0};
          int ncol = 0;
          while (q < g1) {
            const int r = lv.rank[gblk[q] - lv.bbase];
            if (ncol > 0 && ncol + r > e.g_cols) break;
            ncol += r;
            ++q;
          }
          it.q1 = (int32_t)q;
          items.push_back(it);
        }
      }
    }
  }
  std::vector<BlockDev> blk(std::max<int64_t>(e.nblk_all, 1));
  if (e.nblk_all)
    H3D_HIP(hipMemcpy(blk.data(), e.d_blk, sizeof(BlockDev) * e.nblk_all, hipMemcpyDeviceToHost));
  std::vector<GDesc> desc(gblk.size());
  for (size_t q = 0; q < gblk.size(); ++q) {
    const int32_t b = gblk[q];
    desc[q] = GDesc{blk[b].voff, blk[b].chunk, blk[b].r, blk[b].m, btoff[b], pbw[b], 0};
  }
  e.n_gitems = (int64_t)items.size();
  e.d_gitems = upload(items, e.stream);
  e.d_gblk = upload(desc, e.stream);
  e.n_rcols = (int64_t)cols.size();
  e.d_rcols = upload(cols, e.stream);
  e.g_np = np;
  H3D_HIP(hipStreamSynchronize(e.stream));
}

void ensure_block_scratch(Engine& e, int W) {
  if (!e.d_gitems) build_group_tables(e);
  const int64_t T = std::max<int64_t>(e.T, 1), np = std::max<int64_t>(e.g_np, 1);
  if (e.w_cap >= W && e.w_t >= T && e.w_np >= np) return;
  W = std::max(W, e.w_cap);
  dfree(e.d_xw);
  dfree(e.d_tw);
  dfree(e.d_partw);
  dfree(e.d_ylevw);
  dfree(e.d_ynw);
  e.w_cap = 0;
  e.d_xw = dalloc<double>(2 * padded_points(e.N) * W);
  e.d_tw = dalloc<double>(T * W);
  e.d_partw = dalloc<double>(np * W);
  e.d_ylevw = dalloc<double>(e.N * e.nL * W);
  e.d_ynw = dalloc<double>(e.N * W);
  e.w_cap = W;
  e.w_t = T;
  e.w_np = np;
}

template <class S, class A, int W>
void run_group(Engine& e, int prec, bool acc64, const double* b, double* y, int cnt) {
  hipStream_t st = e.stream;
  const int64_t N = e.N, Npad = padded_points(N);
  A* xq = (A*)e.d_xw;
  A* xr = xq + Npad * W;
  A* tw = (A*)e.d_tw;
  A* partw = (A*)e.d_partw;
  gather_w_kernel<A, W><<<(unsigned)((Npad * W + 255) / 256), 256, 0, st>>>(b, e.d_perm, N, Npad,
                                                                           cnt, xq, xr);
  H3D_LAUNCH_CHECK();
  {
    const int64_t n = N * e.nL * W;
    zero_kernel<<<(unsigned)std::min<int64_t>(8192, (n + 255) / 256), 256, 0, st>>>(e.d_ylevw, n);
    H3D_LAUNCH_CHECK();
  }
  // near field into its own buffer (W per point) on the handle's second
  // stream, overlapping phases 1 and 2 (in line while a graph is captured)
  const int nn = (int)e.nitems.size();
  const bool overlap = nn && e.overlap_near && e.side;
  auto launch_near = [&](hipStream_t ns) {
    const int64_t lo = level_offset(e.nL);
    const KernelDev kd = e.kern;
    const bool invr = kd.q == 7 || kd.q == 16;
    const unsigned g = wave_grid(nn);
    const int32_t* nbr = e.lev[e.nL].d_nbr;
    if constexpr (std::is_same<A, float>::value) {
      if (invr)
        near_w32_kernel<InvR32, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                      e.d_bcount + lo, e.d_lx, e.d_ly, e.d_lz,
                                                      e.d_leafc, kd, xr, e.d_ynw);
      else
        near_w32_kernel<AnyQ32, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                      e.d_bcount + lo, e.d_lx, e.d_ly, e.d_lz,
                                                      e.d_leafc, kd, xr, e.d_ynw);
    } else {
      if (kd.q == kQPlugin)
        near_w64_kernel<Plugin, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                      e.d_bcount + lo, e.d_px, e.d_py, e.d_pz,
                                                      kd, xr, e.d_ynw);
      else if (invr && prec != H3D_PREC_FP64)
        near_w64_kernel<InvRNewton, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                          e.d_bcount + lo, e.d_px, e.d_py,
                                                          e.d_pz, kd, xr, e.d_ynw);
      else if (invr)
        near_w64_kernel<InvR, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                    e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, kd,
                                                    xr, e.d_ynw);
      else
        near_w64_kernel<AnyQ, W><<<g, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                    e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, kd,
                                                    xr, e.d_ynw);
    }
    H3D_LAUNCH_CHECK();
  };
  if (overlap) {
    H3D_HIP(hipEventRecord(e.ev_fork, st));
    H3D_HIP(hipStreamWaitEvent(e.side, e.ev_fork, 0));
    launch_near(e.side);
    H3D_HIP(hipEventRecord(e.ev_join, e.side));
  }
  const Plan& p = e.plan;
  if (e.n_gitems) {
    p1g_kernel<S, A, W><<<(unsigned)e.n_gitems, 256, 0, st>>>(
        (const GItem*)e.d_gitems, (int)e.n_gitems, (const GDesc*)e.d_gblk, e.d_vbase[prec], xq,
        tw, partw);
    H3D_LAUNCH_CHECK();
  }
  if (e.n_rcols) {
    reduce_w_kernel<A, W><<<(unsigned)((e.n_rcols * W + 255) / 256), 256, 0, st>>>(
        (const int4*)e.d_rcols, e.n_rcols, partw, tw);
    H3D_LAUNCH_CHECK();
  }
  if (!p.p2.empty()) {
    p2w_kernel<S, A, W><<<wave_grid(p.p2.size()), 256, 0, st>>>(p.d_p2, (int)p.p2.size(),
                                                                e.d_ubase[prec], tw, e.d_ylevw, N);
    H3D_LAUNCH_CHECK();
  }
  if (nn && !overlap) launch_near(st);
  if (overlap) H3D_HIP(hipStreamWaitEvent(st, e.ev_join, 0));
  (void)acc64;
  combine_w_kernel<W><<<(unsigned)((N + 63) / 64), 256, 0, st>>>(
      e.d_ylevw, e.nL, N, nn ? e.d_ynw : nullptr, e.d_iperm, b, cnt, e.alpha, e.kern.diag,
      e.d_diag, y);
  H3D_LAUNCH_CHECK();
}

template <int W>
void run_group_w(Engine& e, int prec, bool acc64, const double* b, double* y, int cnt) {
  if (prec == H3D_PREC_FP64) run_group<double, double, W>(e, prec, acc64, b, y, cnt);
  else if (prec == H3D_PREC_FP32 && acc64) run_group<float, double, W>(e, prec, acc64, b, y, cnt);
  else if (prec == H3D_PREC_FP32) run_group<float, float, W>(e, prec, acc64, b, y, cnt);
  else if (acc64) run_group<uint16_t, double, W>(e, prec, acc64, b, y, cnt);
  else run_group<uint16_t, float, W>(e, prec, acc64, b, y, cnt);
}

}  // namespace

void release_group_tables(Engine& e) {
  dfree(e.d_gitems);
  dfree(e.d_gblk);
  dfree(e.d_rcols);
  e.n_gitems = e.n_rcols = e.g_np = 0;
}

int block_width(const Engine& e, int nrhs, int prec, bool transpose, bool acc64) {
  if (transpose || prec == kPrecMixed || e.rhs_block <= 0 || nrhs < 2) return 1;
  // the grouped fp32-sum near field works on fp32 offsets of radial kernels
  if (e.kern.q == kQPlugin && prec != H3D_PREC_FP64 && !acc64) return 1;
  return nrhs <= 16 || e.rhs_block == 16 ? 16 : 32;
}

void matvec_block(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool acc64) {
  const int Wmax = e.rhs_block == 16 ? 16 : 32;
  ensure_block_scratch(e, nrhs <= 16 ? 16 : Wmax);
  const int64_t N = e.N;
  for (int r0 = 0; r0 < nrhs; r0 += Wmax) {
    const int cnt = std::min(Wmax, nrhs - r0);
    const double* b = d_b + (int64_t)r0 * N;
    double* y = d_y + (int64_t)r0 * N;
    if (cnt <= 16) run_group_w<16>(e, prec, acc64, b, y, cnt);
    else run_group_w<32>(e, prec, acc64, b, y, cnt);
  }
}

void autotune_rhs(Engine& e, int mode, int nrhs, int reps, double* ms, int* best) {
  const int prec = mode & 0xff;
  if (prec < 0 || prec > 2 || (mode & ~(0xff | H3D_ACC_FP64)))
    throw Error(H3D_ERR_INVALID, "grouped products take fp64, fp32 or bf16 factor copies");
  if (nrhs < 2) throw Error(H3D_ERR_INVALID, "nrhs must be >= 2");
  if (reps < 1) throw Error(H3D_ERR_INVALID, "reps must be >= 1");
  const bool acc64 = prec == H3D_PREC_FP64 || (mode & H3D_ACC_FP64);
  prepare_mode(e, mode);
  const int64_t N = e.N;
  double* B = dalloc<double>(N * nrhs);
  double* Y = dalloc<double>(N * nrhs);
  hipEvent_t ev[2] = {nullptr, nullptr};
  const int saved_w = e.rhs_block, saved_c = e.g_cols;
  const int widths[2] = {16, 32}, cols[4] = {256, 512, 1024, 2048};
  int bw = saved_w > 0 ? saved_w : 32, bc = saved_c;
  try {
    H3D_HIP(hipMemsetAsync(B, 0, sizeof(double) * N * nrhs, e.stream));
    for (auto& x : ev) H3D_HIP(hipEventCreate(&x));
    double bt = 1e300;
    for (int wi = 0; wi < 2; ++wi)
      for (int ci = 0; ci < 4; ++ci) {
        if (ms) ms[wi * 4 + ci] = -1.0;
        if (wi == 1 && nrhs <= 16) continue;  // groups of <= 16 run at width 16 anyway
        e.rhs_block = widths[wi];
        if (e.g_cols != cols[ci]) {
          H3D_HIP(hipStreamSynchronize(e.stream));
          release_group_tables(e);
          e.g_cols = cols[ci];
        }
        matvec_block(e, B, Y, nrhs, prec, acc64);  // warm-up (builds the tables)
        H3D_HIP(hipEventRecord(ev[0], e.stream));
        for (int q = 0; q < reps; ++q) matvec_block(e, B, Y, nrhs, prec, acc64);
        H3D_HIP(hipEventRecord(ev[1], e.stream));
        H3D_HIP(hipEventSynchronize(ev[1]));
        float t = 0;
        H3D_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
        const double avg = t / reps;
        if (ms) ms[wi * 4 + ci] = avg;
        if (avg < bt) {
          bt = avg;
          bw = widths[wi];
          bc = cols[ci];
        }
      }
  } catch (...) {
    (void)hipStreamSynchronize(e.stream);
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
    dfree(B);
    dfree(Y);
    e.rhs_block = saved_w;
    throw;
  }
  H3D_HIP(hipStreamSynchronize(e.stream));
  for (auto& x : ev) (void)hipEventDestroy(x);
  dfree(B);
  dfree(Y);
  e.rhs_block = bw;
  if (e.g_cols != bc) {
    release_group_tables(e);
    e.g_cols = bc;
  }
  if (best) {
    best[0] = bw;
    best[1] = bc;
  }
}

}  // namespace h3d
