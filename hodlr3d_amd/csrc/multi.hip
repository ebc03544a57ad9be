// Products with many right-hand sides: the RHS are taken W (16 or 32) at a
// time and every factor element is read once per group and applied to all W
// of them on the matrix cores (the caller side of MatVecProduct,
// src/HODLR3D.cpp:21-27, called once per column in the reference).
//
//   gather    x by position, the W values of one point contiguous (read by
//             phase 1, staged per row range, and by the near field)
//   phase 1   T_b = V_b^T X_s   per source box s: the V columns of every block
//             with that source concatenated, 32-column tiles x 16-row steps,
//             16x16x4 MFMA: A = 16 columns x 4 rows of V (one 4-row vector
//             per lane), B = 4 rows x 16 RHS of x (from LDS); sources of more
//             than 256 rows in 256-row ranges summed afterwards (reduce)
//   phase 2   Y_k = U_k T_k     per <= 64 rows of a target box, 16x16x4 MFMA:
//             A = 16 row quads x 4 columns of U, B = 4 columns x 16 RHS of T
//   near      y_N = K_near X    fp32 sums: entries of 4 sources x 64 targets
//             per step on the VALU, multiplied on the matrix cores (sources
//             staged in LDS); fp64 sums: lanes own target rows, W sums each;
//             own buffer, on the handle's second stream under phases 1 and 2
//   combine   y[:, j] = alpha * (sum_levels + y_N) + d * b[:, j] (an LDS transpose)
//
// The streaming loops of phases 1 and 2 issue the same loads (and, in phase
// 1, the same stores) on every pass, with no branch that skips one: on gfx950
// loads and stores share one in-order counter, and where two paths with
// different operations in flight meet, the compiler must assume the fewest,
// which made a wave wait for its prefetched loads (and stores) before the
// next MFMAs.  fp64 factors and the fp64-accumulation modes use the fp64
// MFMA, fp32/bf16 factors with fp32 accumulation the fp32 MFMA.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "engine.hpp"
#include "kernels.hpp"
#include "stream.hpp"

namespace h3d {
namespace {

constexpr int kGRows = 256;     // rows of one phase-1 item (x staged in LDS)
constexpr int kTileCols = 32;   // columns of one phase-1 tile (two MFMA tiles)
constexpr int64_t kColMask = (int64_t(1) << 48) - 1;
// chunk bases a phase-1 workgroup keeps in LDS (a handle with more factor
// chunks takes its products one column at a time)
constexpr int kLdsChunks = 1024;
static_assert(kLdsChunks % 256 == 0, "staged by 256 threads");

// Waves per SIMD the streaming kernels are compiled for: with fp32 sums four
// (at most 128 registers, no spills), which keeps four 8 KB load batches in
// flight per SIMD, about what HBM needs; fp64 sums need more registers.
template <class A>
constexpr int kWavesFor = sizeof(A) == 4 ? 4 : 1;

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

// Four consecutive stored elements (4-byte aligned) converted to A, with the
// nontemporal hint (NT: the load's lines are not kept in the cache) or
// without (a line that the next instruction reads again stays there).
// (global address space named: a pointer read from LDS would otherwise give
// flat loads, which count on both wait counters)
template <bool NT, class T>
__device__ inline T vload(const void* p) {
  if constexpr (NT) return ntload<T>(p);
  else return *(const __attribute__((address_space(1))) T*)p;
}
template <class S, class A, bool NT = true>
__device__ inline void load4(const S* p, A (&o)[4]) {
  if constexpr (sizeof(S) == 8) {
    const u32x4a v0 = vload<NT, u32x4a>(p), v1 = vload<NT, u32x4a>(p + 2);
    o[0] = (A)__hiloint2double((int)v0.y, (int)v0.x);
    o[1] = (A)__hiloint2double((int)v0.w, (int)v0.z);
    o[2] = (A)__hiloint2double((int)v1.y, (int)v1.x);
    o[3] = (A)__hiloint2double((int)v1.w, (int)v1.z);
  } else if constexpr (sizeof(S) == 4) {
    const u32x4a v = vload<NT, u32x4a>(p);
    o[0] = (A)__uint_as_float(v.x);
    o[1] = (A)__uint_as_float(v.y);
    o[2] = (A)__uint_as_float(v.z);
    o[3] = (A)__uint_as_float(v.w);
  } else if constexpr (sizeof(S) == 3) {  // fp24: 12 bytes, four packed values
    const u32x2a v = vload<NT, u32x2a>(p);
    const uint32_t w = vload<NT, uint32_t>((const char*)p + 8);
    uint32_t u[4];
    f24x4_bits(v.x, v.y, w, u);
    o[0] = (A)__uint_as_float(u[0]), o[1] = (A)__uint_as_float(u[1]);
    o[2] = (A)__uint_as_float(u[2]), o[3] = (A)__uint_as_float(u[3]);
  } else if constexpr (std::is_same<S, i8s>::value) {  // int8 (column-scaled)
    const uint32_t v = vload<NT, uint32_t>(p);
    float f[4];
    i8x4_to_f(v, f);
    o[0] = (A)f[0], o[1] = (A)f[1], o[2] = (A)f[2], o[3] = (A)f[3];
  } else if constexpr (sizeof(S) == 1) {  // fp8 (e4m3)
    const uint32_t v = vload<NT, uint32_t>(p);
    float a, b, c, d;
    f8x2_to_f<false>(v, a, b);
    f8x2_to_f<true>(v, c, d);
    o[0] = (A)a, o[1] = (A)b, o[2] = (A)c, o[3] = (A)d;
  } else if constexpr (std::is_same<S, _Float16>::value) {  // IEEE half
    const u32x2a v = vload<NT, u32x2a>(p);
    o[0] = (A)(float)__builtin_bit_cast(_Float16, (uint16_t)(v.x & 0xffffu));
    o[1] = (A)(float)__builtin_bit_cast(_Float16, (uint16_t)(v.x >> 16));
    o[2] = (A)(float)__builtin_bit_cast(_Float16, (uint16_t)(v.y & 0xffffu));
    o[3] = (A)(float)__builtin_bit_cast(_Float16, (uint16_t)(v.y >> 16));
  } else {  // bf16: element 0 in the low half of each word
    const u32x2a v = vload<NT, u32x2a>(p);
    o[0] = (A)__uint_as_float(v.x << 16);
    o[1] = (A)__uint_as_float(v.x & 0xffff0000u);
    o[2] = (A)__uint_as_float(v.y << 16);
    o[3] = (A)__uint_as_float(v.y & 0xffff0000u);
  }
}

// Four consecutive x values of one RHS in the LDS quad layout (16 or 32 bytes).
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

// x(p, j) at xr[p W + j]; columns j >= cnt are zero.
template <class A, int W>
__global__ void gather_w_kernel(const double* __restrict__ b, const int32_t* __restrict__ perm,
                                int64_t N, int cnt, A* __restrict__ xr) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= N * W) return;
  const int j = (int)(id % W);
  xr[id] = j < cnt ? (A)b[(int64_t)j * N + perm[id / W]] : (A)0;
}

// Phase 1: one workgroup per item = a row range of at most NC x 64 rows of one
// source box (the whole box if it has at most 256 rows, else one 256-row range
// of it) x at most LC columns of the box's column list (a multiple of 32: the
// table pads every source).  x of the range (zero past it) and the item's
// column table go to LDS once; then every wave streams 32-column tiles
// k = wave, wave + 4, .., each a complete unit of NC chunks of 4 steps of 16
// rows: lane (c = l % 16, q = l / 16) holds rows 16 s + 4q .. + 3 of column c
// of each 16-column half, x of the same rows comes from LDS (16x16x4 MFMA:
// A = 16 columns x 4 rows of V, B = 4 rows x 16 RHS).  A tile's sums are
// stored (W per column) at its slot out0 + range * step: a t index, a partial
// slot (sources of more than 256 rows, after the T + 1 t slots), or the spare
// t slot T (padding).
//
// Every pass issues the same loads and stores and no branch skips one (rows
// past the range are moved to row 0 and meet zero x; tiles past the wave's
// last are clamped to it and their sums never stored), the next chunk's loads
// are issued before the current chunk's MFMAs (two register buffers; the
// scheduling barrier keeps each batch whole), and the first tile (NC > 1) or
// first pair of tiles (NC = 1) is peeled off so that the loop is entered with
// the same operations in flight as it loops.  The compiler's waits then count
// exactly the loads each chunk needs: on gfx950 loads and stores share one
// in-order counter, and where two paths with different operations in flight
// meet, the compiler must assume the fewest, which made a wave wait for its
// prefetched loads and its stores.
template <class S, class A, int W, int NC, int LC>
__global__ __launch_bounds__(256, kWavesFor<A>) void p1r_kernel(const GTItem* __restrict__ items, int nitems,
                                                  const int64_t* __restrict__ colv,
                                                  const int2* __restrict__ colo,
                                                  void* const* __restrict__ vbase, int nchunk,
                                                  const A* __restrict__ xr, A* __restrict__ tw) {
  using M = Mfma<A>;
  using Acc = typename M::acc;
  constexpr int H = W / 16;
  constexpr int CT = 2;
  constexpr int ROWS = 64 * NC;
  constexpr int PER = LC / 256;  // table entries staged per thread
  constexpr int VE = 16 / sizeof(A);
  constexpr int XV = ROWS * W / VE;  // 16-byte vectors of x
  constexpr int XPER = XV / 256;
  static_assert(LC % 256 == 0 && XV % 256 == 0, "whole staging passes");
  static_assert(NC == 1 || NC % 2 == 0, "chunks alternate between two load buffers");
  typedef A vec __attribute__((ext_vector_type(VE)));
  extern __shared__ __align__(16) unsigned char p1r_lds[];  // ROWS W values of x
  A* xl = (A*)p1r_lds;
  __shared__ int64_t lcv[LC];
  __shared__ int32_t lslot[LC];
  __shared__ const void* cbase[kLdsChunks];
  const int item = blockIdx.x;
  if (item >= nitems) return;  // whole workgroup
  const GTItem it = items[item];
  const int tid = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int lc = lane & 15, kq = lane >> 4;

  // 1. staging: every load issued before any is used, at clamped indices; the
  // chunk bases are read from LDS only (a select between an LDS and a global
  // address gives flat loads, which wait for everything in flight)
  {
    int64_t tcv[PER];
    int32_t tsl[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int c = tid + 256 * q;
      const int cc = it.col0 + (c < it.ncols ? c : it.ncols - 1);
      tcv[q] = colv[cc];
      const int2 o = colo[cc];
      tsl[q] = o.x + it.range * o.y;
    }
    const void* cb[kLdsChunks / 256];
#pragma unroll
    for (int q = 0; q < kLdsChunks / 256; ++q) {
      const int c = tid + 256 * q;
      cb[q] = vbase[c < nchunk ? c : nchunk - 1];
    }
    const int nvalid = it.nrows * W / VE;
    const A* src = xr + (int64_t)it.a0 * W;
    vec tmp[XPER];
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int v = tid + 256 * q;
      tmp[q] = *(const vec*)(src + (int64_t)(v < nvalid ? v : 0) * VE);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      lcv[tid + 256 * q] = tcv[q];
      lslot[tid + 256 * q] = tsl[q];
    }
#pragma unroll
    for (int q = 0; q < kLdsChunks / 256; ++q)
      if (tid + 256 * q < nchunk) cbase[tid + 256 * q] = cb[q];
#pragma unroll
    for (int q = 0; q < XPER; ++q) {
      const int v = tid + 256 * q;
#pragma unroll
      for (int u = 0; u < VE; ++u) {
        const int e = v * VE + u, i = e / W, j = e % W;
        xl[((i >> 2) * W + j) * 4 + (i & 3)] = v < nvalid ? tmp[q][u] : (A)0;
      }
    }
  }
  __syncthreads();

  // 2. the wave's tiles (a wave without one has nothing in flight here)
  const int ntiles = it.ncols / kTileCols;
  const int ntw = ntiles > wv ? (ntiles - wv + 3) / 4 : 0;
  if (ntw == 0) return;
  auto load = [&](int i, int c, A (&av)[4][CT][4]) {  // chunk c of the wave's tile i (clamped)
    const int k = wv + 4 * (i < ntw ? i : ntw - 1);
    const S* vp[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int64_t cv = lcv[k * kTileCols + 16 * ct + lc];
      vp[ct] = (const S*)cbase[(int)(cv >> 48)] + (cv & kColMask) + it.s0;
    }
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int r = 64 * c + 16 * st + 4 * kq;
      const int rl = r < it.nrows ? r : 0;
      // cached loads: a column's 128-byte line is read by two instructions
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) load4<S, A, false>(vp[ct] + rl, av[st][ct]);
    }
    __builtin_amdgcn_sched_barrier(0);  // the batch stays ahead of the MFMAs
  };
  Acc acc[CT][H];
  auto clear = [&]() {
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int h = 0; h < H; ++h) acc[ct][h] = Acc{};
  };
  auto compute = [&](int c, const A (&av)[4][CT][4]) {
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      A xb[H][4];
#pragma unroll
      for (int h = 0; h < H; ++h)
        load_quad<A>(xl + ((4 * (4 * c + st) + kq) * W + 16 * h + lc) * 4, xb[h]);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int h = 0; h < H; ++h) {
            // fp32: the transposed tile (A = x, B = V), so that a lane holds
            // four consecutive RHS of one column and stores them as one
            // 16-byte vector (four stores per tile instead of sixteen)
            if constexpr (sizeof(A) == 4) acc[ct][h] = M::mac(xb[h][e], av[st][ct][e], acc[ct][h]);
            else acc[ct][h] = M::mac(av[st][ct][e], xb[h][e], acc[ct][h]);
          }
    }
  };
  auto store = [&](int i) {
    const int k = wv + 4 * i;
    if constexpr (sizeof(A) == 4) {
      // element a of acc[ct][h]: RHS 16 h + 4 kq + a of the tile's column 16 ct + lc
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        A* dst = tw + (int64_t)lslot[k * kTileCols + 16 * ct + lc] * W + 4 * kq;
#pragma unroll
        for (int h = 0; h < H; ++h) *(Acc*)(dst + 16 * h) = acc[ct][h];
      }
    } else {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          A* dst = tw + (int64_t)lslot[k * kTileCols + 16 * ct + M::row(kq, a)] * W;
#pragma unroll
          for (int h = 0; h < H; ++h) dst[16 * h + lc] = acc[ct][h][a];
        }
    }
  };
  A av0[4][CT][4], av1[4][CT][4];
  if constexpr (NC == 1) {
    // tile 0, then pairs (1, 2), (3, 4), .. with no exit inside a pair (an
    // exit there reads, to the compiler, as a second way back to the loop
    // head with the other buffer's loads in flight), then the odd tile left
    load(0, 0, av0);
    load(1, 0, av1);
    clear();
    compute(0, av0);
    store(0);
    const int npairs = (ntw - 1) / 2;
    for (int p = 0; p < npairs; ++p) {
      const int i = 2 * p + 1;
      load(i + 1, 0, av0);
      clear();
      compute(0, av1);
      store(i);
      load(i + 2, 0, av1);
      clear();
      compute(0, av0);
      store(i + 1);
    }
    if ((ntw - 1) & 1) {  // tile ntw - 1, loaded into av1
      clear();
      compute(0, av1);
      store(ntw - 1);
    }
  } else {
    // a tile's chunks 0, 1, .., NC - 1 in av0, av1, av0, ..; the last one
    // prefetches chunk 0 of the next tile
    auto tile = [&](int i) {
      clear();
#pragma unroll
      for (int c = 0; c < NC; c += 2) {
        load(i, c + 1, av1);
        compute(c, av0);
        if (c + 2 < NC) load(i, c + 2, av0);  // compile-time branch
        else load(i + 1, 0, av0);
        compute(c + 1, av1);
      }
      store(i);
    };
    load(0, 0, av0);
    tile(0);
    for (int i = 1; i < ntw; ++i) tile(i);
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
// the MFMA output row).  CT (the transposed product, y_s = sum_b V_b t_b over
// the blocks with source s): column c of the item is column table entry
// it.uoff + c (element offset | chunk << 48 in `ubase`'s chunks) with its t
// index in p2t; otherwise column c of U at ubase[it.chunk] + it.uoff + c ld and
// t index it.toff + c.  With `scale` (the mixed copy's classes) t of slot q is
// multiplied by scale[q] as it is loaded; `accumulate` adds to the level sums
// (the second class) instead of storing them.
template <class S, class A, int W, bool CT = false>
__global__ __launch_bounds__(256, kWavesFor<A>) void p2w_kernel(const P2Item* __restrict__ items, int nitems,
                                                  void* const* __restrict__ ubase,
                                                  const A* __restrict__ tw,
                                                  double* __restrict__ ylevw, int64_t N,
                                                  const int64_t* __restrict__ p2col,
                                                  const int32_t* __restrict__ p2t,
                                                  const double* __restrict__ scale, int accumulate) {
  using M = Mfma<A>;
  using Acc = typename M::acc;
  constexpr int H = W / 16;
  constexpr int UF = sizeof(A) == 8 ? 2 : 4;
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const int lc = lane & 15, kq = lane >> 4;
  const P2Item it = items[w];
  const S* U = CT ? nullptr : (const S*)ubase[it.chunk] + it.uoff;
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
        if constexpr (CT) {
          const int64_t cv = p2col[it.uoff + cl];
          const int64_t tslot = p2t[it.uoff + cl];
#pragma unroll
          for (int h = 0; h < H; ++h) tb[st][h] = tw[tslot * W + 16 * h + lc];
          load4<S, A>((const S*)ubase[(int)(cv >> 48)] + (cv & kColMask) + il, av[st]);
        } else {
          const A sc = scale ? (A)scale[it.toff + cl] : (A)1;
#pragma unroll
          for (int h = 0; h < H; ++h) tb[st][h] = tw[(int64_t)(it.toff + cl) * W + 16 * h + lc] * sc;
          load4<S, A>(U + (int64_t)cl * it.ld + il, av[st]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);  // the batch stays ahead of the MFMAs
    };
    Acc acc[4][H];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int h = 0; h < H; ++h) acc[u][h] = Acc{};
    auto compute = [&](int c0, const A (&av)[UF][4], const A (&tb)[UF][H]) {
#pragma unroll
      for (int st = 0; st < UF; ++st) {
        // every step runs (a skipped one would leave its loads unwaited on
        // one path, and the compiler then waits for everything in flight at
        // the loop head): only t is masked (clamped columns give zero terms);
        // rows past the box give output rows that are never stored
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
    // loads on every pass (columns clamped, so past the end they reread the
    // last column): the wait before a step's MFMAs then counts only its own
    // loads, not the prefetched ones
    // two steps per pass and no exit between them (an exit there reads, to
    // the compiler, as a second way back to the loop head with loads in
    // flight); a pass past R adds zeros
    load(0, av0, tb0);
    for (int c0 = 0; c0 < it.R; c0 += 2 * STEP) {
      load(c0 + STEP, av1, tb1);
      compute(c0, av0, tb0);
      load(c0 + 2 * STEP, av0, tb0);
      compute(c0 + STEP, av1, tb1);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = it.i0 + r0 + 4 * M::row(kq, a) + u;
        if (row < lim) {
#pragma unroll
          for (int h = 0; h < H; ++h) {
            double* yp = yb + (int64_t)row * W + 16 * h + lc;
            *yp = accumulate ? *yp + (double)acc[u][h][a] : (double)acc[u][h][a];
          }
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

// Near field with fp32 entries and sums on the matrix cores: Y(t, j) = sum_s
// K(t, s) X(s, j) for the <= 64 targets t of an item.  The sources of each
// neighbour leaf are staged 64 at a time in the wave's LDS slice (coordinates
// as fp32 offsets from their leaf centre, and their W values); each step takes
// 4 of them: lane (c = l % 16, q = l / 16) forms K(target 16 mt + c, source
// s0 + q) for the four 16-target tiles mt on the VALU (the A operands), reads
// X(source s0 + q, RHS 16 h + c) from LDS (the B operands), and the 16x16x4
// MFMA accumulates D(target, RHS); lane l then holds targets 16 mt + 4 q + a.
// Every tile is multiplied on every step (rows past the item are not stored),
// sources past the chunk and the self pair meet a zero entry.  Each sum runs
// over the sources in the same order whatever the group width.
template <class KF, int W>
__global__ __launch_bounds__(256) void near_wm_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const float* __restrict__ lx, const float* __restrict__ ly, const float* __restrict__ lz,
    const double* __restrict__ leafc, KernelDev kd, const float* __restrict__ xr,
    double* __restrict__ yNw) {
  using M = Mfma<float>;
  using Acc = M::acc;
  constexpr int H = W / 16;
  // row stride of the staged values: 16 banks apart from one source to the
  // next, so that the two 16-lane groups of a half-wave read distinct banks
  constexpr int XS = W % 32 == 0 ? W + 16 : W;
  __shared__ float4 scoord[kWavesPerBlock][64];
  __shared__ __align__(16) float sval[kWavesPerBlock][64 * XS];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63, lc = lane & 15, kq = lane >> 4;
  const NearItem it = items[w];
  const int rs = bstart[it.leaf];
  // the lane's targets (clamped to the item's rows; the extra rows are not stored)
  float tx[4], ty[4], tz[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int t = min(16 * mt + lc, it.nr - 1);
    const int ip = rs + it.i0 + t;
    tx[mt] = lx[ip], ty[mt] = ly[ip], tz[mt] = lz[ip];
  }
  const double cxk = leafc[3 * (int64_t)it.leaf], cyk = leafc[3 * (int64_t)it.leaf + 1],
               czk = leafc[3 * (int64_t)it.leaf + 2];
  Acc acc[4][H];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int h = 0; h < H; ++h) acc[mt][h] = Acc{};
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    const float dcx = (float)(cxk - leafc[3 * (int64_t)nn]);
    const float dcy = (float)(cyk - leafc[3 * (int64_t)nn + 1]);
    const float dcz = (float)(czk - leafc[3 * (int64_t)nn + 2]);
    float ax[4], ay[4], az[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) ax[mt] = tx[mt] + dcx, ay[mt] = ty[mt] + dcy, az[mt] = tz[mt] + dcz;
    const bool own = nn == it.leaf;
    for (int c0 = 0; c0 < m; c0 += 64) {
      const int cnt = min(64, m - c0);
      {  // stage the chunk: lane i takes source c0 + i (clamped)
        const int jp = cs + c0 + (lane < cnt ? lane : cnt - 1);
        const float4 pc = make_float4(lx[jp], ly[jp], lz[jp], 0.0f);
        const float4* src = (const float4*)(xr + (int64_t)jp * W);
        float4 v[W / 4];
#pragma unroll
        for (int u = 0; u < W / 4; ++u) v[u] = src[u];
        __builtin_amdgcn_wave_barrier();  // after the previous chunk's reads
        scoord[wv][lane] = pc;
#pragma unroll
        for (int u = 0; u < W / 4; ++u) *(float4*)&sval[wv][lane * XS + 4 * u] = v[u];
        __builtin_amdgcn_wave_barrier();  // before this chunk's reads
      }
      for (int s0 = 0; s0 < cnt; s0 += 4) {
        const int j = s0 + kq;
        const bool valid = j < cnt;
        const int jc = valid ? j : cnt - 1;
        const float4 p = scoord[wv][jc];
        float xb[H];
#pragma unroll
        for (int h = 0; h < H; ++h) xb[h] = sval[wv][jc * XS + 16 * h + lc];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const float dx = ax[mt] - p.x, dy = ay[mt] - p.y, dz = az[mt] - p.z;
          const float R2 = dx * dx + dy * dy + dz * dz;
          float kv = KF::f(kd.q, R2);
          const bool self = own && c0 + j == it.i0 + 16 * mt + lc;
          kv = (valid && !self) ? kv : 0.0f;
#pragma unroll
          for (int h = 0; h < H; ++h) acc[mt][h] = M::mac(kv, xb[h], acc[mt][h]);
        }
      }
    }
  }
  double* yb = yNw + (int64_t)(rs + it.i0) * W;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int t = 16 * mt + M::row(kq, a);
      if (t < it.nr) {
#pragma unroll
        for (int h = 0; h < H; ++h) yb[(int64_t)t * W + 16 * h + lc] = kd.scale * (double)acc[mt][h][a];
      }
    }
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

// Phase-1 items of one group of columns [c0, c1) (a multiple of kTileCols)
// whose rows are the m points at positions a0 ..: every group of at most
// kGRows rows whole, a larger one in ranges of kGRows rows, each range in the
// class of 64 NC rows that holds it (NC = 1, 2, 4), in slices of g_cols / NC
// columns.
void p1_items(std::vector<GTItem> (&items)[3], int64_t c0, int64_t c1, int m, int32_t a0, int gc) {
  for (int s0 = 0; s0 < m; s0 += kGRows) {
    const int nr = std::min(kGRows, m - s0);
    const int cls = nr <= 64 ? 0 : (nr <= 128 ? 1 : 2);
    const int lcols = gc >> cls;
    for (int64_t cc = c0; cc < c1; cc += lcols)
      items[cls].push_back(GTItem{(int32_t)cc, (int32_t)std::min<int64_t>(lcols, c1 - cc), a0 + s0,
                                  nr, s0, s0 / kGRows, 0, 0});
  }
}

// Appends one column to the phase-1 table: its factor column (element offset
// | chunk << 48) and output slot {out0, step}: t index t (and step 0) for a
// group of at most kGRows rows, else partial slot pbase + np + c with step
// ncols (the partial sums of range q of column c at pbase + np + q ncols + c)
// and a reduce entry {np + c, nsplit, ncols, t}.
struct P1Table {
  std::vector<int64_t> colv;
  std::vector<int2> colo;
  std::vector<int4> cols;
  int64_t np = 0;
  void add(int32_t chunk, int64_t off, int nsplit, int64_t pbase, int c, int ncols, int64_t t) {
    if (chunk >= (1 << 15) || off > kColMask) throw Error(H3D_ERR_INVALID, "problem too large");
    colv.push_back(((int64_t)chunk << 48) | off);
    if (nsplit > 1) {
      colo.push_back(make_int2((int)(pbase + np + c), ncols));
      cols.push_back(make_int4((int)(np + c), nsplit, ncols, (int)t));
    } else {
      colo.push_back(make_int2((int)t, 0));
    }
  }
  // pads the group begun at c0 to a multiple of kTileCols with copies of its
  // last column writing the spare t slot T
  void pad(int64_t c0, int32_t spare) {
    while ((colv.size() - c0) % kTileCols) {
      colv.push_back(colv.back());
      colo.push_back(make_int2(spare, 0));
    }
    if ((int64_t)colv.size() > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
  }
};

// Phase 1 forward: the V columns of every block of rank >= 1, grouped by
// source box in by-source block order (x of the source staged once per item).
// Phase 2 forward: every target box in items of 128 rows (its U is one
// column-major matrix).  Transposed: phase 1 over the U columns of every
// target box (U_k^T x_k, the box's rows staged once), phase 2 over the rows
// of every source box, columns from a table: the V columns of the blocks with
// that source (in different chunks) and their t indices.  Items of phase 2
// largest first in both directions.
void build_group_tables(Engine& e, int dir) {
  GroupTables& g = e.grp[dir];
  g.free_all();
  std::vector<GTItem> items[3];
  std::vector<int32_t> btoff(std::max<int64_t>(e.nblk_all, 1), 0);
  std::vector<BlockDev> blk(std::max<int64_t>(e.nblk_all, 1));
  if (e.nblk_all) {
    H3D_HIP(hipMemcpy(btoff.data(), e.d_btoff, sizeof(int32_t) * e.nblk_all, hipMemcpyDeviceToHost));
    H3D_HIP(hipMemcpy(blk.data(), e.d_blk, sizeof(BlockDev) * e.nblk_all, hipMemcpyDeviceToHost));
  }
  P1Table tab;
  tab.colv.reserve(e.rank_sum + e.rank_sum / 16);
  tab.colo.reserve(e.rank_sum + e.rank_sum / 16);
  const int gc = e.g_cols == 1024 ? 1024 : 2048;
  const int64_t pbase = (int64_t)e.T + 1;  // first partial slot
  std::vector<P2Item> p2;
  std::vector<int64_t> p2col;
  std::vector<int32_t> p2t;
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    std::vector<int64_t> first(lv.nboxes + 1, 0);  // by-source offsets within the level
    for (int32_t k : lv.ilidx) ++first[k + 1];
    for (int64_t k = 0; k < lv.nboxes; ++k) first[k + 1] += first[k];
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];  // rows of box k (as target or as source)
      if (n == 0) continue;
      const int nsplit = (n + kGRows - 1) / kGRows;
      const int64_t c0 = (int64_t)tab.colv.size();
      if (dir == 0) {  // V columns of the blocks with source k
        const int64_t ldv = n + (n & 1);
        for (int64_t q = lv.bbase + first[k]; q < lv.bbase + first[k + 1]; ++q) {
          const int32_t b = e.by_src[q];
          const int r = lv.rank[b - lv.bbase];
          for (int c = 0; c < r; ++c)
            tab.add(blk[b].chunk, blk[b].voff + c * ldv, nsplit, pbase, c, r, btoff[b] + c);
          if (nsplit > 1 && r) tab.np += (int64_t)nsplit * r;
        }
      } else {  // U columns of target box k
        const int R = lv.Rt[k];
        const int64_t ld = aca_ldu(n);
        for (int c = 0; c < R; ++c)
          tab.add(lv.chunk[k], lv.uoff[k] + c * ld, nsplit, pbase, c, R, lv.toff[k] + c);
        if (nsplit > 1 && R) tab.np += (int64_t)nsplit * R;
      }
      if (pbase + tab.np > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
      if ((int64_t)tab.colv.size() == c0) continue;
      tab.pad(c0, e.T);
      p1_items(items, c0, (int64_t)tab.colv.size(), n, e.bstart[lo + k], gc);
    }
    // phase 2
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      if (n <= 0) continue;
      P2Item it{};
      if (dir == 0) {
        const int R = lv.Rt[k];
        if (R <= 0) continue;
        it.uoff = lv.uoff[k];
        it.chunk = lv.chunk[k];
        it.ld = aca_ldu(n);
        it.R = R;
        it.toff = lv.toff[k];
      } else {  // the columns of V_s: blocks with source k
        const int64_t ldv = n + (n & 1);
        const int64_t c0 = (int64_t)p2col.size();
        for (int64_t q = lv.bbase + first[k]; q < lv.bbase + first[k + 1]; ++q) {
          const int32_t b = e.by_src[q];
          const int r = lv.rank[b - lv.bbase];
          for (int c = 0; c < r; ++c) {
            p2col.push_back(((int64_t)blk[b].chunk << 48) | (blk[b].voff + c * ldv));
            p2t.push_back(btoff[b] + c);
          }
        }
        const int64_t R = (int64_t)p2col.size() - c0;
        if (R <= 0) continue;
        if (R > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
        it.uoff = c0;
        it.R = (int32_t)R;
      }
      it.yoff = e.bstart[lo + k];
      it.level = j;
      for (int i0 = 0; i0 < n; i0 += 128) {
        it.i0 = i0;
        it.nr = std::min(128, n - i0);
        p2.push_back(it);
      }
    }
  }
  std::vector<GTItem> all;
  for (int c = 0; c < 3; ++c) {
    g.n_p1[c] = (int64_t)items[c].size();
    all.insert(all.end(), items[c].begin(), items[c].end());
  }
  std::stable_sort(p2.begin(), p2.end(), [](const P2Item& a, const P2Item& b) {
    return (int64_t)a.R * a.nr > (int64_t)b.R * b.nr;
  });
  g.n_gitems = (int64_t)all.size();
  g.d_gitems = upload(all, e.stream);
  g.d_gcolv = upload(tab.colv, e.stream);
  g.d_gcolo = upload(tab.colo, e.stream);
  g.n_rcols = (int64_t)tab.cols.size();
  g.d_rcols = upload(tab.cols, e.stream);
  g.g_np = tab.np;
  g.n_gp2 = (int64_t)p2.size();
  g.d_gp2 = upload(p2, e.stream);
  if (dir == 1) {
    g.d_p2col = upload(p2col, e.stream);
    g.d_p2t = upload(p2t, e.stream);
  }
  H3D_HIP(hipStreamSynchronize(e.stream));
  g.built = true;
}

// Tables of class c of the mixed copy (forward): phase 1 over the class's V
// columns grouped by source box (V_s, one matrix per source in its V chunk),
// writing t in the class's target order; phase 2 over each target box's
// class U.  The columns' factors c_k are applied in phase 2 (`scale`).
void build_mixed_group_tables(Engine& e, int c) {
  GroupTables& g = e.grpm[c];
  g.free_all();
  const FactorClass& fc = e.mix.cls[c];
  std::vector<int32_t> btoff(std::max<int64_t>(e.nblk_all, 1), 0);
  std::vector<BlockDev> blk(std::max<int64_t>(e.nblk_all, 1));
  if (e.nblk_all) {
    H3D_HIP(hipMemcpy(btoff.data(), fc.d_btoff, sizeof(int32_t) * e.nblk_all, hipMemcpyDeviceToHost));
    H3D_HIP(hipMemcpy(blk.data(), fc.d_blk, sizeof(BlockDev) * e.nblk_all, hipMemcpyDeviceToHost));
  }
  std::vector<GTItem> items[3];
  P1Table tab;
  tab.colv.reserve(fc.rank_sum + fc.rank_sum / 16);
  tab.colo.reserve(fc.rank_sum + fc.rank_sum / 16);
  const int gc = e.g_cols == 1024 ? 1024 : 2048;
  const int64_t pbase = (int64_t)fc.T + 1;
  std::vector<P2Item> p2;
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    const LevelFactors& lf = fc.lf[j];
    const int64_t lo = level_offset(j);
    std::vector<int64_t> first(lv.nboxes + 1, 0);
    for (int32_t k : lv.ilidx) ++first[k + 1];
    for (int64_t k = 0; k < lv.nboxes; ++k) first[k + 1] += first[k];
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      if (n == 0) continue;
      const int nsplit = (n + kGRows - 1) / kGRows;
      const int64_t c0 = (int64_t)tab.colv.size();
      const int64_t ldv = class_ldv(n, fc.elem);
      for (int64_t q = lv.bbase + first[k]; q < lv.bbase + first[k + 1]; ++q) {
        const int32_t b = e.by_src[q];
        const int r = lf.rank[b - lv.bbase];
        for (int cc = 0; cc < r; ++cc)
          tab.add(blk[b].chunk, blk[b].voff + cc * ldv, nsplit, pbase, cc, r, btoff[b] + cc);
        if (nsplit > 1 && r) tab.np += (int64_t)nsplit * r;
      }
      if (pbase + tab.np > INT32_MAX) throw Error(H3D_ERR_INVALID, "problem too large");
      if ((int64_t)tab.colv.size() == c0) continue;
      tab.pad(c0, fc.T);
      p1_items(items, c0, (int64_t)tab.colv.size(), n, e.bstart[lo + k], gc);
    }
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      const int R = lf.Rt[k];
      if (n <= 0 || R <= 0) continue;
      P2Item it{};
      it.uoff = lf.uoff[k];
      it.chunk = lv.chunk[k];
      it.ld = class_ldu(n, fc.elem);
      it.R = R;
      it.toff = lf.toff[k];
      it.yoff = e.bstart[lo + k];
      it.level = j;
      for (int i0 = 0; i0 < n; i0 += 128) {
        it.i0 = i0;
        it.nr = std::min(128, n - i0);
        p2.push_back(it);
      }
    }
  }
  std::vector<GTItem> all;
  for (int q = 0; q < 3; ++q) {
    g.n_p1[q] = (int64_t)items[q].size();
    all.insert(all.end(), items[q].begin(), items[q].end());
  }
  std::stable_sort(p2.begin(), p2.end(), [](const P2Item& a, const P2Item& b) {
    return (int64_t)a.R * a.nr > (int64_t)b.R * b.nr;
  });
  g.n_gitems = (int64_t)all.size();
  g.d_gitems = upload(all, e.stream);
  g.d_gcolv = upload(tab.colv, e.stream);
  g.d_gcolo = upload(tab.colo, e.stream);
  g.n_rcols = (int64_t)tab.cols.size();
  g.d_rcols = upload(tab.cols, e.stream);
  g.g_np = tab.np;
  g.n_gp2 = (int64_t)p2.size();
  g.d_gp2 = upload(p2, e.stream);
  H3D_HIP(hipStreamSynchronize(e.stream));
  g.built = true;
}

// x (fp64), per-level y and near field of W right-hand sides per point.
void ensure_w_bufs(Engine& e, int W) {
  if (e.w_cap >= W) return;
  dfree(e.d_xw);
  dfree(e.d_ylevw);
  dfree(e.d_ynw);
  e.w_cap = 0;
  e.d_xw = dalloc<double>(e.N * W);
  e.d_ylevw = dalloc<double>(e.N * e.nL * W);
  e.d_ynw = dalloc<double>(e.N * W);
  e.w_cap = W;
}

void ensure_block_scratch(Engine& e, int W, int dir) {
  if (!e.grp[dir].built) build_group_tables(e, dir);
  ensure_w_bufs(e, W);
  W = e.w_cap;
  // t slots 0..T (T: spare for padded columns), then the partial sums (the
  // larger count of the directions built)
  int64_t np = 1;
  for (const GroupTables& g : e.grp) np = std::max(np, g.g_np);
  const int64_t T = e.T + 1;
  if (e.tw_w >= W && e.w_t >= T && e.w_np >= np) return;
  dfree(e.d_tw);
  e.tw_w = 0;
  e.d_tw = dalloc<double>((T + np) * W);
  e.tw_w = W;
  e.w_t = T;
  e.w_np = np;
}

void ensure_mixed_scratch(Engine& e, int W) {
  for (int c = 0; c < kMixClasses; ++c)
    if (!e.grpm[c].built) build_mixed_group_tables(e, c);
  ensure_w_bufs(e, W);
  W = e.w_cap;
  if (e.xw32_w < W) {
    dfree(e.d_xw32);
    e.xw32_w = 0;
    e.d_xw32 = dalloc<float>(e.N * W);
    e.xw32_w = W;
  }
  for (int c = 0; c < kMixClasses; ++c) {
    const int64_t need = ((int64_t)e.mix.cls[c].T + 1 + std::max<int64_t>(e.grpm[c].g_np, 1)) * W *
                         (class_acc64(c) ? (int64_t)sizeof(double) : (int64_t)sizeof(float));
    if (e.mtw_cap[c] >= need) continue;
    dfree(e.d_mtw[c]);
    e.mtw_cap[c] = 0;
    dmalloc(&e.d_mtw[c], (size_t)need);
    e.mtw_cap[c] = need;
  }
}

template <class S, class A, int W, int NC>
void launch_p1(Engine& e, const GroupTables& g, void* const* base, int nchunk, int64_t first,
               int64_t count, const A* xr, A* tw, hipStream_t st) {
  if (count == 0) return;
  auto f = e.g_cols == 1024 ? p1r_kernel<S, A, W, NC, 1024 / NC> : p1r_kernel<S, A, W, NC, 2048 / NC>;
  const size_t lds = sizeof(A) * 64 * NC * W;  // x; the tables are static
  const size_t stat = (sizeof(int64_t) + sizeof(int32_t)) * (e.g_cols == 1024 ? 1024 : 2048) / NC +
                      sizeof(void*) * kLdsChunks;
  if (lds + stat > 65536)
    H3D_HIP(hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  f<<<(unsigned)count, 256, lds, st>>>((const GTItem*)g.d_gitems + first, (int)count, g.d_gcolv,
                                       g.d_gcolo, base, nchunk, xr, tw);
  H3D_LAUNCH_CHECK();
}

template <class S, class A, int W>
void run_group(Engine& e, int prec, bool acc64, bool transpose, const double* b, double* y, int cnt) {
  hipStream_t st = e.stream;
  const int64_t N = e.N;
  const GroupTables& g = e.grp[transpose ? 1 : 0];
  // forward: phase 1 reads V, phase 2 U; transposed the other way round
  void* const* p1base = transpose ? e.d_ubase[prec] : e.d_vbase[prec];
  void* const* p2base = transpose ? e.d_vbase[prec] : e.d_ubase[prec];
  A* xr = (A*)e.d_xw;
  A* tw = (A*)e.d_tw;
  A* partw = tw + (e.T + 1) * W;  // partial sums after the t slots
  gather_w_kernel<A, W><<<(unsigned)((N * W + 255) / 256), 256, 0, st>>>(b, e.d_perm, N, cnt, xr);
  H3D_LAUNCH_CHECK();
  {
    const int64_t n = N * e.nL * W;
    zero_kernel<<<(unsigned)std::min<int64_t>(8192, (n + 255) / 256), 256, 0, st>>>(e.d_ylevw, n);
    H3D_LAUNCH_CHECK();
  }
  // near field into its own buffer (W per point) on the handle's second
  // stream, overlapping phases 1 and 2 (in line while a graph is captured);
  // the grouped transposes take symmetric kernels only, whose near field is
  // its own transpose
  const int nn = (int)e.nitems.size();
  const bool overlap = nn && e.overlap_near && e.side;
  auto launch_near = [&](hipStream_t ns) {
    const int64_t lo = level_offset(e.nL);
    const KernelDev kd = e.kern;
    const bool invr = kd.q == 7 || kd.q == 16;
    const unsigned gn = wave_grid(nn);
    const int32_t* nbr = e.lev[e.nL].d_nbr;
    if constexpr (std::is_same<A, float>::value) {
      if (invr)
        near_wm_kernel<InvR32, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                      e.d_bcount + lo, e.d_lx, e.d_ly, e.d_lz,
                                                      e.d_leafc, kd, xr, e.d_ynw);
      else
        near_wm_kernel<AnyQ32, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                      e.d_bcount + lo, e.d_lx, e.d_ly, e.d_lz,
                                                      e.d_leafc, kd, xr, e.d_ynw);
    } else {
      if (kd.q == kQPlugin)
        near_w64_kernel<Plugin, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                       e.d_bcount + lo, e.d_px, e.d_py, e.d_pz,
                                                       kd, xr, e.d_ynw);
      else if (invr && prec != H3D_PREC_FP64)
        near_w64_kernel<InvRNewton, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                           e.d_bcount + lo, e.d_px, e.d_py,
                                                           e.d_pz, kd, xr, e.d_ynw);
      else if (invr)
        near_w64_kernel<InvR, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                     e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, kd,
                                                     xr, e.d_ynw);
      else
        near_w64_kernel<AnyQ, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
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
  const int nch = (int)e.chunks.size();
  launch_p1<S, A, W, 1>(e, g, p1base, nch, 0, g.n_p1[0], xr, tw, st);
  launch_p1<S, A, W, 2>(e, g, p1base, nch, g.n_p1[0], g.n_p1[1], xr, tw, st);
  launch_p1<S, A, W, 4>(e, g, p1base, nch, g.n_p1[0] + g.n_p1[1], g.n_p1[2], xr, tw, st);
  if (g.n_rcols) {
    reduce_w_kernel<A, W><<<(unsigned)((g.n_rcols * W + 255) / 256), 256, 0, st>>>(
        (const int4*)g.d_rcols, g.n_rcols, partw, tw);
    H3D_LAUNCH_CHECK();
  }
  if (g.n_gp2) {
    if (transpose)
      p2w_kernel<S, A, W, true><<<wave_grid(g.n_gp2), 256, 0, st>>>(
          (const P2Item*)g.d_gp2, (int)g.n_gp2, p2base, tw, e.d_ylevw, N, g.d_p2col, g.d_p2t,
          nullptr, 0);
    else
      p2w_kernel<S, A, W><<<wave_grid(g.n_gp2), 256, 0, st>>>(
          (const P2Item*)g.d_gp2, (int)g.n_gp2, p2base, tw, e.d_ylevw, N, nullptr, nullptr,
          nullptr, 0);
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
void run_group_w(Engine& e, int prec, bool acc64, bool tr, const double* b, double* y, int cnt) {
  if (prec == H3D_PREC_FP64) run_group<double, double, W>(e, prec, acc64, tr, b, y, cnt);
  else if (prec == H3D_PREC_FP32 && acc64) run_group<float, double, W>(e, prec, acc64, tr, b, y, cnt);
  else if (prec == H3D_PREC_FP32) run_group<float, float, W>(e, prec, acc64, tr, b, y, cnt);
  else if (acc64) run_group<uint16_t, double, W>(e, prec, acc64, tr, b, y, cnt);
  else run_group<uint16_t, float, W>(e, prec, acc64, tr, b, y, cnt);
}

// One group of W right-hand sides on the mixed copy: classes 0 and 3 (fp32
// and fp24 factors, fp64 sums and x), classes 1 and 2 (fp16 and fp8 factors,
// fp32 sums and x) each through phase 1, the row-split reduce and phase 2
// (scaled t; every class after the first adds to the level sums); the near
// field as the single product forms it (fp64 sums), beside them on the second
// stream.
template <int W>
void run_group_mixed(Engine& e, const double* b, double* y, int cnt) {
  hipStream_t st = e.stream;
  const int64_t N = e.N;
  double* x64 = e.d_xw;
  float* x32 = e.d_xw32;
  gather_w_kernel<double, W><<<(unsigned)((N * W + 255) / 256), 256, 0, st>>>(b, e.d_perm, N, cnt, x64);
  H3D_LAUNCH_CHECK();
  gather_w_kernel<float, W><<<(unsigned)((N * W + 255) / 256), 256, 0, st>>>(b, e.d_perm, N, cnt, x32);
  H3D_LAUNCH_CHECK();
  {
    const int64_t n = N * e.nL * W;
    zero_kernel<<<(unsigned)std::min<int64_t>(8192, (n + 255) / 256), 256, 0, st>>>(e.d_ylevw, n);
    H3D_LAUNCH_CHECK();
  }
  const int nn = (int)e.nitems.size();
  const bool overlap = nn && e.overlap_near && e.side;
  auto launch_near = [&](hipStream_t ns) {
    const int64_t lo = level_offset(e.nL);
    const KernelDev kd = e.kern;
    const unsigned gn = wave_grid(nn);
    const int32_t* nbr = e.lev[e.nL].d_nbr;
    if (kd.q == kQPlugin)
      near_w64_kernel<Plugin, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                     e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, kd,
                                                     x64, e.d_ynw);
    else if (kd.q == 7 || kd.q == 16)
      near_w64_kernel<InvRNewton, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                         e.d_bcount + lo, e.d_px, e.d_py, e.d_pz,
                                                         kd, x64, e.d_ynw);
    else
      near_w64_kernel<AnyQ, W><<<gn, 256, 0, ns>>>(e.d_near, nn, nbr, e.d_bstart + lo,
                                                   e.d_bcount + lo, e.d_px, e.d_py, e.d_pz, kd,
                                                   x64, e.d_ynw);
    H3D_LAUNCH_CHECK();
  };
  if (overlap) {
    H3D_HIP(hipEventRecord(e.ev_fork, st));
    H3D_HIP(hipStreamWaitEvent(e.side, e.ev_fork, 0));
    launch_near(e.side);
    H3D_HIP(hipEventRecord(e.ev_join, e.side));
  }
  auto run_class = [&](auto SA, int c, const auto* x, auto* tw) {
    using S = typename decltype(SA)::first_type;
    using A = typename decltype(SA)::second_type;
    const FactorClass& fc = e.mix.cls[c];
    const GroupTables& g = e.grpm[c];
    const int nch = (int)fc.vdata.size();
    launch_p1<S, A, W, 1>(e, g, fc.d_vbase, nch, 0, g.n_p1[0], x, tw, st);
    launch_p1<S, A, W, 2>(e, g, fc.d_vbase, nch, g.n_p1[0], g.n_p1[1], x, tw, st);
    launch_p1<S, A, W, 4>(e, g, fc.d_vbase, nch, g.n_p1[0] + g.n_p1[1], g.n_p1[2], x, tw, st);
    if (g.n_rcols) {
      reduce_w_kernel<A, W><<<(unsigned)((g.n_rcols * W + 255) / 256), 256, 0, st>>>(
          (const int4*)g.d_rcols, g.n_rcols, tw + ((int64_t)fc.T + 1) * W, tw);
      H3D_LAUNCH_CHECK();
    }
    if (g.n_gp2) {
      p2w_kernel<S, A, W><<<wave_grid(g.n_gp2), 256, 0, st>>>(
          (const P2Item*)g.d_gp2, (int)g.n_gp2, fc.d_ubase, tw, e.d_ylevw, N, nullptr, nullptr,
          e.mix.d_scale + fc.tbase, c);
      H3D_LAUNCH_CHECK();
    }
  };
  run_class(std::pair<float, double>{}, 0, (const double*)x64, (double*)e.d_mtw[0]);
  run_class(std::pair<f24, double>{}, 3, (const double*)x64, (double*)e.d_mtw[3]);
  run_class(std::pair<_Float16, float>{}, 1, (const float*)x32, (float*)e.d_mtw[1]);
  if (e.mix.int8) run_class(std::pair<i8s, float>{}, 2, (const float*)x32, (float*)e.d_mtw[2]);
  else run_class(std::pair<f8e4m3, float>{}, 2, (const float*)x32, (float*)e.d_mtw[2]);
  if (nn && !overlap) launch_near(st);
  if (overlap) H3D_HIP(hipStreamWaitEvent(st, e.ev_join, 0));
  combine_w_kernel<W><<<(unsigned)((N + 63) / 64), 256, 0, st>>>(
      e.d_ylevw, e.nL, N, nn ? e.d_ynw : nullptr, e.d_iperm, b, cnt, e.alpha, e.kern.diag,
      e.d_diag, y);
  H3D_LAUNCH_CHECK();
}

}  // namespace

void GroupTables::free_all() {
  dfree(d_rcols);
  dfree(d_gitems);
  dfree(d_gp2);
  dfree(d_gcolv);
  dfree(d_gcolo);
  dfree(d_p2col);
  dfree(d_p2t);
  n_rcols = n_gitems = n_gp2 = g_np = 0;
  n_p1[0] = n_p1[1] = n_p1[2] = 0;
  built = false;
}

void release_group_tables(Engine& e) {
  for (GroupTables& g : e.grp) g.free_all();
  for (GroupTables& g : e.grpm) g.free_all();
}

void release_mixed_groups(Engine& e) {
  for (GroupTables& g : e.grpm) g.free_all();
  for (int c = 0; c < kMixClasses; ++c) {
    dfree(e.d_mtw[c]);
    e.mtw_cap[c] = 0;
  }
  dfree(e.d_xw32);
  e.xw32_w = 0;
}

int block_width(const Engine& e, int nrhs, int prec, bool transpose, bool acc64) {
  if (e.rhs_block <= 0 || nrhs < 2) return 1;
  if (prec == kPrecMixed && transpose) return 1;  // the mixed copy's transposes: one at a time
  if ((int64_t)e.chunks.size() > kLdsChunks) return 1;  // phase 1 keeps the chunk bases in LDS
  // a plugin kernel: the grouped fp32-sum near field works on fp32 offsets of
  // radial kernels, and a plugin's transpose needs entry(x_j, x_i) (the mixed
  // copy's grouped near field sums in fp64 with fp64 coordinates)
  if (e.kern.q == kQPlugin && prec != kPrecMixed &&
      (transpose || (prec != H3D_PREC_FP64 && !acc64)))
    return 1;
  return nrhs <= 16 || e.rhs_block == 16 ? 16 : 32;
}

void matvec_block(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool acc64,
                  bool transpose) {
  const int Wmax = e.rhs_block == 16 ? 16 : 32;
  if (prec == kPrecMixed) {
    ensure_mixed_scratch(e, nrhs <= 16 ? 16 : Wmax);
    for (int r0 = 0; r0 < nrhs; r0 += Wmax) {
      const int cnt = std::min(Wmax, nrhs - r0);
      const double* b = d_b + (int64_t)r0 * e.N;
      double* y = d_y + (int64_t)r0 * e.N;
      if (cnt <= 16) run_group_mixed<16>(e, b, y, cnt);
      else run_group_mixed<32>(e, b, y, cnt);
    }
    return;
  }
  ensure_block_scratch(e, nrhs <= 16 ? 16 : Wmax, transpose ? 1 : 0);
  const int64_t N = e.N;
  for (int r0 = 0; r0 < nrhs; r0 += Wmax) {
    const int cnt = std::min(Wmax, nrhs - r0);
    const double* b = d_b + (int64_t)r0 * N;
    double* y = d_y + (int64_t)r0 * N;
    if (cnt <= 16) run_group_w<16>(e, prec, acc64, transpose, b, y, cnt);
    else run_group_w<32>(e, prec, acc64, transpose, b, y, cnt);
  }
}

void autotune_rhs(Engine& e, int mode, int nrhs, int reps, double* ms, int* best) {
  const int prec = mode & 0xff;
  if (prec < 0 || prec > 2 || (mode & ~(0xff | H3D_ACC_FP64)))
    throw Error(H3D_ERR_INVALID, "grouped products take fp64, fp32 or bf16 factor copies");
  if (nrhs < 2) throw Error(H3D_ERR_INVALID, "nrhs must be >= 2");
  if (reps < 1) throw Error(H3D_ERR_INVALID, "reps must be >= 1");
  if ((int64_t)e.chunks.size() > kLdsChunks)
    throw Error(H3D_ERR_STATE, "more than 1024 factor chunks: products run one column at a time");
  const bool acc64 = prec == H3D_PREC_FP64 || (mode & H3D_ACC_FP64);
  prepare_mode(e, mode);
  const int64_t N = e.N;
  double* B = dalloc<double>(N * nrhs);
  double* Y = dalloc<double>(N * nrhs);
  hipEvent_t ev[2] = {nullptr, nullptr};
  const int saved_w = e.rhs_block, saved_c = e.g_cols;
  const int widths[2] = {16, 32}, cols[2] = {1024, 2048};
  int bw = saved_w > 0 ? saved_w : 32, bc = saved_c;
  try {
    H3D_HIP(hipMemsetAsync(B, 0, sizeof(double) * N * nrhs, e.stream));
    for (auto& x : ev) H3D_HIP(hipEventCreate(&x));
    double bt = 1e300;
    for (int wi = 0; wi < 2; ++wi)
      for (int ci = 0; ci < 2; ++ci) {
        if (ms) ms[wi * 2 + ci] = -1.0;
        if (wi == 1 && nrhs <= 16) continue;  // groups of <= 16 run at width 16 anyway
        e.rhs_block = widths[wi];
        if (e.g_cols != cols[ci]) {
          H3D_HIP(hipStreamSynchronize(e.stream));
          release_group_tables(e);
          e.g_cols = cols[ci];
        }
        matvec_block(e, B, Y, nrhs, prec, acc64, false);  // warm-up (builds the tables)
        H3D_HIP(hipEventRecord(ev[0], e.stream));
        for (int q = 0; q < reps; ++q) matvec_block(e, B, Y, nrhs, prec, acc64, false);
        H3D_HIP(hipEventRecord(ev[1], e.stream));
        H3D_HIP(hipEventSynchronize(ev[1]));
        float t = 0;
        H3D_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
        const double avg = t / reps;
        if (ms) ms[wi * 2 + ci] = avg;
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
