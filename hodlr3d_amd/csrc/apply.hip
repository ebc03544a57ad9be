// Fast matvec: batched low-rank applies streamed from HBM, the on-the-fly
// near field, and their transposes.
//
// y = A b is computed as (src/HODLR3DTree.cpp:1647-1805 restated for stored
// factors U = K(:,J) R^-1, V^T = L^-1 K(I,:), which equal the ACA vectors):
//   gather    x[s] = b[perm[s]]                               (assignCharges)
//   phase 1   t = V^T x: one wave per 64 consecutive columns of a target box;
//             each lane owns one column (of whichever block it falls in) and
//             walks that column's rows, so no cross-lane reduction is needed
//   phase 2   y_j = U t: one wave per <= 64 rows of a target box; lanes own
//             rows and walk the box's columns (contiguous column-major reads)
//   near      y_L += sum over self + face/edge leaves of K(i,j) x_j
//   combine   y[i] = sum_j y_j[iperm[i]]                (collect + reorder)
// Every factor element is read exactly once per matvec.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "engine.hpp"
#include "kernels.hpp"
#include "stream.hpp"

namespace h3d {

namespace {

constexpr int kP1Rows = 512;          // rows of a phase-1 item (64 lanes x 8 rows)
constexpr int64_t kP1Elems = 8192;    // factor entries per multi-block phase-1 item
constexpr int kP2Rows = 128;          // rows of one phase-2 item
constexpr int64_t kTItemBytes = 131072;  // factor bytes of a transposed phase-1 item

template <class A>
__device__ inline A wave_sum_xor(A v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}


// One DPP move of a 32- or 64-bit value (CTRL: the instruction's dpp_ctrl).
template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over aligned groups of g lanes (g a power of two <= 64, wave-uniform);
// every lane of a group gets the group's sum.  Within a row of 16 lanes the
// partners come by DPP (quad permutes for 1 and 2 lanes apart, then the
// half-row and row mirrors, whose partners lie in the other half), so that
// the sum takes no LDS round trip; across rows by swizzles.
template <class A>
__device__ inline A group_sum(A a, int g) {
  if (g >= 2) a += dpp_move<0xB1>(a);   // quad_perm [1,0,3,2]
  if (g >= 4) a += dpp_move<0x4E>(a);   // quad_perm [2,3,0,1]
  if (g >= 8) a += dpp_move<0x141>(a);  // row_half_mirror
  if (g >= 16) a += dpp_move<0x140>(a); // row_mirror
  if (g >= 32) a += __shfl_xor(a, 16);
  if (g >= 64) a += __shfl_xor(a, 32);
  return a;
}

__global__ void gather_kernel(const double* __restrict__ b, const int32_t* __restrict__ perm,
                              int64_t N, double* __restrict__ xp, float* __restrict__ xf) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  double v = b[perm[s]];
  xp[s] = v;
  if (xf) xf[s] = (float)v;
}

// Vector loads of 2..16 bytes (4-byte aligned from 4 bytes up), converted to
// the accumulation type.  Factor streams carry the non-temporal hint.

// Element `half` (0: low, 1: high 16 bits) of a 32-bit word of stored S values:
// uint16_t is bf16, _Float16 is IEEE half.
template <class S, class A>
__device__ inline A cvt_elem(uint32_t bits, int half) {
  if constexpr (std::is_same<S, _Float16>::value) {
    const uint16_t h = half ? (uint16_t)(bits >> 16) : (uint16_t)(bits & 0xffffu);
    return (A)(float)__builtin_bit_cast(_Float16, h);
  } else if constexpr (sizeof(S) == 2) {
    return (A)__uint_as_float(half ? (bits & 0xffff0000u) : (bits << 16));
  } else {
    return (A)__uint_as_float(bits);
  }
}

// NV consecutive stored elements of S (NV * sizeof(S) in {2,4,8,16} bytes) as
// raw words.  Kernels that keep their next loads in flight hold raw vectors
// and convert them where they are multiplied, so that the wait for a load
// falls at its use and not right after it.
template <int B>
struct RawOf;
template <>
struct RawOf<16> {
  using T = u32x4a;
};
template <>
struct RawOf<8> {
  using T = u32x2a;
};
template <>
struct RawOf<4> {
  using T = uint32_t;
};
template <>
struct RawOf<2> {
  using T = uint16_t;
};
typedef double f64x2a __attribute__((ext_vector_type(2), aligned(8)));
template <class S, int NV>
struct RawT {
  using T = typename RawOf<NV * (int)sizeof(S)>::T;
};
// packed fp24 entries: four (12 bytes, an 8-byte and a 4-byte load) or
// eight (24 bytes: 16 + 8)
struct Raw12 {
  u32x2a a;
  uint32_t b;
};
struct Raw24 {
  u32x4a a;
  u32x2a b;
};
template <>
struct RawT<f24, 4> {
  using T = Raw12;
};
template <>
struct RawT<f24, 8> {
  using T = Raw24;
};
template <>
struct RawT<double, 2> {
  using T = f64x2a;
};
template <>
struct RawT<double, 1> {
  using T = double;
};
template <class S, int NV>
using Raw = typename RawT<S, NV>::T;

template <class S, int NV>
__device__ inline Raw<S, NV> raw_load(const S* p) {
  if constexpr (sizeof(S) == 3) {
    static_assert(NV == 4 || NV == 8, "fp24 entries are loaded four or eight at a time");
    Raw<S, NV> r;
    if constexpr (NV == 4) {
      r.a = ntload<u32x2a>(p);
      r.b = ntload<uint32_t>((const char*)p + 8);
    } else {
      r.a = ntload<u32x4a>(p);
      r.b = ntload<u32x2a>((const char*)p + 16);
    }
    return r;
  } else {
    return ntload<Raw<S, NV>>(p);
  }
}

template <class S, class A, int NV>
__device__ inline void raw_cvt(const Raw<S, NV>& v, A (&o)[NV]) {
  constexpr int B = NV * (int)sizeof(S);
  if constexpr (sizeof(S) == 8) {
    if constexpr (NV == 2) {
      o[0] = (A)v.x, o[1] = (A)v.y;
    } else {
      o[0] = (A)v;
    }
  } else if constexpr (sizeof(S) == 3) {  // fp24: packed values, four per 12 bytes
    uint32_t u[8];
    if constexpr (NV == 4) {
      f24x4_bits(v.a.x, v.a.y, v.b, *(uint32_t(*)[4])u);
    } else {
      f24x4_bits(v.a.x, v.a.y, v.a.z, *(uint32_t(*)[4])u);
      f24x4_bits(v.a.w, v.b.x, v.b.y, *(uint32_t(*)[4])(u + 4));
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) o[q] = (A)__uint_as_float(u[q]);
  } else if constexpr (sizeof(S) == 1) {  // fp8: two values per conversion; int8: one
    uint32_t w[B >= 4 ? B / 4 : 1];
    if constexpr (B == 16) {
      w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
    } else if constexpr (B == 8) {
      w[0] = v.x, w[1] = v.y;
    } else {
      w[0] = (uint32_t)v;
    }
    if constexpr (std::is_same<S, i8s>::value) {
#pragma unroll
      for (int q = 0; q < NV; ++q)
        o[q] = (A)i8_to_f(w[q / 4], q % 4);
      return;
    }
#pragma unroll
    for (int q = 0; q < NV; q += 2) {
      float a, b;
      if (q & 2) f8x2_to_f<true>(w[q / 4], a, b);
      else f8x2_to_f<false>(w[q / 4], a, b);
      o[q] = (A)a, o[q + 1] = (A)b;
    }
  } else {
    uint32_t w[B >= 4 ? B / 4 : 1];
    if constexpr (B == 16) {
      w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
    } else if constexpr (B == 8) {
      w[0] = v.x, w[1] = v.y;
    } else {
      w[0] = (uint32_t)v;
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) o[q] = cvt_elem<S, A>(w[q * sizeof(S) / 4], (q * sizeof(S) / 2) & 1);
  }
}

template <class S, class A, int NV>
__device__ inline void load_factor(const S* p, A (&o)[NV]) {
  raw_cvt<S, A, NV>(raw_load<S, NV>(p), o);
}

// ---------------------------------------------------------------- phase 1
// One wave walks the blocks [b0, b1) of a target box (consecutive in memory).
// For each block V_b (column-major, ldv = m + (m & 1)) lpc lanes share a
// column: each lane holds up to 8 rows of x in registers (loaded once per
// block) and reads VEC consecutive rows of the column per load; the lanes'
// partial sums are combined inside the group (group_sum).  Rows past
// the item's range see x = 0, so loads may run into padding or the next column.
// With `scale` (indexed like t) the finished t_k is multiplied by scale_k.
template <class S, class A, int VEC>
__device__ __forceinline__ void phase1_item(const P1Item& it, void* const* __restrict__ vbase,
                                            const BlockDev* __restrict__ blk,
                                            const A* __restrict__ x, A* __restrict__ t,
                                            A* __restrict__ part,
                                            const double* __restrict__ scale) {
  constexpr int KMAX = 8 / VEC > 0 ? 8 / VEC : 1;  // rows per lane <= 8 (items have <= 512 rows)
  const int lane = threadIdx.x & 63;
  const S* V = (const S*)vbase[it.chunk];
  for (int b = it.b0; b < it.b1; ++b) {
    const BlockDev bd = blk[b];
    const int s1 = min(it.s1, bd.m);
    const int rows = s1 - it.s0;
    if (bd.r == 0 || rows <= 0) continue;
    const int ldv = bd.m + (bd.m & 1);
    const int need = (rows + VEC - 1) / VEC;
    int lpc = 1;
    while (lpc < need && lpc < 64) lpc <<= 1;
    const int K = (need + lpc - 1) / lpc;
    const int cpp = 64 / lpc, slot = lane / lpc, sub = lane % lpc;
    // Branch-free loads: rows past the block or past K read row 0 (and meet
    // x = 0), columns past r - 1 read column r - 1 (and are not stored); x is
    // loaded at clamped indices and masked by a 0/1 factor afterwards (a
    // select of a loaded value compiles into a branch and a wait per load).
    const A* xs = x + bd.src + it.s0;
    const S* base = V + bd.voff + it.s0;
    auto finish = [&](int l, A acc) {
      acc = group_sum(acc, lpc);
      if (l < bd.r && sub == 0) {
        if (it.pslot < 0)
          t[it.toff + bd.col + l] = scale ? acc * (A)scale[it.toff + bd.col + l] : acc;
        else
          part[(int64_t)it.pslot + l] = acc;
      }
    };
    if (K == 1) {
      // one load vector per lane and column: x and four passes' column loads
      // in flight together
      constexpr int UP = 4;
      const int r0 = sub * VEC;
      const int ro = r0 < rows ? r0 : 0;
      A xv[VEC];
#pragma unroll
      for (int q = 0; q < VEC; ++q) xv[q] = xs[min(r0 + q, rows - 1)];
      for (int c0 = 0; c0 < bd.r; c0 += UP * cpp) {
        Raw<S, VEC> v[UP];
#pragma unroll
        for (int u = 0; u < UP; ++u) {
          const int l = min(c0 + u * cpp + slot, bd.r - 1);
          v[u] = raw_load<S, VEC>(base + (int64_t)l * ldv + ro);
        }
        if (c0 == 0) {
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int q = 0; q < VEC; ++q) xv[q] *= r0 + q < rows ? (A)1 : (A)0;
        }
#pragma unroll
        for (int u = 0; u < UP; ++u) {
          A f[VEC];
          raw_cvt<S, A, VEC>(v[u], f);
          A acc = 0;
#pragma unroll
          for (int q = 0; q < VEC; ++q) acc += f[q] * xv[q];
          finish(c0 + u * cpp + slot, acc);
        }
      }
    } else {
      A xv[KMAX][VEC];
      int roff[KMAX];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        const int r0 = (sub + k * lpc) * VEC;
        roff[k] = (k < K && r0 < rows) ? r0 : 0;
#pragma unroll
        for (int q = 0; q < VEC; ++q) xv[k][q] = xs[min(r0 + q, rows - 1)];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        const int r0 = (sub + k * lpc) * VEC;
#pragma unroll
        for (int q = 0; q < VEC; ++q) xv[k][q] *= (k < K && r0 + q < rows) ? (A)1 : (A)0;
      }
      for (int c0 = 0; c0 < bd.r; c0 += cpp) {
        const int l = min(c0 + slot, bd.r - 1);
        const S* col = base + (int64_t)l * ldv;
        Raw<S, VEC> v[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) v[k] = raw_load<S, VEC>(col + roff[k]);
        A acc = 0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
          A f[VEC];
          raw_cvt<S, A, VEC>(v[k], f);
#pragma unroll
          for (int q = 0; q < VEC; ++q) acc += f[q] * xv[k][q];
        }
        finish(c0 + slot, acc);
      }
    }
  }
}

template <class S, class A, int VEC>
__global__ __launch_bounds__(256) void phase1_kernel(const P1Item* __restrict__ items, int nitems,
                                                     void* const* __restrict__ vbase,
                                                     const BlockDev* __restrict__ blk,
                                                     const A* __restrict__ x, A* __restrict__ t,
                                                     A* __restrict__ part,
                                                     const double* __restrict__ scale) {
  const int w = wave_id();
  if (w >= nitems) return;
  phase1_item<S, A, VEC>(items[w], vbase, blk, x, t, part, scale);
}

// The persistent form: a grid of as many waves as stay resident, each taking
// groups of items from the queue (an atomic counter cleared before the
// launch) until none is left; every wave exits once the counter has passed
// nitems.
template <class S, class A, int VEC>
__global__ __launch_bounds__(256) void phase1_queue_kernel(const P1Item* __restrict__ items,
                                                           int nitems,
                                                           void* const* __restrict__ vbase,
                                                           const BlockDev* __restrict__ blk,
                                                           const A* __restrict__ x,
                                                           A* __restrict__ t, A* __restrict__ part,
                                                           const double* __restrict__ scale,
                                                           unsigned* __restrict__ queue) {
  // four items per take: phase 1 has many small items (runs of blocks)
  queue_loop<4>(queue, nitems,
                [&](int w) { phase1_item<S, A, VEC>(items[w], vbase, blk, x, t, part, scale); });
}

// Sum of the nsplit partials part[base + q * nc] (q = 0, 1, ...) in q order,
// eight loads issued ahead of their adds (a plain loop waits for each load).
template <class A>
__device__ inline A split_sum(const A* __restrict__ part, int64_t base, int nsplit, int nc) {
  A acc = 0;
  int q = 0;
  for (; q + 8 <= nsplit; q += 8) {
    A v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[base + (int64_t)(q + u) * nc];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; q < nsplit; ++q) acc += part[base + (int64_t)q * nc];
  return acc;
}

// Sums the row-split partials of one block (nsplit x nc values, row ranges in
// order) into t[toff + l] (times scale[toff + l] when scaled).
template <class A>
__global__ void reduce_kernel(const P1Reduce* __restrict__ rs, int n, const A* __restrict__ part,
                              A* __restrict__ t, const double* __restrict__ scale) {
  const int w = wave_id();
  if (w >= n) return;
  const int lane = threadIdx.x & 63;
  const P1Reduce r = rs[w];
  for (int l = lane; l < r.nc; l += 64) {
    const A acc = split_sum(part, (int64_t)r.pbase + l, r.nsplit, r.nc);
    t[r.toff + l] = scale ? acc * (A)scale[r.toff + l] : acc;
  }
}

// VEC consecutive x values (fp32 or fp64; VEC * sizeof(A) a multiple of 16
// bytes) by 16-byte loads from a 4-byte aligned address, kept in the cache
// (x is read again by every item of the box).
template <class A, int VEC>
__device__ inline void load_xvec(const A* p, A (&o)[VEC]) {
  constexpr int NW = VEC * (int)sizeof(A) / 16;
  static_assert(NW * 16 == VEC * (int)sizeof(A), "whole 16-byte vectors");
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const u32x4a v = *(const __attribute__((address_space(1))) u32x4a*)((const char*)p + 16 * w);
    if constexpr (sizeof(A) == 4) {
      o[4 * w] = __uint_as_float(v.x);
      o[4 * w + 1] = __uint_as_float(v.y);
      o[4 * w + 2] = __uint_as_float(v.z);
      o[4 * w + 3] = __uint_as_float(v.w);
    } else {
      o[2 * w] = __hiloint2double((int)v.y, (int)v.x);
      o[2 * w + 1] = __hiloint2double((int)v.w, (int)v.z);
    }
  }
}

// ---------------------------------------------------------- mixed phase 1
// t = V^T x for one class of the recompressed copy, grouped by source box: an
// item is a run of columns of V_s (the columns of every block with source s,
// concatenated) over at most 64 x KMAX load vectors of rows (1024 fp32, 2048
// fp16 or 2048 fp8 rows with 16-byte loads), so x is loaded once per item.  lpc lanes
// share a column, each holding K <= kmax load vectors of VEC rows (x in
// registers), and reduce with group_sum.  t is written in source order
// (t_src[tsb + c], contiguous stores); a gather puts it in target order.  A
// load starts inside the rows, so it ends less than VEC elements past them
// (next column, padding or the chunk's zero tail).
template <class S, class A, int VEC>
__global__ __launch_bounds__(256) void mphase1_kernel(const MixP1Item* __restrict__ items,
                                                      int nitems, void* const* __restrict__ vbase,
                                                      const A* __restrict__ x,
                                                      A* __restrict__ tsrc, A* __restrict__ part,
                                                      int kmax, const double* __restrict__ scale) {
  // 64 lanes x KMAX x VEC >= the item's rows (mix_item_rows); two vectors for
  // 16 fp8 entries per lane (x of their rows stays in registers)
  constexpr int KMAX = VEC >= 16 ? 2 : (VEC >= 2 ? 4 : 8);
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const MixP1Item it = items[w];
  const int rows = it.s1 - it.s0;
  const int need = (rows + VEC - 1) / VEC;
  // kmax 0: by item, lanes across the rows (one vector each) for items of
  // at most 32 vectors, so that a wave instruction reads whole columns
  const int kk = kmax > 0 ? min(kmax, KMAX) : (need <= 32 ? 1 : KMAX);
  int lpc = 1;
  while (lpc * kk < need && lpc < 64) lpc <<= 1;
  const int K = (need + lpc - 1) / lpc;
  const int cpp = 64 / lpc, slot = lane / lpc, sub = lane % lpc;
  A xv[KMAX][VEC];
  const A* xs = x + it.xoff + it.s0;
  // x of the lane's rows, NX vectors (one for K == 1): all loads first, as
  // 16-byte vectors; a vector starts inside the item's rows (lanes past them
  // take the last one), so it ends less than VEC entries past them, in the
  // next box's x or in the zero tail of at least 64 entries after x.  Then a
  // 0/1 factor for the entries outside (a select of the loaded value would be
  // compiled into a branch around each load, each followed by a wait for it).
  const int slast = (rows - 1) / VEC * VEC;
  auto load_x = [&](auto NXC) {
    constexpr int NX = decltype(NXC)::value;
#pragma unroll
    for (int k = 0; k < NX; ++k) load_xvec<A, VEC>(xs + min((sub + k * lpc) * VEC, slast), xv[k]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int s = (sub + k * lpc) * VEC;
#pragma unroll
      for (int q = 0; q < VEC; ++q) xv[k][q] *= (k < K && s + q < rows) ? (A)1 : (A)0;
    }
  };
  const S* base = (const S*)vbase[it.vchunk] + it.voff + it.s0;
  // The columns are taken in units (UP column passes with one load vector per
  // lane when K == 1, else one pass with KMAX vectors per lane), two units'
  // loads in flight: the next unit's loads are issued before the current
  // unit's multiplies, reduction and stores.  Loads are branch-free at
  // in-bounds addresses (rows past the item or past K read the column's first
  // rows and meet x = 0; columns past c1 - 1 read column c1 - 1 and are not
  // stored); the values are consumed only by the multiplies.  Each column's
  // sum keeps the order of one pass at a time.
  constexpr int UP = 4;
  const int cols = K == 1 ? UP * cpp : cpp;  // columns per unit
  const int nu = (it.c1 - it.c0 + cols - 1) / cols;
  if (nu <= 0) return;  // no columns (the first prefetch would read column c0 - 1)
  auto finish = [&](int c, A a) {
    a = group_sum(a, lpc);
    if (c < it.c1 && sub == 0) {
      if (it.pslot < 0) tsrc[it.tsb + c] = scale ? a * (A)scale[it.tsb + c] : a;
      else part[(int64_t)it.pslot + (c - it.c0)] = a;
    }
  };
  // one straight loop per shape (ONE: K == 1), so that the loop body holds
  // no branch between a unit's loads and their use
  auto run = [&](auto ONE) {
    constexpr bool one = decltype(ONE)::value;
    constexpr int NL = one ? UP : KMAX;  // load vectors per lane and unit
    load_x(std::integral_constant<int, one ? 1 : KMAX>{});
    int roff[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int r = (sub + k * lpc) * VEC;
      roff[k] = one ? (sub * VEC < rows ? sub * VEC : 0) : ((k < K && r < rows) ? r : 0);
    }
    auto load = [&](int u, Raw<S, VEC> (&v)[NL]) {
      const int cu = it.c0 + u * cols;
#pragma unroll
      for (int k = 0; k < NL; ++k) {
        const int c = min(cu + (one ? k * cpp : 0) + slot, it.c1 - 1);
        v[k] = raw_load<S, VEC>(base + (int64_t)c * it.ld + roff[k]);
      }
    };
    auto compute = [&](int u, const Raw<S, VEC> (&v)[NL]) {
      const int cu = it.c0 + u * cols;
      if constexpr (one) {
#pragma unroll
        for (int k = 0; k < UP; ++k) {
          A f[VEC];
          raw_cvt<S, A, VEC>(v[k], f);
          A a = 0;
#pragma unroll
          for (int q = 0; q < VEC; ++q) a += f[q] * xv[0][q];
          finish(cu + k * cpp + slot, a);
        }
      } else {
        A a = 0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
          A f[VEC];
          raw_cvt<S, A, VEC>(v[k], f);
#pragma unroll
          for (int q = 0; q < VEC; ++q) a += f[q] * xv[k][q];
        }
        finish(cu + slot, a);
      }
    };
    // two units per iteration, one exit at the bottom, barriers against the
    // scheduler sinking the prefetch (an odd count's last unit lies past c1)
    Raw<S, VEC> v0[NL], v1[NL];
    load(0, v0);
    for (int u = 0; u < nu; u += 2) {
      load(min(u + 1, nu - 1), v1);
      __builtin_amdgcn_sched_barrier(0);
      compute(u, v0);
      load(min(u + 2, nu - 1), v0);
      __builtin_amdgcn_sched_barrier(0);
      compute(u + 1, v1);
    }
  };
  if (K == 1)
    run(std::true_type{});
  else
    run(std::false_type{});
}

// Row-split partials of the mixed phase 1 summed in range order into t_src
// (with `scale`, each times its factor, as the phase-1 kernel stores).
template <class A>
__global__ void mreduce_kernel(const MixReduce* __restrict__ rs, int n, const A* __restrict__ part,
                               A* __restrict__ tsrc, const double* __restrict__ scale) {
  const int w = wave_id();
  if (w >= n) return;
  const int lane = threadIdx.x & 63;
  const MixReduce r = rs[w];
  for (int l = lane; l < r.nc; l += 64) {
    const A v = split_sum(part, (int64_t)r.pbase + l, r.nsplit, r.nc);
    tsrc[r.tsb + l] = scale ? v * (A)scale[r.tsb + l] : v;
  }
}

// t_src[map[i]] = t[i]: target-order t into source order (the transposed
// phase 2 of a recompressed class reads each source box's V_s against it).
template <class A>
__global__ void tscatter_kernel(const int32_t* __restrict__ map, int n, const A* __restrict__ t,
                                A* __restrict__ tsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tsrc[map[i]] = t[i];
}

// t[i] = scale[i] * t_src[map[i]]: source-order phase-1 results into target
// order (contiguous map, scale and t, gathered t_src reads), four consecutive
// entries per thread with 16-byte map / scale / t accesses (the arrays start
// 256-byte aligned), the last n % 4 entries one at a time.
template <class A>
__global__ void tgather_kernel(const int32_t* __restrict__ map, int n, const A* __restrict__ tsrc,
                               const A* __restrict__ scale, A* __restrict__ t) {
  const int i = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (i + 3 < n) {
    const int4 m = *(const int4*)(map + i);
    A sc[4], v[4];
    if constexpr (sizeof(A) == 4) {
      const float4 q = *(const float4*)(scale + i);
      sc[0] = q.x, sc[1] = q.y, sc[2] = q.z, sc[3] = q.w;
    } else {
      const double2 q0 = *(const double2*)(scale + i), q1 = *(const double2*)(scale + i + 2);
      sc[0] = q0.x, sc[1] = q0.y, sc[2] = q1.x, sc[3] = q1.y;
    }
    v[0] = tsrc[m.x] * sc[0];
    v[1] = tsrc[m.y] * sc[1];
    v[2] = tsrc[m.z] * sc[2];
    v[3] = tsrc[m.w] * sc[3];
    if constexpr (sizeof(A) == 4) {
      *(float4*)(t + i) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
      *(double2*)(t + i) = make_double2(v[0], v[1]);
      *(double2*)(t + i + 2) = make_double2(v[2], v[3]);
    }
  } else {
    for (int k = i; k < n; ++k) t[k] = tsrc[map[k]] * scale[k];
  }
}

// ---------------------------------------------------------------- phase 2
// U of a target box is column-major with the item's ld.  A lane owns VR
// consecutive rows (one vector load per column); the wave's lanes are split
// into G groups that walk different columns, so one wave instruction reads G
// consecutive columns of the box (contiguous when the item spans the box).
// accumulate != 0 adds to y_level (a second factor copy) instead of storing.
// QUEUE: the persistent form's item body (valid always; the wave's own LDS
// slice is ordered by wave barriers instead of a workgroup barrier).
template <class S, class A, int VR, bool QUEUE>
__device__ __forceinline__ void phase2_item(const P2Item& it, bool valid, A* __restrict__ red,
                                            void* const* __restrict__ ubase,
                                            const A* __restrict__ t, double* __restrict__ ylev,
                                            int64_t N, int accumulate) {
  const int lane = threadIdx.x & 63;
  int G = 1, rcn = 1;
  A acc[VR];
#pragma unroll
  for (int q = 0; q < VR; ++q) acc[q] = 0;
  if (valid) rcn = (it.nr + VR - 1) / VR;
  // more row chunks than lanes (narrow loads): lanes take 64-chunk passes, one column at a time
  const bool multi = valid && rcn > 64;
  if (multi) {
    const S* U0 = (const S*)ubase[it.chunk] + it.uoff + it.i0;
    const A* tt = t + it.toff;
    const int64_t ld = it.ld;
    for (int p0 = 0; p0 < rcn; p0 += 64) {
      const int rho = p0 + lane;
      if (rho >= rcn) break;
      const S* U = U0 + rho * VR;
      A a[VR];
#pragma unroll
      for (int q = 0; q < VR; ++q) a[q] = 0;
      int c = 0;
      for (; c + 3 < it.R; c += 4) {
        A v0[VR], v1[VR], v2[VR], v3[VR];
        load_factor<S, A, VR>(U + (int64_t)c * ld, v0);
        load_factor<S, A, VR>(U + (int64_t)(c + 1) * ld, v1);
        load_factor<S, A, VR>(U + (int64_t)(c + 2) * ld, v2);
        load_factor<S, A, VR>(U + (int64_t)(c + 3) * ld, v3);
#pragma unroll
        for (int q = 0; q < VR; ++q) {
          a[q] += v0[q] * tt[c];
          a[q] += v1[q] * tt[c + 1];
          a[q] += v2[q] * tt[c + 2];
          a[q] += v3[q] * tt[c + 3];
        }
      }
      for (; c < it.R; ++c) {
        A v0[VR];
        load_factor<S, A, VR>(U + (int64_t)c * ld, v0);
#pragma unroll
        for (int q = 0; q < VR; ++q) a[q] += v0[q] * tt[c];
      }
#pragma unroll
      for (int q = 0; q < VR; ++q) {
        const int row = rho * VR + q;
        if (row < it.nr) {
          double* yp = ylev + (int64_t)(it.level - 1) * N + it.yoff + it.i0 + row;
          *yp = accumulate ? *yp + (double)a[q] : (double)a[q];
        }
      }
    }
  } else if (valid) {
    // Units of 4G columns (lane group g takes columns u*4G + g + kG), two
    // units' loads in flight: the next unit's loads are issued before the
    // current unit's multiplies.  Loads are branch-free at in-bounds
    // addresses (column clamped to R - 1, which meets t = 0; a lane past the
    // last group repeats group G - 1, whose result is not read); the values
    // are consumed only by the multiplies.  Each row's sum runs over its
    // columns in increasing order, as one column at a time would.
    G = 64 / rcn;
    const int rho = lane % rcn, g = min(lane / rcn, G - 1);
    const S* U = (const S*)ubase[it.chunk] + it.uoff + it.i0 + rho * VR;
    const A* tt = t + it.toff;
    const int64_t ld = it.ld;
    const int R = it.R, G4 = 4 * G;
    const int nu = (R + G4 - 1) / G4;
    auto load = [&](int u, Raw<S, VR> (&v)[4], A (&tv)[4]) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = u * G4 + k * G + g;
        const int cl = c < R ? c : R - 1;
        v[k] = raw_load<S, VR>(U + (int64_t)cl * ld);
        tv[k] = tt[cl];
      }
    };
    auto compute = [&](int u, const Raw<S, VR> (&v)[4], const A (&tv)[4]) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const A tm = u * G4 + k * G + g < R ? tv[k] : (A)0;
        A f[VR];
        raw_cvt<S, A, VR>(v[k], f);
#pragma unroll
        for (int q = 0; q < VR; ++q) acc[q] += f[q] * tm;
      }
    };
    // the prefetch is unconditional (a unit past the last one repeats the
    // last one's addresses), so that every path holds the same loads in
    // flight and the compiler's waits stay partial
    // Two units per iteration and one exit at the bottom: a mid-loop exit
    // would let the compiler move the next unit's loads past it, a buffer
    // copy at the end would wait for them, and the scheduler would sink them
    // below the multiplies without the barriers.  With an odd count the last
    // iteration's second unit lies past column R - 1 and adds zeros.
    if (nu > 0) {
      Raw<S, VR> v0[4], v1[4];
      A t0[4], t1[4];
      load(0, v0, t0);
      for (int u = 0; u < nu; u += 2) {
        load(min(u + 1, nu - 1), v1, t1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the multiplies
        compute(u, v0, t0);
        load(min(u + 2, nu - 1), v0, t0);
        __builtin_amdgcn_sched_barrier(0);
        compute(u + 1, v1, t1);
      }
    }
  }
  if constexpr (QUEUE) __builtin_amdgcn_wave_barrier();  // after the last item's reads
  if (G > 1) {
#pragma unroll
    for (int q = 0; q < VR; ++q) red[lane * VR + q] = acc[q];
  }
  if constexpr (QUEUE) __builtin_amdgcn_wave_barrier();  // a wave's LDS operations stay in order
  else __syncthreads();
  if (!valid || multi) return;
  double* y0 = ylev + (int64_t)(it.level - 1) * N + it.yoff + it.i0;
  if (G > 1) {
    // row o (held by lane o / VR of every group, entry o % VR) summed over
    // the groups in group order by lane o % 64: every lane reads G
    // consecutive-address values and the stores are contiguous
    for (int o = lane; o < it.nr; o += 64) {
      A s = red[o];
      for (int g = 1; g < G; ++g) s += red[g * rcn * VR + o];
      y0[o] = accumulate ? y0[o] + (double)s : (double)s;
    }
    return;
  }
  if (lane >= rcn) return;
#pragma unroll
  for (int q = 0; q < VR; ++q) {
    const int row = lane * VR + q;
    if (row >= it.nr) break;
    y0[row] = accumulate ? y0[row] + (double)acc[q] : (double)acc[q];
  }
}

template <class S, class A, int VR>
__global__ __launch_bounds__(256) void phase2_kernel(const P2Item* __restrict__ items, int nitems,
                                                     void* const* __restrict__ ubase,
                                                     const A* __restrict__ t,
                                                     double* __restrict__ ylev, int64_t N,
                                                     int accumulate) {
  __shared__ A red[kWavesPerBlock][64 * VR];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave_id();
  const bool valid = w < nitems;
  P2Item it;
  if (valid) it = items[w];
  phase2_item<S, A, VR, false>(it, valid, red[wv], ubase, t, ylev, N, accumulate);
}

// The persistent form of phase 2 (as phase1_queue_kernel).
template <class S, class A, int VR>
__global__ __launch_bounds__(256) void phase2_queue_kernel(const P2Item* __restrict__ items,
                                                           int nitems,
                                                           void* const* __restrict__ ubase,
                                                           const A* __restrict__ t,
                                                           double* __restrict__ ylev, int64_t N,
                                                           int accumulate,
                                                           unsigned* __restrict__ queue) {
  __shared__ A red[kWavesPerBlock][64 * VR];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // one item per take: phase 2 has fewer, larger items (largest first)
  queue_loop<1>(queue, nitems, [&](int w) {
    phase2_item<S, A, VR, true>(items[w], true, red[wv], ubase, t, ylev, N, accumulate);
  });
}


// --------------------------------------------------------- transpose phases
// y[s] = sum over blocks b with this source box of sum_l V_b[s, l] t[btoff_b + l]:
// lanes over rows, every block's row s read as r contiguous values.
template <class S, class A>
__global__ __launch_bounds__(256) void tphase2_kernel(const TP2Item* __restrict__ items, int nitems,
                                                      void* const* __restrict__ vbase,
                                                      const BlockDev* __restrict__ blk,
                                                      const int32_t* __restrict__ by_src,
                                                      const int32_t* __restrict__ btoff,
                                                      const A* __restrict__ t,
                                                      double* __restrict__ ylev, int64_t N,
                                                      int accumulate) {
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const TP2Item it = items[w];
  const bool act = lane < it.nr;
  const int s = it.i0 + (act ? lane : 0);
  A acc = 0;
  for (int q = 0; q < it.scount; ++q) {
    const int b = by_src[it.sfirst + q];
    const BlockDev bd = blk[b];
    if (bd.r == 0) continue;
    const int64_t ldv = class_ldv(bd.m, (int)sizeof(S));
    const S* col = (const S*)vbase[bd.chunk] + bd.voff + s;
    const A* tt = t + btoff[b];
    for (int l = 0; l < bd.r; ++l) acc += (A)Load<S>::f(col + l * ldv) * tt[l];
  }
  if (!act) return;
  double* yp = ylev + (int64_t)(it.level - 1) * N + it.yoff + s;
  *yp = accumulate ? *yp + (double)acc : (double)acc;
}

}  // namespace

namespace {

// ---------------------------------------------------------------- near field
// One wave per <= 64 rows of a leaf; lanes own target rows and walk the source
// points of the leaf and its face/edge neighbours; the i == j term is removed
// by a select in the leaf's own box (the diagonal enters in the combine).
//
// fp64: exact entries (the reference's arithmetic, scale * K(r^2) * x_j per
// pair).  The sources come in chunks of <= 64 (one coalesced load per lane)
// staged in the wave's own LDS slice; an item of at most 32 rows splits each
// chunk between the half-waves (lanes 0-31 the even sources, lanes 32-63 the
// odd ones, for the same targets; the halves are added at the end), so that
// every lane works.  A wave's LDS operations complete in order and no other
// wave uses its slice, so the staging needs no barrier, only to stay in
// program order (wave barriers for the compiler).
template <class KF, bool SWAP = false>
__global__ __launch_bounds__(256) void near64_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const int32_t* __restrict__ perm, KernelDev kd, const double* __restrict__ x,
    double* __restrict__ yN) {
  __shared__ double2 lxy[kWavesPerBlock][64];
  __shared__ double2 lzq[kWavesPerBlock][64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const bool two = it.nr <= 32;
  const int tl = two ? (lane & 31) : lane;  // target of this lane
  const int h = two ? (lane >> 5) : 0;      // first source of this lane in a chunk
  const int step = two ? 2 : 1;
  const int rs = bstart[it.leaf];
  const bool act = tl < it.nr;
  const int ip = rs + it.i0 + (act ? tl : 0);
  const double xi = px[ip], yi = py[ip], zi = pz[ip];
  double acc = 0.0;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    // the self pair can only lie in the leaf's own box (SELF)
    auto sweep = [&](auto SELF) {
      constexpr bool self_box = decltype(SELF)::value;
      for (int s0 = 0; s0 < m; s0 += 64) {
        const int cnt = min(64, m - s0);
        const int jp = cs + s0 + (lane < cnt ? lane : 0);
        const double2 sxy = make_double2(px[jp], py[jp]), szq = make_double2(pz[jp], x[jp]);
        __builtin_amdgcn_wave_barrier();  // after the previous chunk's reads
        lxy[wv][lane] = sxy;
        lzq[wv][lane] = szq;
        __builtin_amdgcn_wave_barrier();  // before this chunk's reads
        const int self = ip - cs - s0;
        auto pair = [&](const double2& a, const double2& b, int u) {
          // entry (target, source); SWAP: (source, target), the transpose's
          const double k = SWAP ? kpair<KF>(kd.q, a.x, a.y, b.x, xi, yi, zi)
                                : kpair<KF>(kd.q, xi, yi, zi, a.x, a.y, b.x);
          const double v = kd.scale * k * b.y;
          if constexpr (self_box) acc += u == self ? 0.0 : v;
          else acc += v;
        };
        int u = h;
        for (; u + 3 * step < cnt; u += 4 * step) {
          double2 a[4], b[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) a[k] = lxy[wv][u + k * step], b[k] = lzq[wv][u + k * step];
#pragma unroll
          for (int k = 0; k < 4; ++k) pair(a[k], b[k], u + k * step);
        }
        for (; u < cnt; u += step) pair(lxy[wv][u], lzq[wv][u], u);
      }
    };
    if (nn == it.leaf)
      sweep(std::true_type{});
    else
      sweep(std::false_type{});
  }
  if (two) acc += __shfl_xor(acc, 32);  // the other half's sources
  if (!act || h != 0) return;
  yN[ip] = acc;
}

// Near field with fp32 entries: coordinates as fp32 offsets from the leaf
// centre plus the fp32 difference of the two leaf centres, so near pairs keep
// ~1e-7 relative accuracy in r.  A is the type of x and of the sums (float, or
// double for fp64 accumulation of the fp32 entries).  One wave
// per item (<= 64 target rows of a leaf).  The sources come in chunks of <= 64
// (one coalesced load per lane) staged in the wave's own LDS slice; an item of
// at most 32 rows splits each chunk between the half-waves (lanes 0-31 the
// even sources, lanes 32-63 the odd ones, for the same targets; the halves
// are added at the end), so that every lane works.  A wave's LDS operations
// complete in order and no other wave uses its slice, so the staging needs no
// barrier, only to stay in program order (wave barriers for the compiler).
template <class A, class KF>
__global__ __launch_bounds__(256) void near32_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const float* __restrict__ lx, const float* __restrict__ ly, const float* __restrict__ lz,
    const double* __restrict__ leafc, const int32_t* __restrict__ perm, KernelDev kd,
    const A* __restrict__ x, double* __restrict__ yN) {
  __shared__ float4 ls[kWavesPerBlock][64];
  __shared__ A lq[kWavesPerBlock][64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const bool two = it.nr <= 32;
  const int tl = two ? (lane & 31) : lane;  // target of this lane
  const int h = two ? (lane >> 5) : 0;      // first source of this lane in a chunk
  const int step = two ? 2 : 1;
  const int rs = bstart[it.leaf];
  const bool act = tl < it.nr;
  const int ip = rs + it.i0 + (act ? tl : 0);
  const float xi = lx[ip], yi = ly[ip], zi = lz[ip];
  const double cxk = leafc[3 * (int64_t)it.leaf], cyk = leafc[3 * (int64_t)it.leaf + 1],
               czk = leafc[3 * (int64_t)it.leaf + 2];
  A acc = 0;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    const float dcx = (float)(cxk - leafc[3 * (int64_t)nn]);
    const float dcy = (float)(cyk - leafc[3 * (int64_t)nn + 1]);
    const float dcz = (float)(czk - leafc[3 * (int64_t)nn + 2]);
    const float ax = xi + dcx, ay = yi + dcy, az = zi + dcz;
    A part = 0;
    // the self pair can only lie in the leaf's own box (SELF)
    auto sweep = [&](auto SELF) {
      constexpr bool self_box = decltype(SELF)::value;
      for (int s0 = 0; s0 < m; s0 += 64) {
        const int cnt = min(64, m - s0);
        const int jp = cs + s0 + (lane < cnt ? lane : 0);
        const float4 sp = make_float4(lx[jp], ly[jp], lz[jp], 0.0f);
        const A sq = x[jp];
        __builtin_amdgcn_wave_barrier();  // after the previous chunk's reads
        ls[wv][lane] = sp;
        lq[wv][lane] = sq;
        __builtin_amdgcn_wave_barrier();  // before this chunk's reads
        // four sources per step (their LDS reads issued together); the self
        // pair is selected away after its entry is formed (no branch)
        const int self = ip - cs - s0;
        auto pair = [&](const float4& sv, A q, int u) {
          const float dx = ax - sv.x, dy = ay - sv.y, dz = az - sv.z;
          const float R2 = dx * dx + dy * dy + dz * dz;
          float kv = KF::f(kd.q, R2);
          if constexpr (self_box) kv = u == self ? 0.0f : kv;
          part += (A)kv * q;
        };
        int u = h;
        for (; u + 3 * step < cnt; u += 4 * step) {
          float4 sv[4];
          A q[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) sv[k] = ls[wv][u + k * step], q[k] = lq[wv][u + k * step];
#pragma unroll
          for (int k = 0; k < 4; ++k) pair(sv[k], q[k], u + k * step);
        }
        for (; u < cnt; u += step) pair(ls[wv][u], lq[wv][u], u);
      }
    };
    if (nn == it.leaf)
      sweep(std::true_type{});
    else
      sweep(std::false_type{});
    acc += part;
  }
  if (two) acc += __shfl_xor(acc, 32);  // the other half's sources
  if (!act || h != 0) return;
  yN[ip] = kd.scale * (double)acc;
}

// 1/r near field with fp64 coordinates and sums: 1/r is the fp32 reciprocal
// square root of r^2 refined by one Newton step in fp64,
//   y1 = y0 (3 - r^2 y0^2) / 2,
// which squares its relative error (~1e-7 -> ~2e-14) at the cost of three
// fp64 operations instead of an fp64 square root and division.
__global__ __launch_bounds__(256) void near_nr_kernel(
    const NearItem* __restrict__ items, int nitems, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ bstart, const int32_t* __restrict__ bcount,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    KernelDev kd, const double* __restrict__ x, double* __restrict__ yN) {
  // sources staged per chunk in the wave's LDS slice and split between the
  // half-waves for items of <= 32 rows, as in near64_kernel
  __shared__ double2 lxy[kWavesPerBlock][64];
  __shared__ double2 lzq[kWavesPerBlock][64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave_id();
  if (w >= nitems) return;
  const int lane = threadIdx.x & 63;
  const NearItem it = items[w];
  const bool two = it.nr <= 32;
  const int tl = two ? (lane & 31) : lane;
  const int h = two ? (lane >> 5) : 0;
  const int step = two ? 2 : 1;
  const int rs = bstart[it.leaf];
  const bool act = tl < it.nr;
  const int ip = rs + it.i0 + (act ? tl : 0);
  const double xi = px[ip], yi = py[ip], zi = pz[ip];
  double acc = 0.0;
  for (int sl = 0; sl <= 26; ++sl) {
    const int nn = sl < 26 ? nbr[(int64_t)it.leaf * 26 + sl] : it.leaf;
    if (nn < 0) continue;
    const int cs = bstart[nn], m = bcount[nn];
    auto sweep = [&](auto SELF) {
      constexpr bool self_box = decltype(SELF)::value;
      for (int s0 = 0; s0 < m; s0 += 64) {
        const int cnt = min(64, m - s0);
        const int jp = cs + s0 + (lane < cnt ? lane : 0);
        const double2 sxy = make_double2(px[jp], py[jp]), szq = make_double2(pz[jp], x[jp]);
        __builtin_amdgcn_wave_barrier();  // after the previous chunk's reads
        lxy[wv][lane] = sxy;
        lzq[wv][lane] = szq;
        __builtin_amdgcn_wave_barrier();  // before this chunk's reads
        const int self = ip - cs - s0;
        auto pair = [&](const double2& a, const double2& b, int u) {
          const double dx = xi - a.x, dy = yi - a.y, dz = zi - b.x;
          const double r2 = dx * dx + dy * dy + dz * dz;
          // raw v_rsq_f32 on a normal fp32 argument (r^2 below FLT_MIN is clamped)
          const double y0 = (double)__builtin_amdgcn_rsqf(fmaxf((float)r2, 1.17549435e-38f));
          double y1 = y0 * (1.5 - 0.5 * r2 * y0 * y0);
          if constexpr (self_box) y1 = u == self ? 0.0 : y1;
          acc += y1 * b.y;
        };
        int u = h;
        for (; u + 3 * step < cnt; u += 4 * step) {
          double2 a[4], b[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) a[k] = lxy[wv][u + k * step], b[k] = lzq[wv][u + k * step];
#pragma unroll
          for (int k = 0; k < 4; ++k) pair(a[k], b[k], u + k * step);
        }
        for (; u < cnt; u += step) pair(lxy[wv][u], lzq[wv][u], u);
      }
    };
    if (nn == it.leaf)
      sweep(std::true_type{});
    else
      sweep(std::false_type{});
  }
  if (two) acc += __shfl_xor(acc, 32);  // the other half's sources
  if (!act || h != 0) return;
  yN[ip] = kd.scale * acc;
}

// y[i] = alpha * (sum over levels of y_j at i's position + the near field
// there) + d_i * b[i], where d_i is the diagonal entry K(i,i) (per point when
// diag_vec is set): the collect and reorder steps of
// src/HODLR3DTree.cpp:1781-1805 plus the i == j term of the self block.
__global__ void combine_kernel(const double* __restrict__ ylev, int nL, int64_t N,
                               const double* __restrict__ ynear,
                               const int32_t* __restrict__ iperm, const double* __restrict__ b,
                               double alpha, double diag, const double* __restrict__ diag_vec,
                               double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t s = iperm[i];
  double acc = 0.0;
  for (int j = 0; j < nL; ++j) acc = acc + ylev[(int64_t)j * N + s];
  if (ynear) acc = acc + ynear[s];
  const double d = diag_vec ? diag_vec[i] : diag;
  y[i] = alpha * acc + d * b[i];
}

template <class S>
__global__ void convert_kernel(const void* __restrict__ src, int sprec, int64_t n, S* __restrict__ dst);
template <>
__global__ void convert_kernel<float>(const void* __restrict__ src, int sprec, int64_t n,
                                      float* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = sprec == 0 ? (float)((const double*)src)[i] : ((const float*)src)[i];
}
template <>
__global__ void convert_kernel<uint16_t>(const void* __restrict__ src, int sprec, int64_t n,
                                         uint16_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = f_to_bf16(sprec == 0 ? (float)((const double*)src)[i] : ((const float*)src)[i]);
}
template <>
__global__ void convert_kernel<double>(const void* __restrict__ src, int sprec, int64_t n,
                                       double* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (double)((const float*)src)[i];
}


// Direct FP64 O(N) sum for one row per workgroup: y_i = alpha * sum_{j != i}
// K(i,j) b_j + d_i b_i with Neumaier-compensated partial sums per thread and a
// compensated tree over the workgroup.  The reference's drivers form exact
// products the same way, from kernel columns (examples/testHODLR3D.cpp:139-143).
template <class KF>
__global__ __launch_bounds__(256) void direct_kernel(
    const int64_t* __restrict__ rows, int64_t nrows, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ iperm, int64_t N, KernelDev kd, double alpha,
    const double* __restrict__ diag_vec, const double* __restrict__ b, double* __restrict__ out) {
  __shared__ double rs[256], rc[256];
  const int64_t row = blockIdx.x;
  if (row >= nrows) return;
  const int64_t i = rows[row];
  const int64_t ip = iperm[i];
  const double xi = px[ip], yi = py[ip], zi = pz[ip];
  double s = 0.0, c = 0.0;
  for (int64_t q = threadIdx.x; q < N; q += 256) {
    if (q == ip) continue;
    const double v = kd.scale * kpair<KF>(kd.q, xi, yi, zi, px[q], py[q], pz[q]) * b[perm[q]];
    const double t = s + v;
    c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  rs[threadIdx.x] = s;
  rc[threadIdx.x] = c;
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) {
      const double a = rs[threadIdx.x], bb = rs[threadIdx.x + d];
      const double t = a + bb;
      rc[threadIdx.x] += rc[threadIdx.x + d] + (fabs(a) >= fabs(bb) ? (a - t) + bb : (bb - t) + a);
      rs[threadIdx.x] = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double dg = diag_vec ? diag_vec[i] : kd.diag;
    out[row] = alpha * (rs[0] + rc[0]) + dg * b[i];
  }
}

}  // namespace

void direct_rows(Engine& e, const double* d_b, const int64_t* h_rows, int64_t nrows, double* d_out) {
  if (nrows <= 0) return;
  for (int64_t r = 0; r < nrows; ++r)
    if (h_rows[r] < 0 || h_rows[r] >= e.N) throw Error(H3D_ERR_INVALID, "row out of range");
  hipStream_t st = e.stream;
  int64_t* d_rows = dalloc<int64_t>(nrows);
  H3D_HIP(hipMemcpyAsync(d_rows, h_rows, sizeof(int64_t) * nrows, hipMemcpyHostToDevice, st));
  for (int64_t off = 0; off < nrows; off += 65535) {
    const unsigned g = (unsigned)std::min<int64_t>(65535, nrows - off);
    if (e.kern.q == 7 || e.kern.q == 16)
      direct_kernel<InvR><<<g, 256, 0, st>>>(d_rows + off, nrows - off, e.d_px, e.d_py, e.d_pz,
                                             e.d_perm, e.d_iperm, e.N, e.kern, e.alpha, e.d_diag,
                                             d_b, d_out + off);
    else if (e.kern.q == kQPlugin)
      direct_kernel<Plugin><<<g, 256, 0, st>>>(d_rows + off, nrows - off, e.d_px, e.d_py,
                                               e.d_pz, e.d_perm, e.d_iperm, e.N, e.kern, e.alpha,
                                               e.d_diag, d_b, d_out + off);
    else
      direct_kernel<AnyQ><<<g, 256, 0, st>>>(d_rows + off, nrows - off, e.d_px, e.d_py, e.d_pz,
                                             e.d_perm, e.d_iperm, e.N, e.kern, e.alpha, e.d_diag,
                                             d_b, d_out + off);
    H3D_LAUNCH_CHECK();
  }
  H3D_HIP(hipStreamSynchronize(st));
  dfree(d_rows);
}

// The per-precision tables of chunk base pointers, rebuilt as a whole: the new
// tables are made first and replace the old ones only once all exist, so a
// failed allocation leaves the previous tables in place.
void upload_chunk_table(Engine& e) {
  const int64_t nc = (int64_t)e.chunks.size();
  void** nu[3] = {nullptr, nullptr, nullptr};
  void** nv[3] = {nullptr, nullptr, nullptr};
  try {
    for (int p = 0; p < 3; ++p) {
      bool have = nc > 0;
      for (const Chunk& c : e.chunks) have = have && c.data[p] != nullptr;
      if (!have) continue;
      std::vector<void*> u(nc), v(nc);
      for (int64_t i = 0; i < nc; ++i) {
        u[i] = e.chunks[i].data[p];
        v[i] = (char*)e.chunks[i].data[p] + e.chunks[i].nU * elem_size(p);
      }
      nu[p] = upload(u, e.stream);
      nv[p] = upload(v, e.stream);
    }
    H3D_HIP(hipStreamSynchronize(e.stream));
  } catch (...) {
    for (int p = 0; p < 3; ++p) dfree(nu[p]), dfree(nv[p]);
    throw;
  }
  for (int p = 0; p < 3; ++p) {
    dfree(e.d_ubase[p]);
    dfree(e.d_vbase[p]);
    e.d_ubase[p] = nu[p];
    e.d_vbase[p] = nv[p];
  }
}

void Plan::free_device() {
  dfree(d_p1);
  dfree(d_p1r);
  dfree(d_p2);
  dfree(d_tp1);
  dfree(d_tp1r);
  dfree(d_tp2);
}

// Phase-2 items of one factor copy: each target box's rows in chunks of p2rows
// (64 lanes x one 16-byte vector: 128 fp64, 256 fp32 or 512 fp16/bf16 rows), so
// that a wave instruction reads 1 KB of one column when the item spans that
// many rows; larger items first (less tail at the end of a launch).
static void p2_items(const Engine& e, const std::vector<const LevelFactors*>& lfs, int p2rows,
                     int class_elem, std::vector<P2Item>& out) {
  out.clear();
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    const LevelFactors& lf = *lfs[j];
    const int64_t lo = level_offset(j);
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      const int R = lf.Rt[k];
      if (n <= 0 || R <= 0) continue;
      for (int i0 = 0; i0 < n; i0 += p2rows) {
        P2Item it{};
        it.uoff = lf.uoff[k];
        it.chunk = lv.chunk[k];
        it.ld = class_elem ? class_ldu(n, class_elem) : aca_ldu(n);
        it.i0 = i0;
        it.nr = std::min(p2rows, n - i0);
        it.R = R;
        it.toff = lf.toff[k];
        it.yoff = e.bstart[lo + k];
        it.level = j;
        out.push_back(it);
      }
    }
  }
  std::stable_sort(out.begin(), out.end(), [](const P2Item& a, const P2Item& b) {
    return (int64_t)a.R * a.nr > (int64_t)b.R * b.nr;
  });
}

void build_plan_host(const Engine& e, const std::vector<const LevelFactors*>& lfs, Plan& out,
                     int p2rows, int class_elem) {
  const int nL = e.nL;
  out.p1.clear();
  out.p1r.clear();
  out.p2.clear();
  out.tp1.clear();
  out.tp1r.clear();
  out.tp2.clear();
  int32_t pslot = 0, tslot = 0;
  int64_t src_cursor = 0;
  for (int j = 1; j <= nL; ++j) {
    const Level& lv = e.lev[j];
    const LevelFactors& lf = *lfs[j];
    const int64_t lo = level_offset(j);
    std::vector<int32_t> nsrc(lv.nboxes, 0);
    for (int32_t ki : lv.ilidx) nsrc[ki]++;
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int n = e.bcount[lo + k];
      const int R = lf.Rt[k];
      if (n > 0 && R > 0) {
        // phase 1: runs of consecutive blocks up to kP1Elems factor entries;
        // a block with more than kP1Rows rows gets one item per row range
        // (partial sums reduced afterwards in range order)
        {
          int32_t g0 = -1;
          int64_t gel = 0;
          auto flush = [&](int32_t end) {
            if (g0 >= 0) {
              P1Item it{};
              it.chunk = lv.chunk[k];
              it.b0 = (int32_t)lv.bbase + g0;
              it.b1 = (int32_t)lv.bbase + end;
              it.s0 = 0;
              it.s1 = kP1Rows;
              it.toff = lf.toff[k];
              it.pslot = -1;
              out.p1.push_back(it);
            }
            g0 = -1;
            gel = 0;
          };
          for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
            const int r = lf.rank[b];
            const int m = e.bcount[lo + lv.ilidx[b]];
            if (r == 0) continue;
            if (m > kP1Rows) {
              flush(b);
              const int nsplit = (m + kP1Rows - 1) / kP1Rows;
              out.p1r.push_back({pslot, nsplit, lf.toff[k] + lf.col[b], r});
              for (int q = 0; q < nsplit; ++q) {
                P1Item it{};
                it.chunk = lv.chunk[k];
                it.b0 = (int32_t)lv.bbase + b;
                it.b1 = it.b0 + 1;
                it.s0 = q * kP1Rows;
                it.s1 = std::min(m, (q + 1) * kP1Rows);
                it.toff = lf.toff[k];
                it.pslot = pslot;
                pslot += r;
                out.p1.push_back(it);
              }
              continue;
            }
            if (g0 >= 0 && gel + (int64_t)r * m > kP1Elems) flush(b);
            if (g0 < 0) g0 = b;
            gel += (int64_t)r * m;
          }
          flush(lv.iloff[k + 1]);
        }
        // transpose phase 1: the box's U columns in runs (about kTItemBytes
        // of factors), rows in ranges of the mixed phase-1 kernel's reach
        // (that of the fp64 copy for the ACA copies, whose plan serves every
        // precision)
        {
          const int es = class_elem ? class_elem : 8;
          const int trows = mix_item_rows(es);
          const int nsplit = (n + trows - 1) / trows;
          const int rows0 = std::min(n, trows);
          const int ncol = (int)std::max<int64_t>(
              16, std::min<int64_t>(4096, kTItemBytes / ((int64_t)rows0 * es)));
          const int ld = class_elem ? class_ldu(n, class_elem) : aca_ldu(n);
          for (int c0 = 0; c0 < R; c0 += ncol) {
            const int c1 = std::min(R, c0 + ncol);
            int32_t pbase = -1;
            if (nsplit > 1) {
              pbase = tslot;
              tslot += nsplit * (c1 - c0);
              out.tp1r.push_back({pbase, nsplit, c1 - c0, lf.toff[k] + c0});
            }
            for (int q = 0; q < nsplit; ++q) {
              MixP1Item it{};
              it.voff = lf.uoff[k];
              it.vchunk = lv.chunk[k];
              it.ld = ld;
              it.s0 = q * trows;
              it.s1 = std::min(n, (q + 1) * trows);
              it.c0 = c0;
              it.c1 = c1;
              it.tsb = lf.toff[k];
              it.xoff = e.bstart[lo + k];
              it.pslot = nsplit > 1 ? pbase + q * (c1 - c0) : -1;
              out.tp1.push_back(it);
            }
          }
        }
      }
    }
    // transpose phase 2 over source boxes (by_src is grouped by level, then box)
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int m = e.bcount[lo + k];
      if (m > 0 && nsrc[k] > 0)
        for (int i0 = 0; i0 < m; i0 += 64)
          out.tp2.push_back({(int32_t)src_cursor, nsrc[k], e.bstart[lo + k], i0, std::min(64, m - i0), j});
      src_cursor += nsrc[k];
    }
  }
  p2_items(e, lfs, p2rows, class_elem, out.p2);
  out.npartial = std::max(pslot, tslot);
}

void upload_plan(Engine& e, Plan& out) {
  hipStream_t st = e.stream;
  out.free_device();
  out.d_p1 = upload(out.p1, st);
  out.d_p1r = upload(out.p1r, st);
  out.d_p2 = upload(out.p2, st);
  out.d_tp1 = upload(out.tp1, st);
  out.d_tp1r = upload(out.tp1r, st);
  out.d_tp2 = upload(out.tp2, st);
  H3D_HIP(hipStreamSynchronize(st));
}

void build_plan_for(Engine& e, const std::vector<const LevelFactors*>& lfs, Plan& out,
                    int p2rows, int class_elem) {
  build_plan_host(e, lfs, out, p2rows, class_elem);
  upload_plan(e, out);
}

void build_plan(Engine& e) {
  std::vector<const LevelFactors*> lfs(e.nL + 1, nullptr);
  for (int j = 1; j <= e.nL; ++j) lfs[j] = &e.lev[j];
  build_plan_for(e, lfs, e.plan, kP2Rows);
  for (int q = 0; q < 2; ++q) {
    p2_items(e, lfs, q == 0 ? 2 * kP2Rows : 4 * kP2Rows, 0, e.p2_narrow[q]);
    dfree(e.d_p2_narrow[q]);
    e.d_p2_narrow[q] = upload(e.p2_narrow[q], e.stream);
  }
  H3D_HIP(hipStreamSynchronize(e.stream));
  e.nitems.clear();
  const int64_t lo = level_offset(e.nL);
  for (int64_t k = 0; k < e.lev[e.nL].nboxes; ++k) {
    const int n = e.bcount[lo + k];
    for (int i0 = 0; i0 < n; i0 += 64) e.nitems.push_back({(int32_t)k, i0, std::min(64, n - i0), 0});
  }
  dfree(e.d_near);
  e.d_near = upload(e.nitems, e.stream);
  H3D_HIP(hipStreamSynchronize(e.stream));
}

void ensure_scratch(Engine& e, int nrhs) {
  int64_t T = e.T, np = e.plan.npartial;
  if (e.mix.ready) {
    // target- and source-order t of every class (mix_t_layout)
    T = std::max<int64_t>(T, mix_t_doubles(e.mix));
    for (const FactorClass& c : e.mix.cls)
      np = std::max<int64_t>(np, std::max(c.plan.npartial, c.mpartial));
  }
  T = std::max<int64_t>(T, 1);
  np = std::max<int64_t>(np, 1);
  if (e.max_nrhs >= nrhs && e.d_xp && e.t_cap >= T && e.part_cap >= np) return;
  nrhs = std::max(nrhs, e.max_nrhs);
  dfree(e.d_xp);
  dfree(e.d_t);
  dfree(e.d_part);
  dfree(e.d_ylev);
  dfree(e.d_io);
  dfree(e.d_xf);
  // x is read in 16-byte vectors that may run up to 7 elements past a box
  e.d_xp = dalloc<double>(e.N * nrhs + 64);
  e.d_xf = dalloc<float>(e.N * nrhs + 64);
  H3D_HIP(hipMemsetAsync(e.d_xp, 0, sizeof(double) * (e.N * nrhs + 64), e.stream));
  H3D_HIP(hipMemsetAsync(e.d_xf, 0, sizeof(float) * (e.N * nrhs + 64), e.stream));
  e.d_t = dalloc<double>(T * nrhs);
  e.d_part = dalloc<double>(np * std::max(nrhs, 2));
  e.d_ylev = dalloc<double>(e.N * e.nL * nrhs);
  e.d_io = dalloc<double>(2 * e.N * nrhs);
  dfree(e.d_ynear);
  e.d_ynear = dalloc<double>(e.N);
  if (!e.side) {  // the near field's stream (default priority) and its fork / join events
    H3D_HIP(hipStreamCreateWithFlags(&e.side, hipStreamNonBlocking));
    H3D_HIP(hipEventCreateWithFlags(&e.ev_fork, hipEventDisableTiming));
    H3D_HIP(hipEventCreateWithFlags(&e.ev_join, hipEventDisableTiming));
  }
  if (!e.side2) {
    H3D_HIP(hipStreamCreateWithFlags(&e.side2, hipStreamNonBlocking));
    H3D_HIP(hipEventCreateWithFlags(&e.ev_join2, hipEventDisableTiming));
  }
  e.max_nrhs = nrhs;
  e.t_cap = T;
  e.part_cap = np;
}

// Builds the precision-`prec` copy of every chunk from the fp64 copy (or the
// fp32 one when fp64 is not resident).
void prepare_precision(Engine& e, int prec, bool move) {
  if (e.fac_bytes[prec] > 0 || e.chunks.empty()) return;
  int src = -1;
  for (int p : {0, 1})
    if (p != prec && e.fac_bytes[p] > 0) {
      src = p;
      break;
    }
  if (src < 0) throw Error(H3D_ERR_STATE, "no factor copy to convert from");
  hipStream_t st = e.stream;
  int64_t bytes = 0;
  size_t done = 0;  // chunks converted so far
  try {
    for (Chunk& c : e.chunks) {
      const int64_t n = c.nU + c.nV + kChunkTail;
      dmalloc(&c.data[prec], (size_t)n * elem_size(prec));
      if (n > 0) {
        if (prec == H3D_PREC_FP32)
          convert_kernel<float><<<2048, 256, 0, st>>>(c.data[src], src, n, (float*)c.data[prec]);
        else if (prec == H3D_PREC_BF16)
          convert_kernel<uint16_t><<<2048, 256, 0, st>>>(c.data[src], src, n,
                                                         (uint16_t*)c.data[prec]);
        else
          convert_kernel<double><<<2048, 256, 0, st>>>(c.data[src], src, n, (double*)c.data[prec]);
        H3D_LAUNCH_CHECK();
      }
      bytes += (c.nU + c.nV) * elem_size(prec);
      if (move) {  // free the other copies of this chunk right away (peak = one chunk)
        H3D_HIP(hipStreamSynchronize(st));
        for (int p = 0; p < 3; ++p)
          if (p != prec && c.data[p]) {
            (void)hipFree(c.data[p]);
            c.data[p] = nullptr;
          }
      }
      ++done;
    }
    H3D_HIP(hipStreamSynchronize(st));
    upload_chunk_table(e);  // replaces the tables only once all new ones exist
    if (move)
      for (int p = 0; p < 3; ++p) e.fac_bytes[p] = 0;
    e.fac_bytes[prec] = std::max<int64_t>(bytes, 1);
  } catch (...) {
    (void)hipStreamSynchronize(st);
    if (!move || done == 0) {
      // nothing of the other copies was freed and the tables are the old ones:
      // drop this call's new copy
      for (Chunk& c : e.chunks)
        if (c.data[prec]) (void)hipFree(c.data[prec]), c.data[prec] = nullptr;
      throw;
    }
    // a move that stopped part way: its first chunks exist only in the new
    // precision, the others only in the old ones, so no copy is whole.  Drop
    // every factor (the handle then answers H3D_ERR_STATE until assembled
    // again) rather than leave a table pointing at freed memory.
    for (Chunk& c : e.chunks)
      for (int p = 0; p < 3; ++p)
        if (c.data[p]) (void)hipFree(c.data[p]), c.data[p] = nullptr;
    for (int p = 0; p < 3; ++p) e.fac_bytes[p] = 0;
    e.assembled = false;
    e.factors_lost = true;
    for (int p = 0; p < 3; ++p) dfree(e.d_ubase[p]), dfree(e.d_vbase[p]);
    throw Error(H3D_ERR_OOM, "a move of the factor copy ran out of device memory part way; the "
                             "factors are lost (create the handle again)");
  }
}

void release_precision(Engine& e, int prec) {
  int others = 0;
  for (int p = 0; p < 3; ++p)
    if (p != prec && e.fac_bytes[p] > 0 && p != 2) ++others;
  if (e.fac_bytes[prec] > 0 && others == 0 && prec != 2)
    throw Error(H3D_ERR_STATE, "cannot release the last fp64/fp32 factor copy");
  H3D_HIP(hipStreamSynchronize(e.stream));
  for (Chunk& c : e.chunks) {
    if (c.data[prec]) (void)hipFree(c.data[prec]);
    c.data[prec] = nullptr;
  }
  e.fac_bytes[prec] = 0;
  upload_chunk_table(e);
}

namespace {

// One factor copy as the far-field kernels see it.
struct FactorView {
  const Plan* plan;
  const P2Item* p2;     // phase-2 items of this copy (device) and their count
  int n2;
  const BlockDev* blk;
  const int32_t* btoff;
  void* const* ubase;
  void* const* vbase;
  const double* scale;  // per t entry of this copy, or null
  int accumulate;       // phase 2 adds into y_level (not the first copy)
  int key;              // autotuning slot
  const FactorClass* fc;  // recompressed class (source-grouped phase 1), or null
  void* tsrc;           // that class's source-order t (phase-1 output)
};

FactorView aca_view(Engine& e, int prec, bool acc64) {
  const P2Item* p2 = prec == 0 ? e.plan.d_p2 : e.d_p2_narrow[prec - 1];
  const int n2 = (int)(prec == 0 ? e.plan.p2.size() : e.p2_narrow[prec - 1].size());
  return {&e.plan, p2, n2, e.d_blk, e.d_btoff, e.d_ubase[prec], e.d_vbase[prec], nullptr, 0,
          prec + (acc64 && prec != 0 ? 3 : 0), nullptr, nullptr};
}

// Class c of the mixed copy: class 0 (fp32 storage) accumulates in fp64;
// classes 1 and 2 (fp16 and fp8 storage, a small part of the operator) work
// in fp32 with their t as floats.
// Scratch t layout of the mixed copy (bytes from d_t, every part 16-byte
// aligned for the gathers' vector stores): each class's target-order t, then
// each class's source-order t (phase-1 output).
inline double* class_td(Engine& e, int c) {  // classes 0 and 3 (fp64 sums)
  return (double*)((char*)e.d_t + mix_t_layout(e.mix).t[c]);
}
inline float* class_tf(Engine& e, int c) {  // classes 1 and 2 (fp32 sums)
  return (float*)((char*)e.d_t + mix_t_layout(e.mix).t[c]);
}
inline void* class_tsrc(Engine& e, int c) {
  return (void*)((char*)e.d_t + mix_t_layout(e.mix).src[c]);
}

// Classes 0 and 3 share one tuning slot, classes 1 and 2 another (fp64 and
// fp32 sums).  Class 0's phase 2 stores y, the others add to it.
FactorView class_view(Engine& e, int c) {
  const FactorClass& fc = e.mix.cls[c];
  return {&fc.plan, fc.plan.d_p2, (int)fc.plan.p2.size(), fc.d_blk, fc.d_btoff, fc.d_ubase,
          fc.d_vbase, e.mix.d_scale + fc.tbase, c, class_acc64(c) ? 6 : 7, &fc, class_tsrc(e, c)};
}

// Source-grouped phase 1 of a mixed class with a load width of `bytes` per lane.
template <class S, class A>
void launch_mp1(Engine& e, const FactorView& f, int bytes, int kmax, const A* x, A* t, A* part) {
  const FactorClass& fc = *f.fc;
  const int n1 = (int)fc.mp1.size();
  if (!n1) return;
  // the items span the rows of 64 lanes x KMAX vectors of 16 bytes (12 for
  // fp24: four entries), which only the 16-byte loads cover
  if (bytes != 16) throw Error(H3D_ERR_INVALID, "the mixed phase 1 takes 16-byte loads");
  constexpr int es = (int)sizeof(S);
  mphase1_kernel<S, A, es == 3 ? 4 : 16 / es><<<wave_grid(n1), 256, 0, e.stream>>>(
      fc.d_mp1, n1, f.vbase, x, (A*)f.tsrc, part, kmax, nullptr);
  H3D_LAUNCH_CHECK();
}

// Workgroups of a persistent kernel: as many as stay resident on the device
// at once (256 threads each), at most enough for `items` waves.
template <class K>
unsigned resident_grid(K kernel, int64_t items) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, unsigned>> known;
  unsigned full = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    for (auto& k : known)
      if (k.first == (const void*)kernel) full = k.second;
    if (!full) {
      int per_cu = 0, cus = 0, dev = 0;
      H3D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
      H3D_HIP(hipGetDevice(&dev));
      H3D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      full = (unsigned)std::max(1, per_cu * cus);
      known.emplace_back((const void*)kernel, full);
    }
  }
  return std::min(full, wave_grid(items));
}

// Phase-1 launch over an item table with a load width of `bytes` per lane;
// with `queue`, the persistent kernel (its counter cleared first).
template <class S, class A, int NV>
void p1_launch(const P1Item* it, int n1, void* const* vbase, const BlockDev* blk, const A* x,
               A* t, A* part, const double* scale, unsigned* queue, hipStream_t st) {
  if (queue) {
    queue_reset_kernel<<<1, 64, 0, st>>>(queue, 1);
    auto k = phase1_queue_kernel<S, A, NV>;
    k<<<resident_grid(k, n1), 256, 0, st>>>(it, n1, vbase, blk, x, t, part, scale, queue);
  } else {
    phase1_kernel<S, A, NV><<<wave_grid(n1), 256, 0, st>>>(it, n1, vbase, blk, x, t, part, scale);
  }
}

template <class S, class A>
void p1_raw(const P1Item* it, int n1, void* const* vbase, const BlockDev* blk, const A* x, A* t,
            A* part, const double* scale, int bytes, hipStream_t st, unsigned* queue = nullptr) {
  if (!n1) return;
  constexpr int es = (int)sizeof(S);
  if constexpr (es == 3) {  // fp24 is a mixed class only (source-grouped phase 1)
    throw Error(H3D_ERR_INVALID, "no block-walking phase 1 for fp24 entries");
  } else if (bytes >= 16 || 16 / es == 1 && bytes >= es)
    p1_launch<S, A, 16 / es>(it, n1, vbase, blk, x, t, part, scale, queue, st);
  else if (bytes == 8 && 8 / es >= 1)
    p1_launch<S, A, (8 / es > 0 ? 8 / es : 1)>(it, n1, vbase, blk, x, t, part, scale, queue, st);
  else if (bytes == 4 && 4 / es >= 1)
    p1_launch<S, A, (4 / es > 0 ? 4 / es : 1)>(it, n1, vbase, blk, x, t, part, scale, queue, st);
  else
    p1_launch<S, A, (2 / es > 0 ? 2 / es : 1)>(it, n1, vbase, blk, x, t, part, scale, queue, st);
  H3D_LAUNCH_CHECK();
}

template <class S, class A, int NV>
void p2_launch(const P2Item* it, int n2, void* const* ubase, const A* t, double* ylev, int64_t N,
               int acc, unsigned* queue, hipStream_t st) {
  if (queue) {
    queue_reset_kernel<<<1, 64, 0, st>>>(queue, 1);
    auto k = phase2_queue_kernel<S, A, NV>;
    k<<<resident_grid(k, n2), 256, 0, st>>>(it, n2, ubase, t, ylev, N, acc, queue);
  } else {
    phase2_kernel<S, A, NV><<<wave_grid(n2), 256, 0, st>>>(it, n2, ubase, t, ylev, N, acc);
  }
}

// Phase-2 launch over an item table with a load width of `bytes` per lane.
template <class S, class A>
void p2_raw(const P2Item* it, int n2, void* const* ubase, const A* t, double* ylev, int64_t N,
            int acc, int bytes, hipStream_t st, unsigned* queue = nullptr) {
  if (!n2) return;
  constexpr int es = (int)sizeof(S);
  if constexpr (es == 3)  // fp24: four entries (12 bytes) per lane whatever the width
    p2_launch<S, A, 4>(it, n2, ubase, t, ylev, N, acc, queue, st);
  else if (bytes >= 16)
    p2_launch<S, A, 16 / es>(it, n2, ubase, t, ylev, N, acc, queue, st);
  else if (bytes == 8)
    p2_launch<S, A, (8 / es > 0 ? 8 / es : 1)>(it, n2, ubase, t, ylev, N, acc, queue, st);
  else if (bytes == 4)
    p2_launch<S, A, (4 / es > 0 ? 4 / es : 1)>(it, n2, ubase, t, ylev, N, acc, queue, st);
  else
    p2_launch<S, A, (2 / es > 0 ? 2 / es : 1)>(it, n2, ubase, t, ylev, N, acc, queue, st);
  H3D_LAUNCH_CHECK();
}

// Phase-1 / phase-2 launch with a load width of `bytes` per lane.
template <class S, class A>
void launch_p1(Engine& e, const FactorView& f, int bytes, const A* x, A* t, A* part) {
  if (f.fc) {
    launch_mp1<S, A>(e, f, bytes, e.tune_k[f.key], x, t, part);
    return;
  }
  p1_raw<S, A>(f.plan->d_p1, (int)f.plan->p1.size(), f.vbase, f.blk, x, t, part, f.scale, bytes,
               e.stream);
}

template <class A>
void launch_p1_reduce(Engine& e, const FactorView& f, const A* part, A* t) {
  if (f.fc) {
    const FactorClass& fc = *f.fc;
    if (!fc.mp1r.empty())
      mreduce_kernel<A><<<wave_grid(fc.mp1r.size()), 256, 0, e.stream>>>(
          fc.d_mp1r, (int)fc.mp1r.size(), part, (A*)f.tsrc, nullptr);
    H3D_LAUNCH_CHECK();
    if (fc.T > 0)
      tgather_kernel<A><<<(fc.T / 4 + 256) / 256, 256, 0, e.stream>>>(
          fc.d_tmap, fc.T, (const A*)f.tsrc,
          sizeof(A) == 4 ? (const A*)fc.d_scale32 : (const A*)f.scale, t);
    H3D_LAUNCH_CHECK();
    return;
  }
  const Plan& p = *f.plan;
  if (p.p1r.empty()) return;
  reduce_kernel<A><<<wave_grid(p.p1r.size()), 256, 0, e.stream>>>(p.d_p1r, (int)p.p1r.size(), part,
                                                                   t, f.scale);
  H3D_LAUNCH_CHECK();
}

template <class S, class A>
void launch_p2(Engine& e, const FactorView& f, int bytes, const A* t) {
  p2_raw<S, A>(f.p2, f.n2, f.ubase, t, e.d_ylev, e.N, f.accumulate, bytes, e.stream);
}

// `mark` (if set) is recorded right after the phase-1 kernel itself.
template <class S, class A>
void run_p1(Engine& e, const FactorView& f, const A* x, A* t, A* part, hipEvent_t mark) {
  launch_p1<S, A>(e, f, e.tune_p1[f.key], x, t, part);
  if (mark) H3D_HIP(hipEventRecord(mark, e.stream));
  launch_p1_reduce<A>(e, f, part, t);
}

template <class S, class A>
void run_p2(Engine& e, const FactorView& f, const A* t) {
  launch_p2<S, A>(e, f, e.tune_p2[f.key], t);
}

// Average ms of `reps` launches of phase 1 (which = 1) or phase 2 (which = 2).
template <class S, class A>
double time_phase(Engine& e, const FactorView& f, int which, int bytes, int reps, const A* x, A* t,
                  A* part) {
  hipEvent_t a, b;
  H3D_HIP(hipEventCreate(&a));
  H3D_HIP(hipEventCreate(&b));
  // one untimed launch first
  if (which == 1) launch_p1<S, A>(e, f, bytes, x, t, part);
  else launch_p2<S, A>(e, f, bytes, t);
  H3D_HIP(hipEventRecord(a, e.stream));
  for (int r = 0; r < reps; ++r) {
    if (which == 1) launch_p1<S, A>(e, f, bytes, x, t, part);
    else launch_p2<S, A>(e, f, bytes, t);
  }
  H3D_HIP(hipEventRecord(b, e.stream));
  H3D_HIP(hipEventSynchronize(b));
  float ms = 0;
  H3D_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// Times every load width (and, for a mixed class's phase 1, rows per lane of
// 4 or 1 load vectors) and keeps the fastest; ms gets the best time per width.
template <class S, class A>
void tune_mode(Engine& e, const FactorView& f, int reps, const A* x, A* t, A* part, double* ms) {
  const int widths[4] = {16, 8, 4, 2};
  for (int which = 1; which <= 2; ++which) {
    double best = 1e300;
    int bw = 16, bk = e.tune_k[f.key];
    for (int q = 0; q < 4; ++q) {
      const int w = widths[q];
      double v = -1;
      if (w >= (int)sizeof(S) && !(which == 1 && f.fc && w != 16)) {
        const int nk = (which == 1 && f.fc) ? 3 : 1;
        for (int ki = 0; ki < nk; ++ki) {
          if (nk == 3) e.tune_k[f.key] = ki == 0 ? 4 : ki == 1 ? 1 : 0;
          const double tv = time_phase<S, A>(e, f, which, w, reps, x, t, part);
          if (v < 0 || tv < v) v = tv;
          if (tv < best) {
            best = tv;
            bw = w;
            if (which == 1 && f.fc) bk = e.tune_k[f.key];
          }
        }
      }
      if (ms) ms[(which - 1) * 4 + q] = v;
    }
    (which == 1 ? e.tune_p1 : e.tune_p2)[f.key] = bw;
    if (which == 1 && f.fc) e.tune_k[f.key] = bk;
  }
}

// Transposed phase 1, t = U^T x over the U columns of every target box: the
// mixed phase-1 kernel on the plan's items (U chunks for V chunks, the box's
// rows for a source's), each t written at its own index (times the copy's
// factor for a recompressed class), row-split partials summed after it.
template <class S, class A>
void run_tp1(Engine& e, const FactorView& f, const A* x, A* t, A* part, hipEvent_t mark) {
  const Plan& p = *f.plan;
  const int n1 = (int)p.tp1.size();
  hipStream_t st = e.stream;
  constexpr int es = (int)sizeof(S);
  if (n1)
    mphase1_kernel<S, A, es == 3 ? 4 : 16 / es><<<wave_grid(n1), 256, 0, st>>>(
        p.d_tp1, n1, f.ubase, x, t, part, 0, f.scale);
  H3D_LAUNCH_CHECK();
  if (mark) H3D_HIP(hipEventRecord(mark, st));
  if (!p.tp1r.empty())
    mreduce_kernel<A><<<wave_grid(p.tp1r.size()), 256, 0, st>>>(p.d_tp1r, (int)p.tp1r.size(), part,
                                                                t, f.scale);
  H3D_LAUNCH_CHECK();
}

// Transposed phase 2, y_s = sum over the blocks with source s of V_b t_b.  A
// recompressed class keeps each source box's V columns together (V_s), so t
// goes to source order and the phase-2 kernel reads V_s as it reads a target
// box's U; the ACA copies' V lie per block, read row by row.
template <class S, class A>
void run_tp2(Engine& e, const FactorView& f, const A* t) {
  if (f.fc) {
    const FactorClass& fc = *f.fc;
    const int n2 = (int)fc.tp2v.size();
    if (!n2) return;
    A* tsrc = (A*)class_tsrc(e, (int)(f.fc - e.mix.cls));
    if (fc.T > 0)
      tscatter_kernel<A><<<(fc.T + 255) / 256, 256, 0, e.stream>>>(fc.d_tmap, fc.T, t, tsrc);
    H3D_LAUNCH_CHECK();
    p2_raw<S, A>(fc.d_tp2v, n2, f.vbase, tsrc, e.d_ylev, e.N, f.accumulate, e.tune_p2[f.key],
                 e.stream);
    return;
  }
  const Plan& p = *f.plan;
  const int n2 = (int)p.tp2.size();
  if (n2)
    tphase2_kernel<S, A><<<wave_grid(n2), 256, 0, e.stream>>>(p.d_tp2, n2, f.vbase, f.blk,
                                                              e.d_by_src, f.btoff, t, e.d_ylev,
                                                              e.N, f.accumulate);
  H3D_LAUNCH_CHECK();
}

// Class 2 of the mixed copy holds e4m3 entries, or int8 ones (kMixInt8):
// f(storage tag).
template <class F>
void with_class2(const Engine& e, F f) {
  if (e.mix.int8) f(i8s{});
  else f(f8e4m3{});
}

// Launches within its scope go to stream `s` (the launch helpers use e.stream).
struct OnStream {
  Engine& e;
  hipStream_t saved;
  OnStream(Engine& en, hipStream_t s) : e(en), saved(en.stream) { e.stream = s; }
  ~OnStream() { e.stream = saved; }
};

// Phase 1 (which = 1) or phase 2 (which = 2) of one copy, forward or transpose.
template <class S, class A>
void run_phase(Engine& e, const FactorView& f, int which, bool transpose, const A* x, A* t, A* part,
               hipEvent_t mark = nullptr) {
  if (!transpose) {
    if (which == 1) run_p1<S, A>(e, f, x, t, part, mark);
    else run_p2<S, A>(e, f, t);
  } else {
    if (which == 1) run_tp1<S, A>(e, f, x, t, part, mark);
    else run_tp2<S, A>(e, f, t);
  }
}

void check_mode(const Engine& e, int mode, int& prec, bool& acc64) {
  prec = mode & 0xff;
  if (prec < 0 || prec > kPrecMixed || (mode & ~(0xff | H3D_ACC_FP64)))
    throw Error(H3D_ERR_INVALID, "unknown precision");
  acc64 = prec == H3D_PREC_FP64 || prec == kPrecMixed || (mode & H3D_ACC_FP64);
}

void ensure_copy(Engine& e, int prec) {
  if (prec == kPrecMixed) {
    if (!e.mix.ready) build_mixed(e, 5.0 * e.tol);
  } else if (e.fac_bytes[prec] == 0) {
    prepare_precision(e, prec);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers over raw tables of the ACA-copy product (the single product and
// the instance batch share them).  x, t and the partial sums are fp64 with
// fp64 accumulation (xp), else fp32 (xf, and t / part read as floats).

NearTables near_tables(Engine& e) {
  NearTables nt;
  const int64_t lo = level_offset(e.nL);
  nt.items = e.d_near;
  nt.n = (int)e.nitems.size();
  nt.nbr = e.lev[e.nL].d_nbr;
  nt.bstart = e.d_bstart + lo;
  nt.bcount = e.d_bcount + lo;
  nt.px = e.d_px, nt.py = e.d_py, nt.pz = e.d_pz;
  nt.lx = e.d_lx, nt.ly = e.d_ly, nt.lz = e.d_lz;
  nt.leafc = e.d_leafc;
  nt.perm = e.d_perm;
  nt.kd = e.kern;
  nt.xp = e.d_xp;
  nt.xf = e.d_xf;
  nt.ynear = e.d_ynear;
  return nt;
}

void far_gather(const double* b, const int32_t* perm, int64_t N, double* xp, float* xf,
                hipStream_t st) {
  if (N <= 0) return;
  gather_kernel<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(b, perm, N, xp, xf);
  H3D_LAUNCH_CHECK();
}

void far_zero(double* p, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  zero_kernel<<<(unsigned)std::min<int64_t>(8192, (n + 255) / 256), 256, 0, st>>>(p, n);
  H3D_LAUNCH_CHECK();
}

void far_phase1(int prec, bool acc64, int bytes, const P1Item* items, int n, void* const* vbase,
                const BlockDev* blk, const double* xp, const float* xf, double* t, double* part,
                hipStream_t st, unsigned* queue) {
  if (acc64) {
    if (prec == H3D_PREC_FP64)
      p1_raw<double, double>(items, n, vbase, blk, xp, t, part, nullptr, bytes, st, queue);
    else if (prec == H3D_PREC_FP32)
      p1_raw<float, double>(items, n, vbase, blk, xp, t, part, nullptr, bytes, st, queue);
    else
      p1_raw<uint16_t, double>(items, n, vbase, blk, xp, t, part, nullptr, bytes, st, queue);
  } else if (prec == H3D_PREC_FP32) {
    p1_raw<float, float>(items, n, vbase, blk, xf, (float*)t, (float*)part, nullptr, bytes, st,
                         queue);
  } else {
    p1_raw<uint16_t, float>(items, n, vbase, blk, xf, (float*)t, (float*)part, nullptr, bytes, st,
                            queue);
  }
}

void far_reduce(bool acc64, const P1Reduce* rs, int n, const double* part, double* t,
                hipStream_t st) {
  if (n <= 0) return;
  if (acc64)
    reduce_kernel<double><<<wave_grid(n), 256, 0, st>>>(rs, n, part, t, nullptr);
  else
    reduce_kernel<float><<<wave_grid(n), 256, 0, st>>>(rs, n, (const float*)part, (float*)t, nullptr);
  H3D_LAUNCH_CHECK();
}

void far_phase2(int prec, bool acc64, int bytes, const P2Item* items, int n, void* const* ubase,
                const double* t, double* ylev, int64_t N, hipStream_t st, unsigned* queue) {
  if (acc64) {
    if (prec == H3D_PREC_FP64)
      p2_raw<double, double>(items, n, ubase, t, ylev, N, 0, bytes, st, queue);
    else if (prec == H3D_PREC_FP32)
      p2_raw<float, double>(items, n, ubase, t, ylev, N, 0, bytes, st, queue);
    else
      p2_raw<uint16_t, double>(items, n, ubase, t, ylev, N, 0, bytes, st, queue);
  } else if (prec == H3D_PREC_FP32) {
    p2_raw<float, float>(items, n, ubase, (const float*)t, ylev, N, 0, bytes, st, queue);
  } else {
    p2_raw<uint16_t, float>(items, n, ubase, (const float*)t, ylev, N, 0, bytes, st, queue);
  }
}

// Near field of an ACA-copy product (every mode but the mixed copy): fp64
// factors take the exact entries, fp32/bf16 factors with fp64 accumulation the
// Newton-refined 1/r (or the exact entries of other kernels), fp32 sums the
// fp32 entries; a plugin kernel's entries are fp64 in every mode, and the
// transpose of a plugin evaluates entry(x_j, x_i).
void far_near(const NearTables& nt, int prec, bool acc64, bool transpose, hipStream_t st) {
  const int nn = nt.n;
  if (nn <= 0) return;
  const KernelDev kd = nt.kd;
  const unsigned g = wave_grid(nn);
  const bool invr = kd.q == 7 || kd.q == 16;
  if (kd.q == kQPlugin && transpose)
    near64_kernel<Plugin, true><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount,
                                                   nt.px, nt.py, nt.pz, nt.perm, kd, nt.xp, nt.ynear);
  else if (kd.q == kQPlugin)
    near64_kernel<Plugin><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount, nt.px,
                                             nt.py, nt.pz, nt.perm, kd, nt.xp, nt.ynear);
  else if (acc64 && prec != H3D_PREC_FP64 && invr)
    near_nr_kernel<<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount, nt.px, nt.py,
                                      nt.pz, kd, nt.xp, nt.ynear);
  else if (acc64 && invr)
    near64_kernel<InvR><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount, nt.px,
                                           nt.py, nt.pz, nt.perm, kd, nt.xp, nt.ynear);
  else if (acc64)
    near64_kernel<AnyQ><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount, nt.px,
                                           nt.py, nt.pz, nt.perm, kd, nt.xp, nt.ynear);
  else if (invr)
    near32_kernel<float, InvR32><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount,
                                                    nt.lx, nt.ly, nt.lz, nt.leafc, nt.perm, kd,
                                                    nt.xf, nt.ynear);
  else
    near32_kernel<float, AnyQ32><<<g, 256, 0, st>>>(nt.items, nn, nt.nbr, nt.bstart, nt.bcount,
                                                    nt.lx, nt.ly, nt.lz, nt.leafc, nt.perm, kd,
                                                    nt.xf, nt.ynear);
  H3D_LAUNCH_CHECK();
}

void autotune(Engine& e, int mode, int reps, double* ms) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "autotune before assemble");
  int prec;
  bool acc64;
  check_mode(e, mode, prec, acc64);
  if (reps < 1) throw Error(H3D_ERR_INVALID, "reps must be >= 1");
  ensure_copy(e, prec);
  ensure_scratch(e, 1);
  // timing does not depend on the values: x is whatever the last matvec left
  if (prec == kPrecMixed) {
    tune_mode<float, double>(e, class_view(e, 0), reps, e.d_xp, class_td(e, 0), e.d_part, ms);
    // the fp16 class's settings serve the fp8 class too (one slot)
    tune_mode<_Float16, float>(e, class_view(e, 1), reps, e.d_xf, class_tf(e, 1),
                               (float*)e.d_part, ms ? ms + 8 : nullptr);
  } else if (acc64) {
    const FactorView f = aca_view(e, prec, true);
    if (prec == H3D_PREC_FP64) tune_mode<double, double>(e, f, reps, e.d_xp, e.d_t, e.d_part, ms);
    else if (prec == H3D_PREC_FP32) tune_mode<float, double>(e, f, reps, e.d_xp, e.d_t, e.d_part, ms);
    else tune_mode<uint16_t, double>(e, f, reps, e.d_xp, e.d_t, e.d_part, ms);
  } else {
    const FactorView f = aca_view(e, prec, false);
    float* t = (float*)e.d_t;
    float* part = (float*)e.d_part;
    if (prec == H3D_PREC_FP32) tune_mode<float, float>(e, f, reps, e.d_xf, t, part, ms);
    else tune_mode<uint16_t, float>(e, f, reps, e.d_xf, t, part, ms);
  }
  H3D_HIP(hipStreamSynchronize(e.stream));
}

void prepare_mode(Engine& e, int mode) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "matvec before assemble");
  int prec;
  bool acc64;
  check_mode(e, mode, prec, acc64);
  ensure_copy(e, prec);
  ensure_scratch(e, 1);
  if (e.timing && !e.ev[0])
    for (auto& ev : e.ev) H3D_HIP(hipEventCreate(&ev));
}

void matvec_device(Engine& e, const double* d_b, double* d_y, int nrhs, int mode, bool transpose) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "matvec before assemble");
  int prec;
  bool acc64;
  check_mode(e, mode, prec, acc64);
  ensure_copy(e, prec);
  ensure_scratch(e, 1);
  if (block_width(e, nrhs, prec, transpose, acc64) > 1) {
    matvec_block(e, d_b, d_y, nrhs, prec, acc64, transpose);
    return;
  }
  hipStream_t st = e.stream;
  const int64_t N = e.N;
  const unsigned gN = (unsigned)((N + 255) / 256);
  if (e.timing && !e.ev[0])
    for (auto& ev : e.ev) H3D_HIP(hipEventCreate(&ev));
  const bool mixed = prec == kPrecMixed;
  // Near field into its own buffer by position (symmetric kernels: the same
  // for A^T).  Without per-kernel timing it runs on the side stream, forked
  // after the gather and joined before the combine, so that this VALU-bound
  // kernel overlaps the HBM-bound far field; with timing, and while a graph
  // of products is captured, it runs in line.
  const int nn = (int)e.nitems.size();
  const bool overlap = nn && !e.timing && e.overlap_near;
  // The mixed copy's classes on two streams (without per-kernel timing, as
  // the near field): each class's phase 2 needs only its own t, and the
  // phase-2 kernels, which add into y, keep the in-line order.
  const bool split = mixed && !e.timing && e.overlap_near && e.side2;
  auto launch_near = [&](hipStream_t ns) {
    const int64_t lo = level_offset(e.nL);
    const KernelDev kd = e.kern;
    const unsigned g = wave_grid(nn);
    // mixed: fp32 entries summed in fp64 (their error, ~2e-8 of ||A x|| at
    // 2^20 points, sits below the copy's fp32 rounding floor); the ACA copies
    // as far_near chooses
    const bool invr = kd.q == 7 || kd.q == 16;
    if (mixed && kd.q != kQPlugin) {
      if (invr)
        near32_kernel<double, InvR32><<<g, 256, 0, ns>>>(
            e.d_near, nn, e.lev[e.nL].d_nbr, e.d_bstart + lo, e.d_bcount + lo, e.d_lx, e.d_ly,
            e.d_lz, e.d_leafc, e.d_perm, kd, e.d_xp, e.d_ynear);
      else
        near32_kernel<double, AnyQ32><<<g, 256, 0, ns>>>(
            e.d_near, nn, e.lev[e.nL].d_nbr, e.d_bstart + lo, e.d_bcount + lo, e.d_lx, e.d_ly,
            e.d_lz, e.d_leafc, e.d_perm, kd, e.d_xp, e.d_ynear);
      H3D_LAUNCH_CHECK();
      return;
    }
    far_near(near_tables(e), prec, acc64, transpose, ns);
  };
  for (int r = 0; r < nrhs; ++r) {
    const double* b = d_b + (int64_t)r * N;
    double* y = d_y + (int64_t)r * N;
    if (e.timing) H3D_HIP(hipEventRecord(e.ev[0], st));
    gather_kernel<<<gN, 256, 0, st>>>(b, e.d_perm, N, e.d_xp, (acc64 && !mixed) ? nullptr : e.d_xf);
    H3D_LAUNCH_CHECK();
    {
      const int64_t n = N * e.nL;
      zero_kernel<<<(unsigned)std::min<int64_t>(4096, (n + 255) / 256), 256, 0, st>>>(e.d_ylev, n);
      H3D_LAUNCH_CHECK();
    }
    if (overlap || split) H3D_HIP(hipEventRecord(e.ev_fork, st));
    if (overlap) {
      H3D_HIP(hipStreamWaitEvent(e.side, e.ev_fork, 0));
      launch_near(e.side);
      H3D_HIP(hipEventRecord(e.ev_join, e.side));
    }
    // event slots: 1 phase 1 start, 8 after the first copy's phase-1 kernel,
    // 13 after its phase 1 (reduce, gather), 15 after the fourth copy's
    // phase-1 kernel, 6 after its phase 1, 9 after the second copy's phase-1
    // kernel, 10 after its phase 1, 12 after the third copy's phase-1 kernel,
    // 2 phase 2 start, 14 after the first copy's phase 2, 7 after the
    // fourth's, 11 after the second's, 3 near start, 4 combine start, 5 end
    // (the copies: the mixed copy's classes 0, 1, 2, 3 = fp32, fp16, fp8,
    // fp24, run in the order 0, 3, 1, 2; one copy in the other modes)
    if (split) {
      // Classes 0 and 3 (fp64 sums), phase 1 then phase 2, on side2 while
      // classes 1 and 2 run phase 1 here; their phase 2 waits for side2's,
      // so y takes the classes in the in-line order 0, 3, 1, 2 (the same
      // sums to the bit).  Each stream has its own region of partial sums.
      H3D_HIP(hipStreamWaitEvent(e.side2, e.ev_fork, 0));
      {
        OnStream on(e, e.side2);
        double* part2 = e.d_part + e.part_cap;
        for (int which = 1; which <= 2; ++which) {
          run_phase<float, double>(e, class_view(e, 0), which, transpose, e.d_xp, class_td(e, 0),
                                   part2);
          run_phase<f24, double>(e, class_view(e, 3), which, transpose, e.d_xp, class_td(e, 3),
                                 part2);
        }
        H3D_HIP(hipEventRecord(e.ev_join2, e.side2));
      }
      float* part = (float*)e.d_part;
      for (int which = 1; which <= 2; ++which) {
        if (which == 2) H3D_HIP(hipStreamWaitEvent(st, e.ev_join2, 0));
        run_phase<_Float16, float>(e, class_view(e, 1), which, transpose, e.d_xf, class_tf(e, 1),
                                   part);
        with_class2(e, [&](auto s8) {
          run_phase<decltype(s8), float>(e, class_view(e, 2), which, transpose, e.d_xf,
                                         class_tf(e, 2), part);
        });
      }
    }
    for (int which = 1; which <= 2 && !split; ++which) {
      if (e.timing) H3D_HIP(hipEventRecord(e.ev[which], st));
      hipEvent_t m1 = (e.timing && which == 1) ? e.ev[8] : nullptr;
      hipEvent_t m2 = (e.timing && which == 1) ? e.ev[9] : nullptr;
      hipEvent_t m3 = (e.timing && which == 1) ? e.ev[12] : nullptr;
      hipEvent_t m4 = (e.timing && which == 1) ? e.ev[15] : nullptr;
      if (mixed) {
        run_phase<float, double>(e, class_view(e, 0), which, transpose, e.d_xp, class_td(e, 0),
                                 e.d_part, m1);
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[12 + which], st));
        run_phase<f24, double>(e, class_view(e, 3), which, transpose, e.d_xp, class_td(e, 3),
                               e.d_part, m4);
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[5 + which], st));
        run_phase<_Float16, float>(e, class_view(e, 1), which, transpose, e.d_xf, class_tf(e, 1),
                                   (float*)e.d_part, m2);
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[9 + which], st));
        with_class2(e, [&](auto s8) {
          run_phase<decltype(s8), float>(e, class_view(e, 2), which, transpose, e.d_xf,
                                         class_tf(e, 2), (float*)e.d_part, m3);
        });
      } else if (acc64) {
        const FactorView f = aca_view(e, prec, true);
        if (prec == H3D_PREC_FP64)
          run_phase<double, double>(e, f, which, transpose, e.d_xp, e.d_t, e.d_part, m1);
        else if (prec == H3D_PREC_FP32)
          run_phase<float, double>(e, f, which, transpose, e.d_xp, e.d_t, e.d_part, m1);
        else
          run_phase<uint16_t, double>(e, f, which, transpose, e.d_xp, e.d_t, e.d_part, m1);
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[12 + which], st));
        if (m4) H3D_HIP(hipEventRecord(m4, st));
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[5 + which], st));
        if (m2) H3D_HIP(hipEventRecord(m2, st));
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[9 + which], st));
        if (m3) H3D_HIP(hipEventRecord(m3, st));
      } else {
        const FactorView f = aca_view(e, prec, false);
        float* t = (float*)e.d_t;
        float* part = (float*)e.d_part;
        if (prec == H3D_PREC_FP32)
          run_phase<float, float>(e, f, which, transpose, e.d_xf, t, part, m1);
        else
          run_phase<uint16_t, float>(e, f, which, transpose, e.d_xf, t, part, m1);
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[12 + which], st));
        if (m4) H3D_HIP(hipEventRecord(m4, st));
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[5 + which], st));
        if (m2) H3D_HIP(hipEventRecord(m2, st));
        if (e.timing) H3D_HIP(hipEventRecord(e.ev[9 + which], st));
        if (m3) H3D_HIP(hipEventRecord(m3, st));
      }
    }
    if (e.timing) H3D_HIP(hipEventRecord(e.ev[3], st));
    if (nn && !overlap) launch_near(st);
    if (overlap) H3D_HIP(hipStreamWaitEvent(st, e.ev_join, 0));
    if (e.timing) H3D_HIP(hipEventRecord(e.ev[4], st));
    combine_kernel<<<gN, 256, 0, st>>>(e.d_ylev, e.nL, N, nn ? e.d_ynear : nullptr, e.d_iperm, b,
                                       e.alpha, e.kern.diag, e.d_diag, y);
    H3D_LAUNCH_CHECK();
    if (e.timing) {
      H3D_HIP(hipEventRecord(e.ev[5], st));
      H3D_HIP(hipEventSynchronize(e.ev[5]));
      float ms[15];
      H3D_HIP(hipEventElapsedTime(&ms[0], e.ev[1], e.ev[2]));  // phase 1 (+ partial reduce)
      H3D_HIP(hipEventElapsedTime(&ms[1], e.ev[2], e.ev[3]));  // phase 2
      H3D_HIP(hipEventElapsedTime(&ms[2], e.ev[3], e.ev[4]));  // near field
      H3D_HIP(hipEventElapsedTime(&ms[3], e.ev[4], e.ev[5]));  // combine
      H3D_HIP(hipEventElapsedTime(&ms[4], e.ev[0], e.ev[5]));  // whole matvec
      H3D_HIP(hipEventElapsedTime(&ms[5], e.ev[6], e.ev[10]));  // phase 1 of the second copy
      H3D_HIP(hipEventElapsedTime(&ms[6], e.ev[7], e.ev[11]));  // phase 2 of the second copy
      H3D_HIP(hipEventElapsedTime(&ms[7], e.ev[1], e.ev[8]));   // first copy's phase-1 kernel
      H3D_HIP(hipEventElapsedTime(&ms[8], e.ev[6], e.ev[9]));   // second copy's phase-1 kernel
      H3D_HIP(hipEventElapsedTime(&ms[9], e.ev[10], e.ev[2]));  // phase 1 of the third copy
      H3D_HIP(hipEventElapsedTime(&ms[10], e.ev[11], e.ev[3]));  // phase 2 of the third copy
      H3D_HIP(hipEventElapsedTime(&ms[11], e.ev[10], e.ev[12]));  // third copy's phase-1 kernel
      H3D_HIP(hipEventElapsedTime(&ms[12], e.ev[13], e.ev[6]));  // phase 1 of the fourth copy
      H3D_HIP(hipEventElapsedTime(&ms[13], e.ev[14], e.ev[7]));  // phase 2 of the fourth copy
      H3D_HIP(hipEventElapsedTime(&ms[14], e.ev[13], e.ev[15]));  // fourth copy's phase-1 kernel
      for (int q = 0; q < 15; ++q) e.ktime[q] += ms[q];
      e.ref_blas2_ms += ms[0] + ms[1];  // far field: the BLAS-2 applies
      e.ref_regen_ms += ms[2];          // near field: entries regenerated in the product
      e.kcount++;
    }
  }
}

// ---------------------------------------------------------------------------
// The allocation-free product (h3d_matvec_noalloc).  Every buffer a product
// uses beyond its input and output -- x by position (fp64 and fp32), t, the
// phase-1 partial sums, the per-level y and the near field -- is carved from
// this array instead of the handle's own scratch.  Each GPU the library runs
// on holds one copy until the process ends; every call on that GPU shares it
// and it keeps the last call's contents, so calls that overlap on one GPU race.
namespace {

constexpr int64_t kNoAllocBytes = int64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current GPU's copy of `scratch` (looked up once
// per device; host code never uses the array's own name as an address).
unsigned char* noalloc_scratch() {
  int dev = 0;
  H3D_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(H3D_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    H3D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(H3D_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return (unsigned char*)g_scratch_ptr[dev];
}

// Clears the 64 entries past N of x by position (fp64 and fp32): vector loads
// of a box's last rows may read them (the handle's own buffers are cleared
// once at allocation; the shared array keeps what the last call left).
__global__ void clear_x_tails(double* __restrict__ xp, float* __restrict__ xf, int64_t N) {
  const int i = threadIdx.x;
  if (i < 64) {
    xp[N + i] = 0.0;
    xf[N + i] = 0.0f;
  }
}

}  // namespace

int64_t noalloc_max_bytes() { return kNoAllocBytes; }

void matvec_noalloc(Engine& e, const double* d_b, double* d_y, int mode) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "matvec before assemble");
  int prec;
  bool acc64;
  check_mode(e, mode, prec, acc64);
  if (prec == kPrecMixed ? !e.mix.ready : e.fac_bytes[prec] == 0)
    throw Error(H3D_ERR_STATE, "the factor copy of this precision is not resident; build it "
                               "first (h3d_prepare_precision / h3d_compress)");
  // the sizes ensure_scratch gives the handle's own buffers for one vector
  int64_t T = e.T, np = e.plan.npartial;
  if (e.mix.ready) {
    T = std::max<int64_t>(T, mix_t_doubles(e.mix));
    for (const FactorClass& c : e.mix.cls)
      np = std::max<int64_t>(np, std::max(c.plan.npartial, c.mpartial));
  }
  T = std::max<int64_t>(T, 1);
  np = std::max<int64_t>(np, 1);
  const int64_t N = e.N;
  auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
  const int64_t o_xf = up(8 * (N + 64)), o_t = o_xf + up(4 * (N + 64)), o_part = o_t + up(8 * T),
                o_ylev = o_part + up(8 * np), o_yn = o_ylev + up(8 * N * e.nL),
                need = o_yn + up(8 * N);
  if (need > kNoAllocBytes)
    throw Error(H3D_ERR_INVALID, "this product needs " + std::to_string(need) +
                                     " bytes of scratch, more than the " +
                                     std::to_string(kNoAllocBytes) + " of h3d_matvec_noalloc");
  unsigned char* base = noalloc_scratch();
  // Swap the carved buffers in for this call (the caps match them, so the
  // product's own scratch check allocates nothing), with the near field in
  // line (no second stream) and no event timing; restored on every exit.
  struct Saved {
    Engine& e;
    double *xp, *t, *part, *ylev, *yn;
    float* xf;
    int max_nrhs;
    int64_t t_cap, part_cap;
    bool overlap, timing;
    explicit Saved(Engine& en)
        : e(en), xp(en.d_xp), t(en.d_t), part(en.d_part), ylev(en.d_ylev), yn(en.d_ynear),
          xf(en.d_xf), max_nrhs(en.max_nrhs), t_cap(en.t_cap), part_cap(en.part_cap),
          overlap(en.overlap_near), timing(en.timing) {}
    ~Saved() {
      e.d_xp = xp, e.d_t = t, e.d_part = part, e.d_ylev = ylev, e.d_ynear = yn, e.d_xf = xf;
      e.max_nrhs = max_nrhs, e.t_cap = t_cap, e.part_cap = part_cap;
      e.overlap_near = overlap, e.timing = timing;
    }
  } saved(e);
  e.d_xp = (double*)base;
  e.d_xf = (float*)(base + o_xf);
  e.d_t = (double*)(base + o_t);
  e.d_part = (double*)(base + o_part);
  e.d_ylev = (double*)(base + o_ylev);
  e.d_ynear = (double*)(base + o_yn);
  e.max_nrhs = 1;
  e.t_cap = T;
  e.part_cap = np;
  e.overlap_near = false;
  e.timing = false;
  clear_x_tails<<<1, 64, 0, e.stream>>>(e.d_xp, e.d_xf, N);
  H3D_LAUNCH_CHECK();
  matvec_device(e, d_b, d_y, 1, mode, false);
}

}  // namespace h3d
