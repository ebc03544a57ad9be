// Device helpers shared by the factor-streaming kernels: one work item per
// wave, global (non-flat) non-temporal loads of factor streams.
#pragma once
#include <cstdint>

#include "common.hpp"

namespace h3d {
namespace {

constexpr int kWavesPerBlock = 4;

inline unsigned wave_grid(int64_t items) {
  return (unsigned)((items + kWavesPerBlock - 1) / kWavesPerBlock);
}

// Global index of the calling wave (one work item per wave), made visibly
// wave-uniform so that the compiler loads the item and its tables with scalar
// loads.
__device__ inline int wave_id() {
  return __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
}

// Factor pointers come from per-chunk tables, so the compiler cannot prove
// they are global memory; loading through an address-space-1 pointer makes
// them global (not flat) loads, which do not share the LDS wait counter.
template <class T>
__device__ inline T ntload(const void* p) {
  return __builtin_nontemporal_load((const __attribute__((address_space(1))) T*)p);
}

// A persistent kernel's queue: an atomic counter of items taken.  A wave
// takes GRAB consecutive items per add (one add by lane 0: with many small
// items, one counter hit once per item by every wave of the device would
// serialize), and asks for its next group before working on the current one,
// so the add's latency is hidden.  queue_take returns lane 0's raw result;
// queue_first makes it wave-uniform where it is needed.
template <int GRAB>
__device__ inline unsigned queue_take(unsigned* queue) {
  unsigned v = 0;
  if ((threadIdx.x & 63) == 0) v = atomicAdd(queue, (unsigned)GRAB);
  return v;
}
__device__ inline int queue_first(unsigned v) { return (int)__builtin_amdgcn_readfirstlane(v); }

// Runs body(w) for every item w < nitems that the calling wave takes from the
// queue; every wave returns once the counter has passed nitems.
template <int GRAB, class F>
__device__ inline void queue_loop(unsigned* queue, int nitems, F&& body) {
  int w0 = queue_first(queue_take<GRAB>(queue));
  while (w0 < nitems) {
    const unsigned next = queue_take<GRAB>(queue);
    const int w1 = min(w0 + GRAB, nitems);
    for (int w = w0; w < w1; ++w) body(w);
    w0 = queue_first(next);
  }
}

// Clears n queue counters (a kernel, not a memset node, so that it may be
// captured into a graph).
__global__ void queue_reset_kernel(unsigned* __restrict__ q, int n) {
  if ((int)threadIdx.x < n) q[threadIdx.x] = 0u;
}

// Zero fill (used instead of a memset node, so that a matvec may be captured
// into a HIP graph).
__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// 8- and 16-byte vectors that need only 4-byte alignment
typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u32x2a __attribute__((ext_vector_type(2), aligned(4)));

// A stored 8-bit float: OCP e4m3 (1 sign, 4 exponent bits of bias 7, 3
// mantissa bits; largest finite 448), gfx950's fp8 conversion format.
struct f8e4m3 {
  uint8_t bits;
};

// A stored int8 entry of a column scaled by its largest entry (the scale is
// carried by the column's factor).
struct i8s {
  int8_t v;
};

// Int8 entry q of a 32-bit word (byte q) as a float: the byte flipped to
// offset binary (v + 128, an unsigned byte), converted, 128 subtracted (all
// exact).  A sign-extended byte converted as a signed integer was compiled for
// gfx950 into an unsigned conversion of the sign-extended byte (wrong for
// v < 0), so the conversion stays unsigned.
__device__ inline float i8_to_f(uint32_t w, int q) {
  return (float)(((w ^ 0x80808080u) >> (8 * q)) & 0xffu) - 128.0f;
}

// Four int8 entries of a 32-bit word (byte 0 first) as floats.
__device__ inline void i8x4_to_f(uint32_t w, float (&o)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = i8_to_f(w, q);
}

// A stored 24-bit float: the upper three bytes of an fp32 (sign, 8 exponent
// bits, 15 mantissa bits), rounded to nearest; entries are packed, 3 bytes each.
struct f24 {
  uint8_t b[3];
};
static_assert(sizeof(f24) == 3, "packed 3-byte entries");

// Four packed fp24 values (12 bytes) as fp32 bit patterns.
__device__ inline void f24x4_bits(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t (&o)[4]) {
  o[0] = w0 << 8;
  o[1] = ((w0 >> 24) << 8) | (w1 << 16);
  o[2] = ((w1 >> 16) << 8) | (w2 << 24);
  o[3] = w2 & 0xffffff00u;
}

// Two fp8 values of a 32-bit word (HI: bytes 2-3) as floats.
template <bool HI>
__device__ inline void f8x2_to_f(uint32_t w, float& a, float& b) {
  const auto v = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, HI);
  a = v[0];
  b = v[1];
}

template <class S>
struct Load;
template <>
struct Load<double> {
  __device__ static inline double f(const double* p) { return ntload<double>(p); }
};
template <>
struct Load<float> {
  __device__ static inline float f(const float* p) { return ntload<float>(p); }
};
template <>
struct Load<uint16_t> {
  __device__ static inline float f(const uint16_t* p) { return bf16_to_f(ntload<uint16_t>(p)); }
};
template <>
struct Load<f24> {
  __device__ static inline float f(const f24* p) {
    const uint8_t* b = (const uint8_t*)p;
    const uint32_t u = ((uint32_t)ntload<uint8_t>(b) << 8) | ((uint32_t)ntload<uint8_t>(b + 1) << 16) |
                       ((uint32_t)ntload<uint8_t>(b + 2) << 24);
    return __uint_as_float(u);
  }
};
template <>
struct Load<f8e4m3> {
  __device__ static inline float f(const f8e4m3* p) {
    return __builtin_amdgcn_cvt_f32_fp8((int)ntload<uint8_t>(p), 0);
  }
};
template <>
struct Load<i8s> {
  __device__ static inline float f(const i8s* p) { return i8_to_f(ntload<uint8_t>(p), 0); }
};
template <>
struct Load<_Float16> {
  __device__ static inline float f(const _Float16* p) {
    return (float)__builtin_bit_cast(_Float16, ntload<uint16_t>(p));
  }
};

}  // namespace
}  // namespace h3d
