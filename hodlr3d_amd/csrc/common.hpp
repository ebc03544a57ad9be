// Shared definitions of the HIP engine: status codes, launch checks and the
// wavefront / workgroup reduction helpers used by every kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/hodlr3d.h"

namespace h3d {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define H3D_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      throw ::h3d::Error(e_ == hipErrorOutOfMemory ? H3D_ERR_OOM : H3D_ERR_HIP,           \
                         std::string(#expr) + ": " + hipGetErrorString(e_));              \
    }                                                                                     \
  } while (0)

#define H3D_LAUNCH_CHECK() H3D_HIP(hipGetLastError())

constexpr int kWave = 64;

// Reads lane `l` (wave-uniform) of a double held in VGPRs.
__device__ inline double readlane_d(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double_as_longlong(v) & 0xffffffff, l);
  int hi = __builtin_amdgcn_readlane((int)(__double_as_longlong(v) >> 32), l);
  return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}
__device__ inline float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Element type traits of stored factors.
template <class S>
struct Elem;
template <>
struct Elem<double> {
  __device__ static inline double ld(const double* p) { return *p; }
};
template <>
struct Elem<float> {
  __device__ static inline float ld(const float* p) { return *p; }
};

__device__ inline float bf16_to_f(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

// Round-to-nearest-even fp32 -> bf16 (finite inputs; NaN kept quiet).
__host__ __device__ inline uint16_t f_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

}  // namespace h3d
