// Device kernel catalogue: entry(i,j) = scale * phi_q(r) for i != j
// (src/kernel.cpp:65-227, dispatch examples/testHODLR3Dsolve.cpp:32-76).
// The fp64 forms keep the reference's operation order (R = sqrt(R2) first,
// (R/b)*(R/b) rather than R2, no contraction in translation units built with
// -ffp-contract=off) so that the sqrt/division-only kernels (Qchoice 2,3,4,6,
// 7,8,16) give the same bits as a glibc build; transcendental ones use the
// device math library and may differ in the last place.
#pragma once

#include <hip/hip_runtime.h>

// The user kernel plugin (h3d_user_kernel), chosen at build time.
#ifdef H3D_KERNEL_PLUGIN
#include H3D_KERNEL_PLUGIN
#else
#include "../../include/h3d_plugin_plummer.hpp"
#endif

namespace h3d {

// Qchoice of the plugin kernel (H3D_QCHOICE_PLUGIN in include/hodlr3d.h).
constexpr int kQPlugin = 100;

static __device__ __noinline__ double phi64_any(int q, double R2) {
  double R = sqrt(R2);
  switch (q) {
    case 0: return log(1.0 + R / 1.0);
    case 1: return exp(-R / 1.0);
    case 2: return 1.0 / (1.0 + (R / 1.0) * (R / 1.0));
    case 3: return 1.0 + (R / 1.0) * (R / 1.0);
    case 4: return 1.0 / sqrt((R / 1.0) * (R / 1.0) + 1);
    case 5: return exp(-(R / 1.0) * (R / 1.0));
    case 6: return sqrt((R / 1.0) * (R / 1.0) + 1);
    case 7: return 1 / R;
    case 8: return 1.0 / (R2 * R2);
    case 9: return log(R);
    case 10: return R * R * log(R);
    case 11: return j0(1.0 * R);
    case 12: return y0(1.0 * R);
    case 13: return cos(1.0 * R) / R;
    // K_{1/2}(R) = sqrt(pi/(2R)) exp(-R), so K_{1/2}(R)/sqrt(R) = sqrt(pi/2) exp(-R)/R
    case 14: return 1.2533141373155002512 * exp(-R) / R;
    case 15: return exp(-50.0 * R) / R;
    case 16: return 1 / R;
    default: return __builtin_nan("");
  }
}

// Kernel functors.  InvR covers Qchoice 7 and 16 (1/r, the hot case) inline;
// AnyQ calls the out-of-line catalogue so the fast kernels keep few registers.
// Catalogue functors are radial (f64 of r^2); Plugin evaluates the user's
// entry from the two points (kPoints).
struct InvR {
  static constexpr bool kPoints = false;
  __device__ static inline double f64(int, double R2) { return 1 / sqrt(R2); }
};
struct AnyQ {
  static constexpr bool kPoints = false;
  __device__ static inline double f64(int q, double R2) { return phi64_any(q, R2); }
};
struct Plugin {
  static constexpr bool kPoints = true;
  __device__ static inline double pts(double xa, double ya, double za, double xb, double yb,
                                      double zb) {
    return h3d_user_kernel::entry(h3d_point{xa, ya, za}, h3d_point{xb, yb, zb});
  }
};

// Unscaled entry between two distinct points a and b (phi(r_ab) for the
// catalogue, with r^2 = (dx*dx + dy*dy) + dz*dz as the reference forms it).
template <class KF>
__device__ inline double kpair(int q, double xa, double ya, double za, double xb, double yb,
                               double zb) {
  if constexpr (KF::kPoints) {
    return KF::pts(xa, ya, za, xb, yb, zb);
  } else {
    // formed without contraction in every translation unit, so the entries
    // equal the reference's (and the ACA's) bit for bit
#pragma clang fp contract(off)
    const double dx = xa - xb, dy = ya - yb, dz = za - zb;
    return KF::f64(q, dx * dx + dy * dy + dz * dz);
  }
}

static __device__ __noinline__ float phi32_any(int q, float R2);

// fp32 form for the on-the-fly fp32 near field.
__device__ inline float phi32(int q, float R2) {
  if (q == 7 || q == 16) return rsqrtf(R2);
  return phi32_any(q, R2);
}

static __device__ __noinline__ float phi32_any(int q, float R2) {
  switch (q) {
    case 7:
    case 16: return rsqrtf(R2);
    case 8: return 1.0f / (R2 * R2);
    default: {
      float R = sqrtf(R2);
      switch (q) {
        case 0: return logf(1.0f + R);
        case 1: return expf(-R);
        case 2: return 1.0f / (1.0f + R * R);
        case 3: return 1.0f + R * R;
        case 4: return rsqrtf(R * R + 1.0f);
        case 5: return expf(-R * R);
        case 6: return sqrtf(R * R + 1.0f);
        case 9: return logf(R);
        case 10: return R * R * logf(R);
        case 11: return j0f(R);
        case 12: return y0f(R);
        case 13: return cosf(R) / R;
        case 14: return 1.2533141f * expf(-R) / R;
        case 15: return expf(-50.0f * R) / R;
        default: return __builtin_nanf("");
      }
    }
  }
}

// fp32 kernel functors for the on-the-fly fp32 near field: 1/r as the raw
// fp32 reciprocal square root (r^2 below FLT_MIN clamped), or the catalogue.
struct InvR32 {
  __device__ static inline float f(int, float R2) {
    return __builtin_amdgcn_rsqf(fmaxf(R2, 1.17549435e-38f));
  }
};
struct AnyQ32 {
  __device__ static inline float f(int q, float R2) { return phi32(q, R2); }
};

// 1/r in fp64 from the fp32 reciprocal square root refined by one Newton step,
// y1 = y0 (3 - r^2 y0^2) / 2 (relative error ~1e-7 -> ~2e-14).
struct InvRNewton {
  static constexpr bool kPoints = false;
  __device__ static inline double f64(int, double R2) {
    const double y0 = (double)__builtin_amdgcn_rsqf(fmaxf((float)R2, 1.17549435e-38f));
    return y0 * (1.5 - 0.5 * R2 * y0 * y0);
  }
};

}  // namespace h3d
