/*
 * h3d_plugin_plummer.hpp -- the default kernel plugin (see h3d_kernel_plugin.hpp):
 * the Plummer-softened Coulomb kernel 1 / sqrt(r^2 + eps^2), eps = 1/64, used
 * for gravitational N-body sums.  Not in the reference's catalogue
 * (src/kernel.cpp:65-227); it is finite at r = 0, so coincident points are
 * allowed with it.  Replace it with a header of the same form through
 * `make KERNEL_PLUGIN=...`.
 */
#ifndef H3D_PLUGIN_PLUMMER_HPP_
#define H3D_PLUGIN_PLUMMER_HPP_

#include <math.h>

#include "h3d_kernel_plugin.hpp"

struct h3d_user_kernel {
  static constexpr const char* name = "plummer 1/sqrt(r^2 + (1/64)^2)";
  H3D_HD static double entry(const h3d_point& a, const h3d_point& b) {
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double eps = 1.0 / 64.0;
    return 1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps * eps);
  }
};

#endif  // H3D_PLUGIN_PLUMMER_HPP_
