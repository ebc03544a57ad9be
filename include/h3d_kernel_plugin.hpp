/*
 * h3d_kernel_plugin.hpp -- the user kernel plugin of libhodlr3d.so.
 *
 * Stands in for the reference's user kernel: a subclass of `kernel` whose
 * virtual getMatrixEntry(i, j) the user defines in their own translation unit
 * (include/kernel.hpp:19-22, examples/testHODLR3D.cpp:33-71), reading the two
 * particles particles[i], particles[j].  Here the entry is a device functor
 * compiled into the library's fused kernels (ACA, near field, direct sums,
 * entry gathers), where it is evaluated in registers from the two points:
 *
 *   struct h3d_user_kernel {
 *     static constexpr const char* name = "my kernel";
 *     // entry(i, j) for i != j of the points a = x_i and b = x_j
 *     H3D_HD static double entry(const h3d_point& a, const h3d_point& b);
 *   };
 *
 * in a header of the user's, built in with
 *
 *   make -C hodlr3d_amd/csrc KERNEL_PLUGIN=/path/to/my_kernel.hpp
 *   make -C oracle KERNEL_PLUGIN=/path/to/my_kernel.hpp   (the CPU checker)
 *
 * and selected at run time with h3d_kernel.qchoice = H3D_QCHOICE_PLUGIN.  As
 * for the catalogue, entry(i, j) = scale * h3d_user_kernel::entry(x_i, x_j)
 * for i != j and the diagonal value (h3d_kernel.diag or diag_vec) for i == j;
 * the i == j rule of the reference's dispatch (examples/testHODLR3D.cpp:34)
 * is that diagonal.  The entry need not be symmetric: transposed products
 * evaluate entry(x_j, x_i).  A plugin whose entry is finite for coincident
 * points (a == b) may be used with duplicate points; otherwise h3d_create
 * refuses them (H3D_ERR_NONFINITE).  Without KERNEL_PLUGIN the library is
 * built with h3d_plugin_plummer.hpp.
 */
#ifndef H3D_KERNEL_PLUGIN_HPP_
#define H3D_KERNEL_PLUGIN_HPP_

#if defined(__HIPCC__) || defined(__HIP__)
#define H3D_HD __host__ __device__
#else
#define H3D_HD
#endif

struct h3d_point {
  double x, y, z;
};

#endif  // H3D_KERNEL_PLUGIN_HPP_
