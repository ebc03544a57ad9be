// Engine state shared by the host orchestration and the kernel launchers.
//
// Layout in HBM (all device arrays; "position" = index in leaf-DFS order):
//   px,py,pz  fp64 coordinates by position; lx,ly,lz fp32 offsets from the
//             point's leaf centre (on-the-fly fp32 near field)
//   perm      position -> original index; iperm its inverse
//   bstart,bcount per level (level j at offset (8^j-1)/7): each box is the
//             contiguous position segment [bstart, bstart+bcount)
//   U (per level, per target box k): column-major n_k x R_k with ld = n_k,
//             the u-vectors of the box's blocks in interaction-list order
//   V (per level, per source box s): row-major m_s x Rp_s (Rp_s = R_s padded
//             to a multiple of 4), the v-vectors of every block whose source
//             is s; tdst maps each of these columns to its slot in t
//   t         one value per (block, rank) in target order
#pragma once

#include <vector>

#include "common.hpp"

namespace h3d {

struct KernelDev {
  int q = 7;
  double scale = 1.0;
  double diag = 0.0;
  const double* diag_vec = nullptr;  // device, original order (or null)
};

// Far-field phase 1 item: t = V^T x over one column chunk of one source box.
struct P1Item {
  int64_t voff;   // element offset (within the level's V) of row 0, column 0
  int32_t level;  // 1..nLevels
  int32_t R;      // row stride (padded columns)
  int32_t c0;     // first column of this chunk
  int32_t nc;     // columns in this chunk (<= 64)
  int32_t xoff;   // position of the box's row 0
  int32_t s0, s1; // row range handled by this item
  int32_t tcol;   // tdst offset of the box's column 0
  int32_t pslot;  // -1: write t; otherwise partial-sum slot (64 values)
  int32_t pad;
};

// Partial-sum reduction for a column chunk split over several row ranges.
struct P1Reduce {
  int32_t pbase, nsplit, tcol, c0, nc, pad;
};

// Far-field phase 2 item: y_level[rows] = U t over one row chunk of a target box.
struct P2Item {
  int64_t uoff;  // element offset (within the level's U) of row 0, column 0
  int32_t ld;    // column stride (= box size)
  int32_t i0;    // first row of this chunk
  int32_t nr;    // rows in this chunk (<= 64)
  int32_t R;     // columns
  int32_t toff;  // offset of the box's t values
  int32_t yoff;  // position of the box's row 0
  int32_t level; // 1..nLevels
  int32_t pad;
};

// Near-field item: one row chunk of a leaf box against itself and its
// face/edge neighbours (src/HODLR3DTree.cpp:1708-1755).
struct NearItem {
  int32_t leaf;
  int32_t i0;
  int32_t nr;
  int32_t pad;
};

struct Level {
  int64_t nboxes = 0;
  std::vector<int32_t> iloff, ilidx;  // host copies of the interaction lists
  int32_t* d_iloff = nullptr;
  int32_t* d_ilidx = nullptr;
  int32_t* d_nbr = nullptr;           // 26 slots per box
  // ACA results, one entry per block (CSR order)
  std::vector<int32_t> rank, status;
  std::vector<int64_t> pivoff;        // offset of the block's pivots
  int32_t* d_pivI = nullptr;
  int32_t* d_pivJ = nullptr;
  // factor layout
  std::vector<int32_t> Rt;            // columns per target box
  std::vector<int32_t> Rs;            // padded columns per source box
  std::vector<int64_t> uoff, voff;    // element offsets per box (global arrays)
  std::vector<int32_t> toff;          // t offset per target box (global)
  std::vector<int32_t> tcol;          // tdst offset per source box (global)
  std::vector<int32_t> bcol_t, bcol_s;  // per block: first column in target / source concat
};

// Factors of one precision: one U and one V array per level (index 1..nL).
struct Factors {
  void* U[8] = {};
  void* V[8] = {};
  int64_t bytes = 0;
};

struct Engine {
  int64_t N = 0;
  int nL = 0;
  int leaf1d = 0;
  double L = 1.0;
  double ctr[3] = {0, 0, 0};
  int tol_pow = 7;
  double tol = 1e-7;
  int64_t ws_budget = 0;
  KernelDev kern;
  std::vector<double> radius;  // boxRadius[j]

  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;

  // points
  double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr;
  float *d_lx = nullptr, *d_ly = nullptr, *d_lz = nullptr;
  int32_t* d_perm = nullptr;
  int32_t* d_iperm = nullptr;
  double* d_diag = nullptr;
  double* d_leafc = nullptr;  // leaf centres, 3 per leaf
  std::vector<int32_t> perm;
  std::vector<int32_t> bstart, bcount;  // all levels
  int32_t* d_bstart = nullptr;
  int32_t* d_bcount = nullptr;
  std::vector<Level> lev;

  bool assembled = false;
  int64_t nU = 0, nV = 0;        // elements of U / V (all levels)
  int64_t nUlev[8] = {}, nVlev[8] = {};
  int32_t T = 0;                 // total t length
  int32_t* d_tdst = nullptr;     // per V column (all levels, padded): t index or -1
  Factors fac[3];

  // matvec plan
  std::vector<P1Item> p1;
  std::vector<P1Reduce> p1r;
  std::vector<P2Item> p2;
  std::vector<NearItem> nitems;
  int32_t npartial = 0;
  P1Item* d_p1 = nullptr;
  P1Reduce* d_p1r = nullptr;
  P2Item* d_p2 = nullptr;
  NearItem* d_near = nullptr;
  // transpose plan (phase 1 over target boxes, phase 2 over source boxes)
  std::vector<P1Item> tp1_items;
  std::vector<P1Reduce> tp1_red;
  std::vector<P2Item> tp2;
  P1Item* d_tp1x = nullptr;
  P1Reduce* d_tp1r = nullptr;
  P2Item* d_tp2 = nullptr;

  // scratch for one matvec
  int max_nrhs = 0;
  double* d_xp = nullptr;     // x by position (nrhs)
  float* d_xf = nullptr;      // fp32 copy of d_xp
  double* d_t = nullptr;      // t (nrhs), fp64 or reinterpreted as fp32
  double* d_part = nullptr;   // phase-1 partial sums
  double* d_ylev = nullptr;   // per-level y by position: nL * N (* nrhs)
  double* d_io = nullptr;     // staging for host vectors

  // stats
  int max_rank = 0;
  int64_t n_blocks = 0, rank_sum = 0, factor_elems = 0;
  double find_memory = 0;
  int64_t near_pairs = 0;
  double t_tree_ms = 0, t_assemble_ms = 0;

  ~Engine();
  void free_all();
};

inline int64_t level_offset(int j) { return ((int64_t(1) << (3 * j)) - 1) / 7; }

// tree.hip
void build_tree(Engine& e, const double* h_xyz);
// aca.hip
void assemble_all(Engine& e);
// apply.hip
void build_plan(Engine& e);
void ensure_scratch(Engine& e, int nrhs);
void prepare_precision(Engine& e, int prec);
void matvec_device(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool transpose);
// krylov.hip
int gmres_reference(Engine& e, const double* d_b, double* d_x, int maxit, double thr,
                    double* err, int* its, double* hist);
int solve_ir(Engine& e, const double* d_b, double* d_x, int restart, int max_inner,
             double inner_tol, int inner_prec, double tol, int max_outer, bool transpose,
             h3d_solve_info* info);

template <class T>
T* dalloc(int64_t n) {
  T* p = nullptr;
  if (n <= 0) return nullptr;
  H3D_HIP(hipMalloc((void**)&p, sizeof(T) * (size_t)n));
  return p;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace h3d
