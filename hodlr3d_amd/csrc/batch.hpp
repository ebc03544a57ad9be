// The instance batch (batch.hip): merged work tables of many engines and
// their scratch.
#pragma once

#include <vector>

#include "engine.hpp"

namespace h3d {

struct Batch {
  std::vector<Engine*> es;
  int prec = 1;
  bool acc64 = false;
  int64_t N = 0;                 // points of all instances
  int nL = 0;                    // most levels of any instance
  std::vector<int64_t> off;      // instance i: [off[i], off[i + 1]) of the packed vectors
  int64_t T = 0, npart = 0, nchunks = 0, factor_bytes = 0;
  KernelDev kd;
  int32_t* d_perm = nullptr;     // position -> packed original index
  int32_t* d_iperm = nullptr;
  double* d_alpha = nullptr;     // per point (packed original order): the instance's alpha
  double* d_diag = nullptr;      // per point: its diagonal entry
  BlockDev* d_blk = nullptr;
  void** d_ubase = nullptr;
  void** d_vbase = nullptr;
  P1Item* d_p1 = nullptr;
  P1Reduce* d_p1r = nullptr;
  P2Item* d_p2 = nullptr;
  int n1 = 0, n1r = 0, n2 = 0;
  NearItem* d_near = nullptr;
  int nn = 0;
  int32_t* d_nbr = nullptr;      // 26 per leaf (merged leaf ids)
  int32_t* d_lstart = nullptr;   // per leaf: first position
  int32_t* d_lcount = nullptr;
  double *d_px = nullptr, *d_py = nullptr, *d_pz = nullptr;
  float *d_lx = nullptr, *d_ly = nullptr, *d_lz = nullptr;
  double* d_leafc = nullptr;
  // scratch of one product
  double* d_xp = nullptr;
  float* d_xf = nullptr;
  double* d_t = nullptr;
  double* d_part = nullptr;
  double* d_ylev = nullptr;
  double* d_ynear = nullptr;
  hipStream_t own = nullptr, stream = nullptr, side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // load width in bytes per lane, | H3D_PERSISTENT for the persistent launch
  int tune_p1 = 16, tune_p2 = 16;
  unsigned* d_queue = nullptr;     // the persistent launches' item counters (phase 1, phase 2)

  ~Batch();
  void free_all();
};

// Builds the merged tables of `count` engines for products in `mode` (an ACA
// copy: fp64, fp32 or bf16, with H3D_ACC_FP64 optional); makes each engine's
// copy and scratch first.
void build_batch(Batch& bt, Engine* const* es, int count, int mode);
// One product of every instance: packed device vectors b -> y.  With `side`
// the near field runs there (fork / join events), else in line on st.
void batch_product(Batch& bt, const double* b, double* y, hipStream_t st, hipStream_t side,
                   hipEvent_t fork, hipEvent_t join);
void batch_matvec(Batch& bt, const double* b, double* y);
// Times every load width of both phases over the merged tables, then the
// persistent launch at the fastest width, and keeps the fastest; ms[10]: as
// h3d_autotune's, then phase 1 and phase 2 persistent.
void batch_autotune(Batch& bt, int reps, double* ms);

}  // namespace h3d
