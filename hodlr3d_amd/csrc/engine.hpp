// Engine state shared by the host orchestration and the kernel launchers.
//
// Layout in HBM (all device arrays; "position" = index in leaf-DFS order):
//   px,py,pz  fp64 coordinates by position; lx,ly,lz fp32 offsets from the
//             point's leaf centre (on-the-fly fp32 near field)
//   perm      position -> original index; iperm its inverse
//   bstart,bcount per level (level j at offset (8^j-1)/7): each box is the
//             contiguous position segment [bstart, bstart+bcount)
//   factors   stored in chunks; one chunk holds every block of a run of
//             consecutive target boxes of one level:
//               U region: per target box k, column-major n_k x R_k with
//                         ld = aca_ldu(n_k) (n_k rounded up to even, or to a
//                         multiple of 32 above 256 rows), the u-vectors of its
//                         blocks in interaction-list order
//               V region: per block, column-major m_b x r_b with ldv = m_b
//                         rounded up to even (the v-vectors)
//             Padding is zero and every chunk ends with 64 zero elements, so
//             16-byte vector loads of any column stay inside the allocation.
//             Each stored precision has its own copy of every chunk with the
//             same element offsets.
//   t         one value per (block, rank) in target order: box k's blocks
//             occupy t[toff_k, toff_k + R_k)
#pragma once

#include <algorithm>
#include <atomic>
#include <vector>

#include "common.hpp"

namespace h3d {

struct KernelDev {
  int q = 7;
  double scale = 1.0;
  double diag = 0.0;
  const double* diag_vec = nullptr;  // device, original order (or null)
};

// Per-block data read by the phase-1 lanes (global block id = level base + CSR index).
struct BlockDev {
  int64_t voff;   // element offset of V_b in its chunk's V region
  int32_t col;    // first column of the block in its target box's concatenation
  int32_t r;      // rank
  int32_t m;      // rows of V_b (source box size)
  int32_t src;    // position of the source box's first point
  int32_t chunk;  // chunk holding the block
  int32_t pad;
};

// Far-field phase 1 item: t = V_b^T x for the consecutive blocks [b0, b1) of
// one target box, over rows [s0, s1) of each block (at most 512 rows).
struct P1Item {
  int32_t chunk;
  int32_t b0, b1;  // global block ids
  int32_t s0, s1;  // row range
  int32_t toff;    // t offset of the target box
  int32_t pslot;   // -1: write t; otherwise element offset of r partial sums
  int32_t pad;
};

// Phase-1 item of the multi-RHS path: rows [s0, s0 + nrows) (at most 256) of
// one source box against the columns [col0, col0 + ncols) of the column table
// (at most 256, all with that source box).  The table lists the V columns of
// every block of rank >= 1, grouped by source box (by-source block order).
struct GTItem {
  int32_t col0, ncols;
  int32_t a0;     // position of the range's first row
  int32_t nrows;
  int32_t s0;     // first row of the range within the source box
  int32_t range;  // s0 / 256: selects the partial-sum slot
  int32_t pad0, pad1;
};

// Partial-sum reduction: nsplit row ranges x nc columns at part[pbase + q*nc + l]
// summed in range order into t[toff + l].
struct P1Reduce {
  int32_t pbase, nsplit, toff, nc;
};

// Far-field phase 2 item: y_level[rows] = U t over one row chunk of a target box.
struct P2Item {
  int64_t uoff;   // element offset of the box's U in its chunk's U region
  int32_t chunk;
  int32_t ld;     // column stride (= box size)
  int32_t i0;     // first row of this chunk
  int32_t nr;     // rows in this chunk (<= the plan's rows per item)
  int32_t R;      // columns
  int32_t toff;   // offset of the box's t values
  int32_t yoff;   // position of the box's row 0
  int32_t level;  // 1..nLevels
};

// Transpose phase 2 item: rows [i0, i0+nr) of one source box gather
// V_b t_b from every block whose source is that box.
struct TP2Item {
  int32_t sfirst, scount;  // range in the by-source block list
  int32_t yoff;            // position of the source box's row 0
  int32_t i0, nr;
  int32_t level;
};

// Near-field item: one row chunk of a leaf box against itself and its
// face/edge neighbours (src/HODLR3DTree.cpp:1708-1755).
struct NearItem {
  int32_t leaf;
  int32_t i0;
  int32_t nr;
  int32_t pad;
};

// Phase-1 item of a recompressed class: rows [s0, s1) of columns [c0, c1) of
// a source box's V_s (the V columns of every block with that source,
// concatenated in by-source order, column-major with ld = box size rounded
// up to even).
struct MixP1Item {
  int64_t voff;    // element offset of V_s column 0 in its V chunk
  int32_t vchunk;
  int32_t ld;
  int32_t s0, s1;  // row range
  int32_t c0, c1;  // column range
  int32_t tsb;     // source-order t index of V_s column 0
  int32_t xoff;    // position of the source box's row 0
  int32_t pslot;   // -1: write t; otherwise element offset of (c1 - c0) partial sums
  int32_t pad;
};

// Row-split partial sums of a mixed phase-1 column range: nsplit x nc values at
// part[pbase + q*nc + l] summed in range order into t_src[tsb + l].
struct MixReduce {
  int32_t pbase, nsplit, nc, tsb;
};

// Per-level factor bookkeeping of one copy, indexed by the level's blocks (CSR order)
// and target boxes.
struct LevelFactors {
  std::vector<int32_t> rank, col;  // per block
  std::vector<int64_t> voff;       // per block
  std::vector<int32_t> Rt, toff;   // per target box
  std::vector<int64_t> uoff;       // per target box
};

// One level of the tree: structure plus the ACA copy's factor bookkeeping
// (rank: block rank; col: first column of the block in its target box; voff:
// V offset in the chunk; Rt: columns of a target box; toff: its global t
// offset; uoff: its U offset in the chunk).
struct Level : LevelFactors {
  int64_t nboxes = 0;
  int64_t bbase = 0;                  // global id of the level's first block
  std::vector<int32_t> iloff, ilidx;  // host copies of the interaction lists
  int32_t* d_iloff = nullptr;
  int32_t* d_ilidx = nullptr;
  int32_t* d_nbr = nullptr;           // 26 slots per box
  std::vector<int32_t> status;        // per block (CSR order)
  std::vector<int64_t> pivoff;        // per block: pivot offset in its chunk
  std::vector<int32_t> chunk;         // per target box: chunk id
};

// One chunk of factors: per precision a U region followed by a V region.
struct Chunk {
  void* data[3] = {nullptr, nullptr, nullptr};
  int64_t nU = 0, nV = 0;  // elements
  int32_t* piv = nullptr;  // pivots: for each block I then J (r each)
  int64_t npiv = 0;
};

// Work lists of one factor copy's matvec (built on the host, mirrored on the
// device): phase-1 items and their partial-sum reductions, phase-2 items, and
// the transpose's items.
struct Plan {
  std::vector<P1Item> p1;
  std::vector<P1Reduce> p1r;
  std::vector<P2Item> p2;
  // transpose phase 1 over the U columns of each target box, in the form of
  // the mixed phase 1 (column runs of a box's U, rows in ranges)
  std::vector<MixP1Item> tp1;
  std::vector<MixReduce> tp1r;
  std::vector<TP2Item> tp2;
  int32_t npartial = 0;
  P1Item* d_p1 = nullptr;
  P1Reduce* d_p1r = nullptr;
  P2Item* d_p2 = nullptr;
  MixP1Item* d_tp1 = nullptr;
  MixReduce* d_tp1r = nullptr;
  TP2Item* d_tp2 = nullptr;
  void free_device();
};

// One storage class of the recompressed copy (fp32, fp16, fp8 or fp24 columns).  Columns
// are stored scaled: the class's U and V columns hold u_k * su / nu_k and
// v_k * sv / nv_k (su, sv powers of two near sqrt(n), sqrt(m)), and the
// factor c_k = nu_k nv_k / (su sv) multiplies t_k after phase 1.  Chunk ids
// and the block/box structure are those of the ACA copy.
struct FactorClass {
  int elem = 4;                      // bytes per stored element (4: fp32, 3: fp24, 2: fp16, 1: fp8)
  std::vector<LevelFactors> lf;      // [level]; voff is relative to the block's V chunk
  std::vector<void*> udata, vdata;   // U chunks (the ACA copy's chunk ids), V chunks
  void* arena = nullptr;             // one allocation holding every chunk of the class
  std::vector<int64_t> nU, nV;       // elements per U chunk / V chunk
  void** d_ubase = nullptr;          // per U chunk
  void** d_vbase = nullptr;          // per V chunk
  BlockDev* d_blk = nullptr;         // chunk = V chunk id
  int32_t* d_btoff = nullptr;
  Plan plan;                         // phase 2 and transpose items (phase 1 below)
  // Phase 1 runs by source box and writes t in source order (t_src, every
  // store contiguous); tmap[i] is the t_src index of target-order t entry i,
  // gathered (and scaled) into target order before phase 2.
  std::vector<MixP1Item> mp1;
  std::vector<MixReduce> mp1r;
  std::vector<int32_t> tmap;
  MixP1Item* d_mp1 = nullptr;
  MixReduce* d_mp1r = nullptr;
  // transpose phase 2 over each source box's V_s (t in source order)
  std::vector<P2Item> tp2v;
  P2Item* d_tp2v = nullptr;
  int32_t* d_tmap = nullptr;
  float* d_scale32 = nullptr;        // classes 1, 2: c_k in target order as fp32 (their sums' type)
  int32_t mpartial = 0;              // partial sums of the phase-1 items
  int32_t T = 0;                     // t entries of this class
  int32_t tbase = 0;                 // offset of this class in the shared t / scale arrays
  int64_t u_elems = 0, v_elems = 0, rank_sum = 0, bytes = 0;
  void free_all();
};

// Rows of a mixed phase-1 item for entries of `elem` bytes: the kernel's reach
// with 16-byte loads, 64 lanes x 4 vectors (x 2 for 1-byte entries, whose
// vectors of 16 rows would otherwise hold 64 x values per lane in registers).
inline int mix_item_rows(int elem) {
  return elem == 1 ? 64 * 2 * 16 : (elem == 3 ? 64 * 4 * 4 : 64 * 4 * 16 / elem);
}

// Storage formats of the recompressed copy besides fp32 (h3d_set_mixed_formats).
// kMixInt8: the 8-bit tier stores int8 with a scale per column side (its
// largest entry) instead of e4m3
constexpr int kMixFp16 = 1, kMixFp8 = 2, kMixFp24 = 4, kMixInt8 = 8;

// Recompressed mixed-precision copy: SVD-truncated factors of every block with
// columns in fp32, fp24, fp16 or fp8 chosen against a global Frobenius error
// budget.  Classes 0 and 3 (fp32, fp24) sum in fp64, classes 1 and 2 (fp16,
// fp8) in fp32.
constexpr int kMixClasses = 4;
inline bool class_acc64(int c) { return c == 0 || c == 3; }
struct MixedCopy {
  bool ready = false;
  double tol = 0;                    // requested relative Frobenius error
  int formats = kMixFp16 | kMixFp8 | kMixFp24 | kMixInt8;  // formats the next build may use
  int built_formats = 0;             // those the copy in place could use
  bool int8 = false;                 // class 2 holds int8 columns (built with kMixInt8)
  // 0: fp32 columns, 1: fp16, 2: fp8 (e4m3, or int8), 3: fp24 (fp32 without its last 8 bits)
  FactorClass cls[kMixClasses];
  double* d_scale = nullptr;         // c_k per t entry (class 0, 1, 2, then 3)
  // statistics of the last build
  double fro2_far = 0, fro2_near = 0, err2_drop = 0, err2_fp16 = 0, err2_fp32 = 0, err2_fp8 = 0,
         err2_fp24 = 0, err2_fp8_model = 0;  // int8: err2_fp8 realized, the model's plan
  double theta_drop = 0, ms = 0;
  bool single_jacobi = false;        // the build kept J between its passes (one Jacobi per block)
  int64_t aca_rank_sum = 0, dropped = 0, raw_blocks = 0, raw_rank = 0;
  void free_all();
};

// Scratch t layout of the mixed copy's product (bytes from the t buffer,
// every part 16-byte aligned for vector stores): each class's target-order t
// (classes 0 and 3 in doubles, 1 and 2 in floats), then each class's
// source-order t (phase-1 output).
struct MixT {
  int64_t t[kMixClasses], src[kMixClasses], end;
};
inline MixT mix_t_layout(const MixedCopy& m) {
  MixT L{};
  int64_t o = 0;
  auto put = [&](int64_t bytes) {
    const int64_t at = o;
    o = (o + bytes + 15) / 16 * 16;
    return at;
  };
  for (int c = 0; c < kMixClasses; ++c) L.t[c] = put((int64_t)m.cls[c].T * (class_acc64(c) ? 8 : 4));
  for (int c = 0; c < kMixClasses; ++c) L.src[c] = put((int64_t)m.cls[c].T * (class_acc64(c) ? 8 : 4));
  L.end = o;
  return L;
}
// doubles of t scratch the mixed copy's product needs
inline int64_t mix_t_doubles(const MixedCopy& m) { return mix_t_layout(m).end / 8 + 2; }

// Work tables of the grouped products (several right-hand sides per factor
// read, multi.hip) in one direction.  Phase 1 takes a column table: forward,
// the V columns of every block grouped by source box; transposed, the U
// columns of every target box.  Phase 2: forward, items over the rows of
// each target box (its U is one matrix); transposed, items over the rows of
// each source box with a column table (the V columns of the blocks with that
// source lie in different chunks) and the t index of each column.
struct GroupTables {
  bool built = false;
  void* d_rcols = nullptr;          // per column of a row-split source: int4 reduce entry
  int64_t n_rcols = 0;
  void* d_gitems = nullptr;         // phase-1 items (GTItem), by class of 64, 128, 256 rows
  int64_t n_gitems = 0;
  int64_t n_p1[3] = {0, 0, 0};      // items of each class
  void* d_gp2 = nullptr;            // phase-2 items (P2Item)
  int64_t n_gp2 = 0;
  int64_t* d_gcolv = nullptr;       // phase-1 column table: element offset | chunk << 48
  int2* d_gcolo = nullptr;          // phase-1 column table: {output slot of range 0, step per range}
  int64_t g_np = 0;                 // partial entries per RHS
  int64_t* d_p2col = nullptr;       // transposed phase 2: per column, element offset | chunk << 48
  int32_t* d_p2t = nullptr;         // transposed phase 2: per column, its t index
  void free_all();  // multi.hip
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
  int store_mask = 1;  // precisions written at assembly (bit p = precision p)
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
  bool factors_lost = false;  // a move ran out of memory part way (no copy is whole)
  int64_t nblk_all = 0;
  int32_t T = 0;                 // total t length
  std::vector<Chunk> chunks;
  void** d_ubase[3] = {nullptr, nullptr, nullptr};  // per precision: U base of every chunk
  void** d_vbase[3] = {nullptr, nullptr, nullptr};  // per precision: V base of every chunk
  int64_t fac_bytes[3] = {0, 0, 0};
  BlockDev* d_blk = nullptr;     // all blocks
  std::vector<int32_t> by_src;   // global block ids grouped by source box
  int32_t* d_by_src = nullptr;
  int32_t* d_btoff = nullptr;    // per block: global t index of its first column

  // matvec plan of the ACA copy; near-field items are shared by every copy
  Plan plan;
  // phase-2 items of the fp32 and bf16 copies (256 / 512 rows, so that a
  // wave instruction reads 1 KB of a column with 16-byte loads, as the plan's
  // 128-row items do for fp64)
  std::vector<P2Item> p2_narrow[2];
  P2Item* d_p2_narrow[2] = {nullptr, nullptr};
  std::vector<NearItem> nitems;
  NearItem* d_near = nullptr;
  MixedCopy mix;

  // scratch for one matvec
  int max_nrhs = 0;
  double* d_xp = nullptr;     // x by position
  float* d_xf = nullptr;      // fp32 copy of d_xp
  double* d_t = nullptr;      // t, fp64 or reinterpreted as fp32
  double* d_part = nullptr;   // phase-1 partial sums
  double* d_ylev = nullptr;   // per-level y by position: nL * N
  double* d_ynear = nullptr;  // near field by position: N
  hipStream_t side = nullptr;                              // the near field's stream
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;         // its fork / join events
  hipStream_t side2 = nullptr;      // the mixed copy's fp64-sum classes (forked at ev_fork)
  hipEvent_t ev_join2 = nullptr;    // its join event
  hipEvent_t ev_switch = nullptr;  // orders a new stream after the old (h3d_set_stream)
  // products fork work to `side` (near field) and `side2` (half of the mixed
  // copy's classes); off while a graph is captured and in h3d_matvec_noalloc
  bool overlap_near = true;
  double* d_io = nullptr;     // staging for host vectors
  // allocated entries of d_t / d_part per right-hand side (d_part holds
  // max(max_nrhs, 2) regions of part_cap: the second is side2's)
  int64_t t_cap = 0, part_cap = 0;
  // products with several right-hand sides: groups of up to rhs_block (16 or
  // 32; 0 = one at a time) applied together (multi.hip); scratch for W per entry
  int rhs_block = 32;
  int g_cols = 2048;                // columns of one short-source phase-1 item (1024 or 2048)
  int w_cap = 0;                    // W of the allocated group scratch
  int64_t w_t = 0, w_np = 0;        // t and partial entries it holds per RHS
  double* d_xw = nullptr;           // x by position, W per point
  double* d_tw = nullptr;           // t (T + 1 slots) then the phase-1 partial sums, W per slot
  double* d_ylevw = nullptr;        // per-level y by position, W per point
  double* d_ynw = nullptr;          // near field by position, W per point
  double* d_iow = nullptr;          // host staging of one group (2 N W)
  int io_w = 0;
  GroupTables grp[2];               // grouped products: forward, transposed
  GroupTables grpm[kMixClasses];    // grouped products on the mixed copy's classes (forward)
  int tw_w = 0;                     // W of the allocated d_tw
  float* d_xw32 = nullptr;          // x by position in fp32, W per point (the fp16 class)
  int xw32_w = 0;
  void* d_mtw[kMixClasses] = {};    // the mixed classes' t and partial sums, W per slot
  int64_t mtw_cap[kMixClasses] = {};  // their allocated bytes

  // apply-time operator: y = alpha * (off-diagonal part) + diag * b
  double alpha = 1.0;
  // autotuned load width in bytes per lane for phase 1 / phase 2, indexed by
  // storage precision + 3 * (fp64 accumulation over fp32/bf16); 6 and 7 are
  // the fp32 and fp16 classes of the mixed copy
  int tune_p1[8] = {16, 16, 16, 16, 16, 16, 16, 16};
  int tune_p2[8] = {16, 16, 16, 16, 16, 16, 16, 16};
  // rows per lane of the mixed phase-1 kernel, in load vectors (1..4), slots 6, 7
  int tune_k[8] = {4, 4, 4, 4, 4, 4, 4, 4};
  // per-kernel event timing of matvecs
  bool timing = false;
  hipEvent_t ev[16] = {};
  double ktime[15] = {};
  int kcount = 0;
  // reference-semantics timers (never reset), fed by the per-kernel timing
  double ref_regen_ms = 0, ref_blas2_ms = 0;

  // stats
  int max_rank = 0;
  int64_t n_blocks = 0, rank_sum = 0, factor_elems = 0, u_elems = 0, v_elems = 0;
  double find_memory = 0;
  int64_t near_pairs = 0;
  double t_tree_ms = 0, t_assemble_ms = 0;

  ~Engine();
  void free_all();
};

inline int64_t level_offset(int j) { return ((int64_t(1) << (3 * j)) - 1) / 7; }

// Column stride of the ACA copies' U for a target box of n rows (element
// offsets are shared by the fp64, fp32 and bf16 copies).  Columns of more than
// 256 rows span several phase-2 items of every copy (1 KB of rows each), so
// they start on multiples of 32 elements (256 / 128 / 64 bytes); shorter
// columns are packed (n rounded up to even).
__host__ __device__ inline int aca_ldu(int n) { return n > 256 ? (n + 31) / 32 * 32 : n + (n & 1); }
// Element offset (in its chunk) of the U of such a box placed at `off` or after.
inline int64_t aca_ustart(int64_t off, int n) { return n > 256 ? (off + 31) / 32 * 32 : off; }

// Column stride of a recompressed class's U for a target box of n rows (elem
// bytes per entry).  Columns of more than 1 KB span several phase-2 items
// (1 KB of rows each), so they start on 128-byte lines: an item's piece of a
// column then touches 8 lines, not 9.  Shorter columns are packed (n rounded
// up to even): one item reads the box's columns as one region.
inline int class_ldu(int n, int elem) {
  // 3-byte entries: 128 entries (3 lines) per step for long columns
  const int line = elem == 3 ? 128 : 128 / elem;
  return n * elem > 1024 ? (n + line - 1) / line * line
                         : (elem == 1 || elem == 3 ? (n + 3) & ~3 : n + (n & 1));
}
// Column stride of V for a source box of m rows (the ACA copies and the
// recompressed classes; elem bytes per entry): even, a multiple of 4 for 1- and
// 3-byte entries, so that every column starts 4-byte aligned for vector loads.
__host__ __device__ inline int class_ldv(int m, int elem) {
  return elem == 1 || elem == 3 ? (m + 3) & ~3 : m + (m & 1);
}
// Element offset (in its chunk) of the U of such a box placed at `off` or after.
inline int64_t class_ustart(int64_t off, int n, int elem) {
  const int line = elem == 3 ? 128 : 128 / elem;
  return n * elem > 1024 ? (off + line - 1) / line * line : off;
}
constexpr int64_t kChunkTail = 64;  // zero elements after every factor chunk
inline int elem_size(int prec) { return prec == 0 ? 8 : (prec == 1 ? 4 : 2); }
// Precision code of the recompressed mixed copy (fp32 + fp16 columns, fp64 accumulation).
constexpr int kPrecMixed = 3;

// The near field's inputs: items over the leaves, each leaf's 26 neighbour
// slots (-1: none) and position segment, the points (fp64 by position, fp32
// offsets from the leaf centre) and x by position.
struct NearTables {
  const NearItem* items = nullptr;
  int n = 0;
  const int32_t* nbr = nullptr;
  const int32_t* bstart = nullptr;
  const int32_t* bcount = nullptr;
  const double *px = nullptr, *py = nullptr, *pz = nullptr;
  const float *lx = nullptr, *ly = nullptr, *lz = nullptr;
  const double* leafc = nullptr;
  const int32_t* perm = nullptr;
  KernelDev kd;
  const double* xp = nullptr;
  const float* xf = nullptr;
  double* ynear = nullptr;
};
NearTables near_tables(Engine& e);

// Launchers of the ACA-copy product over raw tables (apply.hip), shared by
// the single product and the instance batch (batch.hip).  With fp64
// accumulation x, t and the partial sums are fp64 (xp); otherwise x is xf and
// t / part hold floats.
void far_gather(const double* b, const int32_t* perm, int64_t N, double* xp, float* xf,
                hipStream_t st);
void far_zero(double* p, int64_t n, hipStream_t st);
// queue: persistent launch (a resident grid taking items from this counter;
// null: one wave per item)
void far_phase1(int prec, bool acc64, int bytes, const P1Item* items, int n, void* const* vbase,
                const BlockDev* blk, const double* xp, const float* xf, double* t, double* part,
                hipStream_t st, unsigned* queue = nullptr);
void far_reduce(bool acc64, const P1Reduce* rs, int n, const double* part, double* t,
                hipStream_t st);
void far_phase2(int prec, bool acc64, int bytes, const P2Item* items, int n, void* const* ubase,
                const double* t, double* ylev, int64_t N, hipStream_t st,
                unsigned* queue = nullptr);
void far_near(const NearTables& nt, int prec, bool acc64, bool transpose, hipStream_t st);

// tree.hip
void build_tree(Engine& e, const double* h_xyz);
// aca.hip
void assemble_all(Engine& e);
// Derived tables and statistics from the per-level factor bookkeeping and chunks.
void finish_assembly(Engine& e);
// store.hip: the saved assembled representation (load fills a fresh engine)
void save_engine(Engine& e, const char* path);
void load_engine(Engine& e, const char* path, int64_t ws_budget);
// multi.hip: group width for nrhs right-hand sides (1: one at a time), and
// the grouped product (forward, ACA factor copies)
int block_width(const Engine& e, int nrhs, int prec, bool transpose, bool acc64);
void matvec_block(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool acc64,
                  bool transpose);
// Frees the grouped tables and scratch of the mixed copy (it is rebuilt or released).
void release_mixed_groups(Engine& e);
// times the grouped product of nrhs columns for every group width and
// phase-1 item size, keeps the fastest; ms[8] (width-major), best[2]
void autotune_rhs(Engine& e, int mode, int nrhs, int reps, double* ms, int* best);
// apply.hip
void build_plan(Engine& e);
void ensure_scratch(Engine& e, int nrhs);
void prepare_precision(Engine& e, int prec, bool move = false);
void release_precision(Engine& e, int prec);
void upload_chunk_table(Engine& e);
void matvec_device(Engine& e, const double* d_b, double* d_y, int nrhs, int prec, bool transpose);
// One product with every scratch buffer carved from a static per-GPU array
// (no allocation); the precision's copy must be resident.
void matvec_noalloc(Engine& e, const double* d_b, double* d_y, int mode);
int64_t noalloc_max_bytes();
// Builds whatever a matvec in `mode` needs (factor copy, scratch) so that the
// matvec itself allocates nothing and can be captured into a graph.
void prepare_mode(Engine& e, int mode);
void direct_rows(Engine& e, const double* d_b, const int64_t* h_rows, int64_t nrows, double* d_out);
// Kernel entries K(rows[r], cols[c]) (original indices, host arrays) into the
// column-major nrows x ncols d_out (kernel::getMatrix, src/kernel.cpp:37-50).
void get_entries(Engine& e, const int64_t* h_rows, int64_t nrows, const int64_t* h_cols,
                 int64_t ncols, double* d_out);
// Times every load width of phase 1 and phase 2 for one precision mode on the
// assembled problem and keeps the fastest; ms[8] = phase 1 at 16/8/4/2 bytes,
// then phase 2 at 16/8/4/2 bytes (negative when the width does not apply).
void autotune(Engine& e, int mode, int reps, double* ms);
// compress.hip: builds e.mix with relative Frobenius error budget `tol`
void build_mixed(Engine& e, double tol);
void release_mixed(Engine& e);
// Plan of one copy from its per-level factor bookkeeping (apply.hip).
// lf[j] describes level j (1..nLevels).
// p2rows: rows of one phase-2 item (128 for the ACA copies, whose plan serves
// every precision; 256 / 512 for the fp32 / fp16 classes of the mixed copy);
// class_elem: 0 for the ACA layout (U stride aca_ldu), else the
// bytes per entry of a mixed class (U stride class_ldu)
void build_plan_for(Engine& e, const std::vector<const LevelFactors*>& lf, Plan& out, int p2rows,
                    int class_elem = 0);
// The same in two steps: the work lists on the host (reads e only; safe on a
// host thread while the stream works), then their device copies.
void build_plan_host(const Engine& e, const std::vector<const LevelFactors*>& lf, Plan& out,
                     int p2rows, int class_elem);
void upload_plan(Engine& e, Plan& out);
// krylov.hip
int gmres_reference(Engine& e, const double* d_b, double* d_x, int maxit, double thr,
                    double* err, int* its, double* hist);
int solve_ir(Engine& e, const double* d_b, double* d_x, int restart, int max_inner,
             double inner_tol, int inner_prec, double tol, int max_outer, bool transpose,
             h3d_solve_info* info);

// Every device allocation of the library goes through dmalloc, so that a test
// can make one of them fail (h3d_debug_fail_alloc): the countdown counts the
// allocations still allowed; the one found at zero fails and disarms it.
extern std::atomic<int64_t> g_alloc_countdown;  // < 0: no failure armed
inline void dmalloc(void** p, size_t bytes) {
  int64_t c = g_alloc_countdown.load();
  while (c >= 0 && !g_alloc_countdown.compare_exchange_weak(c, c - 1)) {
  }
  if (c == 0) throw Error(H3D_ERR_OOM, "device allocation failed (injected by h3d_debug_fail_alloc)");
  H3D_HIP(hipMalloc(p, bytes));
}

template <class T>
T* dalloc(int64_t n) {
  T* p = nullptr;
  if (n <= 0) return nullptr;
  dmalloc((void**)&p, sizeof(T) * (size_t)n);
  return p;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}
// Owner of a device buffer for the duration of one call: freed on every exit,
// exceptions included (hipFree waits for the device first).
template <class T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(int64_t n) : p(dalloc<T>(n)) {}
  ~DevBuf() { dfree(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Device copy of a host vector (at least one element), queued on `st`.
template <class T>
T* upload(const std::vector<T>& v, hipStream_t st) {
  T* d = dalloc<T>(std::max<int64_t>((int64_t)v.size(), 1));
  if (!v.empty())
    H3D_HIP(hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
  return d;
}

}  // namespace h3d
