/*
 * hodlr3d.h -- C ABI of the MI355X HODLR3D engine (libhodlr3d.so).
 *
 * The reference (SAFRAN-LAB/HODLR3D) exposes C++ classes only; each entry
 * point below names the reference interface it stands in for.  All functions
 * return an int status (H3D_OK == 0).  Pointers are host memory unless a
 * `where` argument says H3D_DEVICE.  Calls on one handle are ordered on the
 * handle's HIP stream and are not re-entrant; distinct handles are independent.
 * (A product's near field runs on a second stream of the handle, forked after
 * the gather of x and joined before the combine with events on the handle's
 * stream, so each product still completes in order on that stream.)
 * Vectors are in the caller's (original) point order, like the reference's
 * MatVecProduct after `reorder` (src/HODLR3DTree.cpp:1795-1805).
 */
#ifndef HODLR3D_H_
#define HODLR3D_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: void APIs, UB on bad input) ---- */
#define H3D_OK 0
#define H3D_ERR_INVALID 1      /* bad argument, N < 1, nLevels outside [1,7], ... */
#define H3D_ERR_OOM 2          /* device or host allocation failed */
#define H3D_ERR_NONFINITE 3    /* non-finite kernel entry (duplicate points, NaN) */
#define H3D_ERR_HIP 4          /* HIP runtime error */
#define H3D_ERR_STATE 5        /* call out of order, e.g. matvec before assemble */
#define H3D_ERR_NOT_CONVERGED 6
#define H3D_ERR_INTERNAL 7
#define H3D_ERR_IO 8           /* file missing, unreadable, truncated or corrupt */

/* ---- stored-factor precisions ---- */
#define H3D_PREC_FP64 0  /* fp64 factors, fp64 accumulation, fp64 near field */
#define H3D_PREC_FP32 1  /* fp32 factors, fp32 accumulation, fp32 near field */
#define H3D_PREC_BF16 2  /* bf16 factors, fp32 accumulation, fp32 near field */
/* Recompressed copy (see h3d_compress): SVD-truncated factors with fp32 and
 * fp16 columns, fp64 accumulation, fp32 near-field entries summed in fp64. */
#define H3D_PREC_MIXED 3
/* Flag or-ed into a precision: accumulate in fp64 (fp64 near field) over
 * fp32/bf16-stored factors. */
#define H3D_ACC_FP64 0x100

/* Qchoice of the user kernel plugin compiled into the library
 * (include/h3d_kernel_plugin.hpp). */
#define H3D_QCHOICE_PLUGIN 100

/* ---- pointer locations ---- */
#define H3D_HOST 0
#define H3D_DEVICE 1

typedef struct h3d_handle h3d_handle;

/* Kernel plugin.  Replaces `userkernel::getMatrixEntry` (include/kernel.hpp:112,
 * dispatch examples/testHODLR3Dsolve.cpp:32-76): entry(i,j) = scale * phi_q(r_ij)
 * for i != j and the diagonal value for i == j.  qchoice follows the
 * reference's Qchoice numbering 0..16; scale = 1, diag = 0 reproduces the
 * reference for Qchoice 0..15.  For Qchoice 16 (IE_CUBE_3D,
 * src/kernel.cpp:65-73) pass scale = diag = 0 to get h3 = 1/N and
 * Kii = 1 + the self-cell integral (src/Integral3D.cpp:9-43).  With
 * qchoice = H3D_QCHOICE_PLUGIN the entry is scale * h3d_user_kernel::entry of
 * the two points, a device functor built into the library
 * (include/h3d_kernel_plugin.hpp): the reference's user-defined
 * getMatrixEntry.  Coincident points are accepted for kernels finite at
 * r = 0 (Qchoice 0-6, 11, a plugin finite there) and refused otherwise. */
typedef struct h3d_kernel {
  int qchoice;             /* 0..16, or H3D_QCHOICE_PLUGIN */
  double scale;
  double diag;
  const double* diag_vec; /* optional per-point diagonal (host, original order) */
} h3d_kernel;

typedef struct h3d_options {
  double center[3];        /* root cube centre; the reference uses the origin */
  int nlevels;             /* < 0: the reference formula (src/HODLR3D.cpp:6-7) */
  int64_t workspace_bytes; /* device workspace budget for assembly, 0 = automatic */
  int store_mask;          /* precisions written by h3d_assemble, bit p = H3D_PREC_p;
                              0 means fp64 only */
} h3d_options;

typedef struct h3d_stats {
  int64_t N;
  int nlevels;
  int max_rank;            /* HODLR3DTree::getMaxRank (src/HODLR3DTree.cpp:1807-1818) */
  int64_t n_blocks;        /* low-rank blocks (all levels) */
  int64_t rank_sum;        /* sum of block ranks */
  int64_t factor_elems;    /* sum over blocks of r*(n+m): U and V entries */
  double find_memory;      /* HODLR3DTree::findMemory entries (src/HODLR3DTree.cpp:1757-1778) */
  int64_t near_pairs;      /* sum over leaves of n_k * (sum of near-box sizes) */
  double t_tree_ms;        /* octree, bucketing, lists (device, wall) */
  double t_assemble_ms;    /* ACA + factor layout (device, wall) */
  int64_t factor_bytes[3]; /* resident factor bytes per precision (0 if absent) */
  int64_t u_elems;         /* sum over blocks of r*n (U entries, read by phase 2) */
  int64_t v_elems;         /* sum over blocks of r*m (V entries, read by phase 1) */
  /* Timers with the reference's meaning (HODLR3DTree::assTime / matVecTime,
   * include/HODLR3DTree.hpp:46, accumulated in src/HODLR3DTree.cpp:1660-1755
   * and never reset), in seconds, summed over every product made while
   * per-kernel timing is on (h3d_kernel_timing): ref_ass_time = regeneration
   * of kernel entries inside products (here the on-the-fly near field; the far
   * field stores its factors), ref_matvec_time = the low-rank BLAS-2 applies
   * (the far-field kernels). */
  double ref_ass_time;
  double ref_matvec_time;
} h3d_stats;

typedef struct h3d_mixed_stats {
  double tol;               /* requested relative Frobenius error */
  double predicted_rel_err; /* estimate of ||A_mixed - A_aca||_F / ||A||_F actually spent */
  double err_drop, err_fp16, err_fp32;   /* its parts: dropped columns, fp16 and fp32 rounding */
  double fro_far, fro_near; /* ||.||_F of the far (low-rank) and near (dense) parts */
  int64_t aca_rank_sum;     /* columns before recompression */
  int64_t rank_fp32, rank_fp16, dropped; /* columns per class and dropped */
  int64_t raw_blocks, raw_rank;          /* blocks kept as ACA columns in fp32 */
  int64_t bytes;            /* resident bytes of the copy */
  int64_t u_elems[2], v_elems[2];        /* per class (0: fp32, 1: fp16) */
  double build_ms;
  int single_jacobi;        /* 1: the build kept each block's Jacobi rotations between its
                               two passes (one Jacobi per block); 0: it ran them twice */
  int formats;              /* formats the copy could use (H3D_MIXED_*) */
  double err_fp8;           /* part of predicted_rel_err from fp8 rounding */
  int64_t rank_fp8;         /* columns in fp8 (e4m3) */
  int64_t u_elems_fp8, v_elems_fp8;
  double err_fp24;          /* part of predicted_rel_err from fp24 rounding */
  int64_t rank_fp24;        /* columns in fp24 (an fp32 without its last 8 bits) */
  int64_t u_elems_fp24, v_elems_fp24;
  double err_fp8_model;     /* with H3D_MIXED_INT8: the selection's model of err_fp8 (err_fp8
                               is then the realized value, from each column's largest entry) */
} h3d_mixed_stats;

/* Storage formats of the recompressed copy besides fp32 (h3d_set_mixed_formats). */
#define H3D_MIXED_FP16 1
#define H3D_MIXED_FP8 2
#define H3D_MIXED_FP24 4
/* The 8-bit tier in int8 with a scale per column side (its largest entry)
 * instead of e4m3: an error of rho / (127 sqrt 12) per side, rho = max / rms
 * of the column, modelled from the box size when the tiers are chosen. */
#define H3D_MIXED_INT8 8

typedef struct h3d_solve_info {
  int outer_iterations;    /* refinement steps */
  int inner_iterations;    /* total inner GMRES iterations */
  int outer_prec;          /* precision mode of the outer residual operator */
  double rel_residual;     /* final ||b - A x|| / ||b|| with the outer operator, outer_prec:
                              the fp64 factors if resident, else fp64 arithmetic over the
                              fp32 factors (H3D_PREC_FP32 | H3D_ACC_FP64) */
  double history[64];      /* relative residual after each refinement step */
} h3d_solve_info;

/* nLevels = ceil(3 log(N^(1/3)/leaf1d) / log 8): src/HODLR3D.cpp:6-7. */
int h3d_nlevels(int64_t N, int leaf1d);

/* Replaces HODLR3D::HODLR3D (src/HODLR3D.cpp:3-15): copies the points
 * (N x 3, pts3D layout, include/definitions.hpp:7-9), builds the octree,
 * interaction lists and point buckets on the GPU.  `opt` may be NULL. */
int h3d_create(const double* xyz, int64_t N, int leaf1d, double L, int tol_pow,
               const h3d_kernel* kernel, const h3d_options* opt, h3d_handle** out);

/* Replaces HODLR3D::assemble (src/HODLR3D.cpp:17-19 -> getUVtTree,
 * src/HODLR3DTree.cpp:1621-1645): pivot-exact ACA of every block on the GPU,
 * factors stored in fp64. */
int h3d_assemble(h3d_handle* h);

/* Saved assembled representation (no reference counterpart: the reference
 * keeps nothing on disk; the caller side of HODLR3D::assemble for repeated
 * runs).  h3d_save writes points, kernel, operator scaling, diagonal, tree
 * permutation, per-level ranks and offsets, factor chunks (the fp64 copy, else
 * the fp32 one) and ACA pivots to `path`, with a checksum.  h3d_load rebuilds
 * the tree from the saved points, checks it against the saved permutation and
 * every offset against its chunk, and returns an assembled handle whose
 * products equal the saved handle's bit for bit.  Only opt->workspace_bytes is
 * read from `opt` (may be NULL).  The mixed copy is not saved: h3d_compress
 * rebuilds it deterministically.  H3D_ERR_IO: missing, truncated, foreign or
 * corrupt file, or one of another format version (version 2 stores U columns
 * of more than 256 rows on 32-element strides; version 1 files are refused). */
int h3d_save(h3d_handle* h, const char* path);
int h3d_load(const char* path, const h3d_options* opt, h3d_handle** out);

/* Builds a copy of the factors in `prec` (from the fp64 copy, else the fp32
 * one), or frees one (H3D_PREC_MIXED too); the last fp64/fp32 copy cannot be
 * freed.  On H3D_ERR_OOM nothing of the new copy is kept and the others are
 * unchanged. */
int h3d_prepare_precision(h3d_handle* h, int prec);
/* Like h3d_prepare_precision, but frees every other copy chunk by chunk, so the
 * peak is one chunk above the larger of the two copies.  If memory runs out
 * after the first chunk has moved, no copy is whole any more: the factors are
 * dropped, the call returns H3D_ERR_OOM and the handle answers H3D_ERR_STATE
 * (create it again). */
int h3d_move_precision(h3d_handle* h, int prec);
int h3d_release_precision(h3d_handle* h, int prec);

/* Builds the recompressed mixed-precision copy (no reference counterpart; an
 * extension of the reference's HODLR3D::assemble, src/HODLR3D.cpp:17-19) from
 * the fp64 factors, else from the fp32 copy (same blocks; its 2^-24 rounding
 * per entry stays far below any useful budget): every block B = U V^T is
 * recompressed through the
 * eigendecomposition of its Gram matrices (B = U' V'^T, V' orthonormal), and
 * its columns are dropped, stored in fp8, fp16, fp24 or fp32 so that the expected
 * added relative error ||A_mixed - A||_F / ||A||_F stays within `tol` (one
 * budget over all blocks, spent on the columns with the least error per stored
 * element).  Needs the fp64 or the fp32 copy resident; stats may be NULL.  A matvec with
 * H3D_PREC_MIXED builds the copy with tol = 5 * 10^-TOL_POW if absent. */
int h3d_compress(h3d_handle* h, double tol, h3d_mixed_stats* stats);
/* Formats the next build of the recompressed copy may store columns in besides
 * fp32 (no reference counterpart): any of H3D_MIXED_FP16, H3D_MIXED_FP8 and
 * H3D_MIXED_FP24, or 0 (fp32 columns only), plus H3D_MIXED_INT8 (the 8-bit
 * tier in int8; it enables that tier); all four by default.  Each
 * format is a tier of the budget's choice: a column of the least error per
 * stored element is dropped, the next ones stored in fp8 (e4m3, 1 byte), fp16
 * (2 bytes), fp24 (3 bytes: fp32 rounded to 16 significant bits) or fp32, each
 * tier over a fixed range of that error measure (the ranges minimise the bytes
 * for the budget); a format turned off is skipped.  The copy in place is kept
 * until the next h3d_compress. */
int h3d_set_mixed_formats(h3d_handle* h, int formats);
int h3d_get_mixed_stats(h3d_handle* h, h3d_mixed_stats* stats);

/* Replaces HODLR3D::MatVecProduct (src/HODLR3D.cpp:21-27): y = A b for nrhs
 * column-major right-hand sides of length N. */
int h3d_matvec(h3d_handle* h, const double* b, double* y, int nrhs, int prec, int where);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the product of h3d_matvec for one right-hand side, b and
 * y in device memory, whose scratch (x by position, t, partial sums, per-level
 * y, near field) is carved from a static device array of 64 MiB that each GPU
 * holds until the process ends, not from the handle's buffers; it calls no
 * hipMalloc or hipFree.  A product needing more than 64 MiB of scratch
 * (N * (20 + 8 nLevels) bytes, plus 8 bytes per low-rank column -- twice that
 * for the mixed copy -- and the partial sums: up to about 10^5 points)
 * returns H3D_ERR_INVALID, and a precision whose factor copy is
 * not resident H3D_ERR_STATE (build it first with h3d_prepare_precision or
 * h3d_compress).  Results equal h3d_matvec's bit for bit.  Every call on one
 * GPU shares its array: calls that overlap on one GPU (two streams, two
 * threads) race, so serialise them. */
int h3d_matvec_noalloc(h3d_handle* h, const double* b, double* y, int prec);

/* y = A^T b with the same stored factors (no reference counterpart; the
 * adjoint of MatVecProduct). */
int h3d_matvec_transpose(h3d_handle* h, const double* b, double* y, int nrhs, int prec,
                         int where);

/* Replaces the per-point diagonal K(i,i) after creation (NULL: back to the
 * kernel's scalar diagonal). */
int h3d_set_diag(h3d_handle* h, const double* diag_vec, int where);

/* Apply-time operator: y = alpha * (assembled off-diagonal operator) b + d b,
 * d = the per-point diagonal if set, else `diag`.  Defaults: alpha = 1 and the
 * kernel's diagonal.  Lets one assembly of K serve A = diag + alpha K. */
int h3d_set_operator(h3d_handle* h, double alpha, double diag);

/* Autotuner: times the far-field kernels of precision mode `prec` with every
 * per-lane load width (16/8/4/2 bytes) on the assembled problem, `reps`
 * launches each, and keeps the fastest for later matvecs.  ms (may be NULL)
 * receives phase-1 then phase-2 averages per width, -1 where a width does not
 * apply: 8 values, or 16 for H3D_PREC_MIXED (fp32 class, then fp16 class; the
 * fp24 class runs with the fp32 class's settings, the fp8 class with the fp16
 * class's, and fp24 always loads four entries per lane in both phases). */
int h3d_autotune(h3d_handle* h, int prec, int reps, double* ms);

/* Products with several right-hand sides (h3d_matvec with nrhs >= 2 on an
 * ACA factor copy) take them `width` at a time (16 or 32; 0: one at a time)
 * and read each factor element once per group, on the matrix cores (default
 * 32).  No reference counterpart: MatVecProduct (src/HODLR3D.cpp:21-27) takes
 * one vector.  Transposes and the mixed copy always go one RHS at a time. */
int h3d_set_rhs_block(h3d_handle* h, int width);

/* Autotuner of the grouped path: times the product of `nrhs` (>= 2) columns
 * in mode `prec` (an ACA copy) for group widths 16 and 32 and phase-1 items
 * of 1024 and 2048 columns on sources of at most 64 points (the leaves of a
 * uniform tree), `reps` products each, and keeps the fastest (the
 * results do not depend on the choice).  ms[4] (may be NULL) receives the
 * averages, width-major, -1 where a width does not apply; best[2] (may be
 * NULL) the chosen width and item size. */
int h3d_autotune_rhs(h3d_handle* h, int prec, int nrhs, int reps, double* ms, int* best);

/* Pins the per-lane load widths (bytes, >= the element size) for one mode.
 * With H3D_PREC_MIXED the widths apply to both classes, phase 1 takes 16
 * bytes only (its items span 64 lanes x 4 vectors of 16 bytes of rows) and
 * phase 2 at least 4. */
int h3d_set_tuning(h3d_handle* h, int prec, int p1_bytes, int p2_bytes);

/* The complete tuning state of mode `prec` (the autotuner's choice or what was
 * pinned): v[0], v[1] = phase-1 / phase-2 load bytes per lane (for
 * H3D_PREC_MIXED those of its fp32 and fp24 classes), v[2], v[3] = those of the
 * mixed copy's fp16 and fp8 classes, v[4], v[5] = the mixed phase-1 kernel's
 * load vectors of rows per lane for each of the two groups (4 or 1; 0: chosen
 * per item); 0 where a slot does not apply.  h3d_set_tuning_state pins a state read back earlier, so that a
 * profiled run can time exactly the kernel variants an autotuned run used. */
int h3d_get_tuning(h3d_handle* h, int prec, int* v);
int h3d_set_tuning_state(h3d_handle* h, int prec, const int* v);

/* Per-kernel timing of subsequent matvecs with HIP events on the handle's
 * stream (enable != 0).  ms[15] receives the accumulated milliseconds of
 * phase 1 (V^T x, with its partial-sum reduction and, for H3D_PREC_MIXED,
 * the gather into target order), phase 2 (U t), near field, combine and whole
 * matvecs since the last read; then the parts of phase 1 and phase 2 spent on
 * the second factor copy (the fp16 class of H3D_PREC_MIXED; 0 otherwise); then
 * the phase-1 kernel alone of the first and of the second copy; then the
 * parts of phase 1 and phase 2 spent on the third copy (the fp8 class) and its
 * phase-1 kernel alone; then the same three for the fourth copy (the fp24
 * class).  *count gets
 * the number of matvecs; all are then reset.  Grouped products of several
 * right-hand sides (h3d_set_rhs_block) are not timed.  While timing is on,
 * every kernel of a matvec runs in line on the handle's stream; otherwise the
 * near field runs on a second stream of the handle, overlapping the far field
 * (forked after the gather of x, joined before the combine). */
int h3d_kernel_timing(h3d_handle* h, int enable, double* ms, int* count);

/* Replaces classGMRES::gmres (src/gmres.cpp:40-94): unrestarted GMRES, x0 = 0,
 * modified Gram-Schmidt, Givens rotations.  hist receives error values (its+1
 * entries, at most maxit+1). */
int h3d_gmres(h3d_handle* h, const double* b, double* x, int maxit, double threshold,
              double* error, int* iterations, double* hist, int where);

/* GMRES(restart) with the `inner_prec` operator inside fp64 iterative
 * refinement: r = b - A x in fp64 (fp64 factors if resident, else fp64
 * arithmetic over the fp32 factors), inner solve to `inner_tol`, x += d, until
 * ||r||/||b|| <= tol.  transpose != 0 solves A^T x = b. */
int h3d_solve_ir(h3d_handle* h, const double* b, double* x, int restart, int max_inner,
                 double inner_tol, int inner_prec, double tol, int max_outer, int transpose,
                 h3d_solve_info* info, int where);

/* Direct O(N) FP64 evaluation of rows of the same operator (every entry from
 * the kernel, compensated summation): out[r] = (A b)[rows[r]].  rows are
 * original indices in host memory; b and out follow `where`.  The reference's
 * drivers form exact products from kernel columns (examples/testHODLR3D.cpp:139-143). */
int h3d_direct(h3d_handle* h, const double* b, const int64_t* rows, int64_t nrows, double* out,
               int where);

/* Kernel entries: out[r + c * nrows] = K(rows[r], cols[c]) for original
 * indices (host arrays), the diagonal value where they are equal; out follows
 * `where`.  Replaces kernel::getMatrix(row_indices, col_indices)
 * (src/kernel.cpp:37-50); kernel::getRow(j, cols) (:7-15) and getCol(k, rows)
 * / getCol(k) (:17-35) are its one-row and one-column cases.  Entries of the
 * kernel only: the apply-time alpha of h3d_set_operator is not applied. */
int h3d_get_matrix(h3d_handle* h, const int64_t* rows, int64_t nrows, const int64_t* cols,
                   int64_t ncols, double* out, int where);

/* Name of the kernel plugin compiled into the library (Qchoice
 * H3D_QCHOICE_PLUGIN). */
const char* h3d_plugin_name(void);

/* Fault injection for tests (no reference counterpart): after `after` more
 * device allocations of the library (process-wide), the next one fails with
 * H3D_ERR_OOM and the hook disarms itself; a negative value disarms it. */
int h3d_debug_fail_alloc(int64_t after);

/* Structure export for bit-exact checks (the reference keeps these in
 * HODLR3DBox, include/HODLR3DTree.hpp:11-24). */
int h3d_export_perm(h3d_handle* h, int32_t* perm);                 /* leaf-DFS order */
int h3d_export_boxes(h3d_handle* h, int level, int32_t* start, int32_t* count);
int h3d_export_neighbors(h3d_handle* h, int level, int32_t* nbr26);
int64_t h3d_il_size(h3d_handle* h, int level);
int h3d_export_il(h3d_handle* h, int level, int32_t* offsets, int32_t* boxes);
int h3d_export_ranks(h3d_handle* h, int level, int32_t* rank, int32_t* status);
/* Local pivots of block b (row within the target box, column within the
 * source box) and, if U/V are non-NULL, the ACA vectors (rank x n, rank x m). */
int h3d_export_block(h3d_handle* h, int level, int64_t b, int32_t* I, int32_t* J, double* U,
                     double* V);

int h3d_get_stats(h3d_handle* h, h3d_stats* s);

/* The construction parameters of a handle (those of h3d_create, or read back
 * from a saved file by h3d_load); kernel_scale/kernel_diag are the effective
 * values (Qchoice 16 with scale = diag = 0 resolves to h^3 = 1/N and K_ii).
 * alpha and the scalar diagonal are the current operator (h3d_set_operator). */
typedef struct h3d_config {
  int64_t N;
  int nlevels;
  int leaf1d;
  double L;
  double center[3];
  int tol_pow;
  int qchoice;
  double kernel_scale;
  double kernel_diag;
  double alpha;
  int has_diag_vec;
} h3d_config;
int h3d_get_config(h3d_handle* h, h3d_config* c);

/* Batched products in one HIP graph (no reference counterpart; the caller side
 * of MatVecProduct for many instances).  h3d_graph_create runs and captures one
 * matvec (or transpose) of each handle, hs[i]: b[i] -> y[i] (device pointers,
 * baked into the graph), in mode `prec`; h3d_graph_launch replays all of them
 * with one launch on `stream` (NULL: the legacy default stream); products of
 * distinct handles are independent branches of the graph and may run
 * concurrently (those of one handle stay in order).  The caller rewrites the
 * contents of b[i] between launches; the y[i] must not overlap.  The graph stays valid while
 * every handle keeps the factor copies and scratch it had at capture: no
 * prepare/move/release of precisions, compress or larger nrhs in between.  The
 * handles' own streams are restored after the capture. */
typedef struct h3d_graph h3d_graph;
int h3d_graph_create(h3d_handle* const* hs, int count, const double* const* b, double* const* y,
                     int prec, int transpose, h3d_graph** out);
int h3d_graph_launch(h3d_graph* g, void* stream);
void h3d_graph_destroy(h3d_graph* g);

/* Instance batch (no reference counterpart; the caller side of MatVecProduct,
 * src/HODLR3D.cpp:21-27, for many independent problems): one product of each
 * of `count` assembled handles in one launch per step for all of them.  The
 * work tables of the instances (far-field items, row-split reductions,
 * near-field items, chunk base pointers into each handle's own factors) are
 * merged at creation, so a product is gather -> phase 1 -> reduce -> phase 2
 * -> combine over every instance at once, with the near field of all leaves on
 * a second stream.  Each instance's result equals its own h3d_matvec in the
 * same mode bit for bit when both use the same load widths.
 *   mode: H3D_PREC_FP64, H3D_PREC_FP32 or H3D_PREC_BF16, optionally
 *         | H3D_ACC_FP64 (the ACA copies; the copy is made where missing).
 * All handles share the kernel (Qchoice and scale); each keeps its own operator
 * (alpha, diagonal) as set at creation.  A handle may appear once.
 * Vectors are packed device arrays of float64: instance i owns entries
 * [off[i], off[i+1]) of b and y in its own point order (h3d_batch_layout
 * writes the count + 1 offsets).  The batch reads the handles' factors: it
 * stays valid while every handle lives and keeps the copy of `mode` (no
 * release, move or destroy in between).  Products run on the batch's own
 * stream until h3d_batch_set_stream names another (NULL: the legacy default
 * stream; the new stream first waits for the old one's work); products and
 * graph replays of one batch share its scratch, so the caller keeps them in
 * order.
 * Each far-field phase launches either one wave per item or persistently
 * (H3D_PERSISTENT): a grid of as many workgroups as stay resident, whose
 * waves take items from a work queue (an atomic counter, cleared by a kernel
 * before the launch) until every item of every instance is taken.
 * h3d_batch_autotune times the load widths (2..16 bytes per lane) of both
 * phases over the merged tables, then the persistent launch at the fastest
 * width, and keeps the fastest (ms[10]: ms[0..7] as h3d_autotune's, ms[8],
 * ms[9] phase 1 / phase 2 persistent); h3d_batch_set_tuning / get_tuning pin
 * or read the choice (load bytes | H3D_PERSISTENT).
 * h3d_batch_graph_create captures one product b -> y (device pointers baked
 * in) into a graph replayed by h3d_graph_launch and freed by
 * h3d_graph_destroy. */
#define H3D_PERSISTENT 0x100
typedef struct h3d_batch h3d_batch;
int h3d_batch_create(h3d_handle* const* hs, int count, int mode, h3d_batch** out);
int h3d_batch_layout(h3d_batch* bt, int64_t* offsets);
int h3d_batch_matvec(h3d_batch* bt, const double* d_b, double* d_y);
int h3d_batch_set_stream(h3d_batch* bt, void* stream);
int h3d_batch_synchronize(h3d_batch* bt);
int h3d_batch_autotune(h3d_batch* bt, int reps, double* ms);
int h3d_batch_set_tuning(h3d_batch* bt, int p1_bytes, int p2_bytes);
int h3d_batch_get_tuning(h3d_batch* bt, int* p1_bytes, int* p2_bytes);
int h3d_batch_graph_create(h3d_batch* bt, const double* d_b, double* d_y, h3d_graph** out);
void h3d_batch_destroy(h3d_batch* bt);

/* Run subsequent calls on a caller-owned hipStream_t (passed as void*; NULL is
 * the legacy default stream); h3d_use_own_stream returns to the handle's own.
 * Either switch makes the new stream wait for the work already queued on the
 * old one (an event, also valid under stream capture), since all products of
 * a handle share its scratch. */
int h3d_set_stream(h3d_handle* h, void* stream);
int h3d_use_own_stream(h3d_handle* h);
int h3d_synchronize(h3d_handle* h);

/* Replaces HODLR3D::~HODLR3D (src/HODLR3D.cpp:29-32). */
void h3d_destroy(h3d_handle* h);

const char* h3d_status_string(int status);
/* Text of the last error raised on this thread. */
const char* h3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HODLR3D_H_ */
