// Products of many independent instances in one set of launches.  There is no
// reference counterpart: this is the caller side of MatVecProduct
// (src/HODLR3D.cpp:21-27, one vector of one instance per call) for many
// problems at once.
//
// A batch concatenates the work tables of its instances: the phase-1 items,
// row-split reductions, phase-2 items and near-field items of every instance,
// with chunk, block, t, partial-sum, point and leaf indices moved by the
// instance's offsets, plus one table of chunk base pointers into the
// instances' own factor memory (nothing of the factors is copied).  One
// product of every instance is then one launch per step for all of them:
//
//   gather (all points) -> phase 1 -> reduce -> phase 2 -> combine
//
// with the near field of all leaves on a second stream.  The kernels and the
// items are those of each instance's own product, so every instance's result
// equals its own product bit for bit (at the same load widths).  Small
// instances alone leave most of the GPU idle; merged, the launches hold every
// instance's waves together and fill it.
//
// Vectors are packed: instance i holds [off_i, off_{i+1}) of b and y, in its
// own original point order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "stream.hpp"

namespace h3d {

namespace {

// y[i] = alpha_i * (sum over levels of y_level at i's position + the near
// field there) + d_i * b[i], with alpha and d per point (each instance's
// operator and diagonal entries), as combine_kernel does for one instance.
__global__ void bcombine_kernel(const double* __restrict__ ylev, int nL, int64_t N,
                                const double* __restrict__ ynear,
                                const int32_t* __restrict__ iperm,
                                const double* __restrict__ alpha, const double* __restrict__ diag,
                                const double* __restrict__ b, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t s = iperm[i];
  double acc = 0.0;
  for (int j = 0; j < nL; ++j) acc = acc + ylev[(int64_t)j * N + s];
  if (ynear) acc = acc + ynear[s];
  y[i] = alpha[i] * acc + diag[i] * b[i];
}

template <class T>
std::vector<T> download(const T* d, int64_t n) {
  std::vector<T> h(std::max<int64_t>(n, 0));
  if (n > 0) H3D_HIP(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
  return h;
}

void check_i32(int64_t v, const char* what) {
  if (v > INT32_MAX) throw Error(H3D_ERR_INVALID, std::string("batch too large: ") + what);
}

}  // namespace

Batch::~Batch() { free_all(); }

void Batch::free_all() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (side) (void)hipStreamSynchronize(side);
  for (void** p : {(void**)&d_perm, (void**)&d_iperm, (void**)&d_alpha, (void**)&d_diag,
                   (void**)&d_blk, (void**)&d_ubase, (void**)&d_vbase, (void**)&d_p1,
                   (void**)&d_p1r, (void**)&d_p2, (void**)&d_near, (void**)&d_nbr,
                   (void**)&d_lstart, (void**)&d_lcount, (void**)&d_px, (void**)&d_py,
                   (void**)&d_pz, (void**)&d_lx, (void**)&d_ly, (void**)&d_lz, (void**)&d_leafc,
                   (void**)&d_xp, (void**)&d_xf, (void**)&d_t, (void**)&d_part, (void**)&d_ylev,
                   (void**)&d_ynear, (void**)&d_queue})
    dfree(*p);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (own) (void)hipStreamDestroy(own);
  ev_fork = ev_join = nullptr;
  side = own = stream = nullptr;
}

void build_batch(Batch& bt, Engine* const* es, int count, int mode) {
  if (count < 1) throw Error(H3D_ERR_INVALID, "a batch needs at least one instance");
  const int prec = mode & 0xff;
  if (prec < 0 || prec > 2 || (mode & ~(0xff | H3D_ACC_FP64)))
    throw Error(H3D_ERR_INVALID, "batches take the fp64, fp32 or bf16 factor copies");
  bt.prec = prec;
  bt.acc64 = prec == H3D_PREC_FP64 || (mode & H3D_ACC_FP64);
  bt.es.assign(es, es + count);
  for (int i = 0; i < count; ++i) {
    if (!es[i]) throw Error(H3D_ERR_INVALID, "null handle");
    for (int j = 0; j < i; ++j)
      if (es[j] == es[i]) throw Error(H3D_ERR_INVALID, "a handle appears twice in the batch");
    prepare_mode(*es[i], mode);
    H3D_HIP(hipStreamSynchronize(es[i]->stream));
    if (es[i]->kern.q != es[0]->kern.q || es[i]->kern.scale != es[0]->kern.scale)
      throw Error(H3D_ERR_INVALID, "the instances of a batch share the kernel (Qchoice and scale)");
  }
  bt.kd = es[0]->kern;
  bt.kd.diag_vec = nullptr;

  // offsets of every instance in the merged tables
  const int C = count;
  std::vector<int64_t> poff(C + 1, 0), boff(C + 1, 0), coff(C + 1, 0), toff(C + 1, 0),
      qoff(C + 1, 0), loff(C + 1, 0);
  int nL = 0;
  for (int i = 0; i < C; ++i) {
    const Engine& e = *es[i];
    poff[i + 1] = poff[i] + e.N;
    boff[i + 1] = boff[i] + e.nblk_all;
    coff[i + 1] = coff[i] + (int64_t)e.chunks.size();
    toff[i + 1] = toff[i] + e.T;
    qoff[i + 1] = qoff[i] + e.plan.npartial;
    loff[i + 1] = loff[i] + e.lev[e.nL].nboxes;
    nL = std::max(nL, e.nL);
  }
  check_i32(poff[C] + 64, "points");
  check_i32(boff[C], "blocks");
  check_i32(toff[C] + 1, "t entries");
  check_i32(qoff[C] + 1, "partial sums");
  check_i32(loff[C], "leaves");
  bt.N = poff[C];
  bt.nL = nL;
  bt.off = poff;
  bt.T = toff[C];
  bt.npart = qoff[C];
  hipStream_t st = nullptr;
  H3D_HIP(hipStreamCreateWithFlags(&bt.own, hipStreamNonBlocking));
  bt.stream = st = bt.own;
  H3D_HIP(hipStreamCreateWithFlags(&bt.side, hipStreamNonBlocking));
  H3D_HIP(hipEventCreateWithFlags(&bt.ev_fork, hipEventDisableTiming));
  H3D_HIP(hipEventCreateWithFlags(&bt.ev_join, hipEventDisableTiming));

  // per point (positions and original order) and per leaf
  std::vector<int32_t> perm(bt.N), iperm(bt.N);
  std::vector<double> alpha(bt.N), diag(bt.N);
  std::vector<int32_t> lstart(loff[C]), lcount(loff[C]), nbr(26 * loff[C]);
  std::vector<NearItem> near;
  for (int i = 0; i < C; ++i) {
    const Engine& e = *es[i];
    const int64_t p0 = poff[i], l0 = loff[i];
    for (int64_t s = 0; s < e.N; ++s) {
      perm[p0 + s] = (int32_t)(e.perm[s] + p0);
      iperm[p0 + e.perm[s]] = (int32_t)(p0 + s);
    }
    std::vector<double> dv;
    if (e.d_diag) dv = download(e.d_diag, e.N);
    for (int64_t j = 0; j < e.N; ++j) {
      alpha[p0 + j] = e.alpha;
      diag[p0 + j] = e.d_diag ? dv[j] : e.kern.diag;
    }
    const int64_t lo = level_offset(e.nL), nleaf = e.lev[e.nL].nboxes;
    const std::vector<int32_t> nb = download(e.lev[e.nL].d_nbr, 26 * nleaf);
    for (int64_t k = 0; k < nleaf; ++k) {
      lstart[l0 + k] = (int32_t)(e.bstart[lo + k] + p0);
      lcount[l0 + k] = e.bcount[lo + k];
      for (int q = 0; q < 26; ++q) {
        const int32_t v = nb[26 * k + q];
        nbr[26 * (l0 + k) + q] = v < 0 ? v : (int32_t)(v + l0);
      }
    }
    for (NearItem it : e.nitems) {
      it.leaf += (int32_t)l0;
      near.push_back(it);
    }
  }
  bt.nn = (int)near.size();
  bt.d_perm = upload(perm, st);
  bt.d_iperm = upload(iperm, st);
  bt.d_alpha = upload(alpha, st);
  bt.d_diag = upload(diag, st);
  bt.d_lstart = upload(lstart, st);
  bt.d_lcount = upload(lcount, st);
  bt.d_nbr = upload(nbr, st);
  bt.d_near = upload(near, st);
  // coordinates by position and leaf centres, copied from the instances
  bt.d_px = dalloc<double>(bt.N);
  bt.d_py = dalloc<double>(bt.N);
  bt.d_pz = dalloc<double>(bt.N);
  bt.d_lx = dalloc<float>(bt.N);
  bt.d_ly = dalloc<float>(bt.N);
  bt.d_lz = dalloc<float>(bt.N);
  bt.d_leafc = dalloc<double>(3 * std::max<int64_t>(loff[C], 1));
  for (int i = 0; i < C; ++i) {
    const Engine& e = *es[i];
    const int64_t p0 = poff[i], n = e.N;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (bytes) H3D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
    };
    cp(bt.d_px + p0, e.d_px, sizeof(double) * n);
    cp(bt.d_py + p0, e.d_py, sizeof(double) * n);
    cp(bt.d_pz + p0, e.d_pz, sizeof(double) * n);
    cp(bt.d_lx + p0, e.d_lx, sizeof(float) * n);
    cp(bt.d_ly + p0, e.d_ly, sizeof(float) * n);
    cp(bt.d_lz + p0, e.d_lz, sizeof(float) * n);
    cp(bt.d_leafc + 3 * loff[i], e.d_leafc, sizeof(double) * 3 * e.lev[e.nL].nboxes);
  }

  // factor tables: chunk bases, blocks and the phase items
  std::vector<void*> ub(std::max<int64_t>(coff[C], 1)), vb(std::max<int64_t>(coff[C], 1));
  std::vector<BlockDev> blk(std::max<int64_t>(boff[C], 1));
  std::vector<P1Item> p1;
  std::vector<P1Reduce> p1r;
  std::vector<P2Item> p2;
  for (int i = 0; i < C; ++i) {
    const Engine& e = *es[i];
    const int32_t c0 = (int32_t)coff[i], b0 = (int32_t)boff[i], t0 = (int32_t)toff[i],
                  q0 = (int32_t)qoff[i], p0 = (int32_t)poff[i];
    for (size_t c = 0; c < e.chunks.size(); ++c) {
      ub[c0 + c] = e.chunks[c].data[prec];
      vb[c0 + c] = (char*)e.chunks[c].data[prec] + e.chunks[c].nU * elem_size(prec);
    }
    if (e.nblk_all) {
      const std::vector<BlockDev> hb = download(e.d_blk, e.nblk_all);
      for (int64_t b = 0; b < e.nblk_all; ++b) {
        BlockDev d = hb[b];
        d.src += p0;
        d.chunk += c0;
        blk[b0 + b] = d;
      }
    }
    for (P1Item it : e.plan.p1) {
      it.chunk += c0;
      it.b0 += b0;
      it.b1 += b0;
      it.toff += t0;
      if (it.pslot >= 0) it.pslot += q0;
      p1.push_back(it);
    }
    for (P1Reduce r : e.plan.p1r) {
      r.pbase += q0;
      r.toff += t0;
      p1r.push_back(r);
    }
    const std::vector<P2Item>& src2 = prec == 0 ? e.plan.p2 : e.p2_narrow[prec - 1];
    for (P2Item it : src2) {
      it.chunk += c0;
      it.toff += t0;
      it.yoff += p0;
      p2.push_back(it);
    }
  }
  // larger phase-2 items first across all instances (less tail), as within one
  std::stable_sort(p2.begin(), p2.end(), [](const P2Item& a, const P2Item& b) {
    return (int64_t)a.R * a.nr > (int64_t)b.R * b.nr;
  });
  check_i32((int64_t)p1.size(), "phase-1 items");
  check_i32((int64_t)p2.size(), "phase-2 items");
  bt.n1 = (int)p1.size();
  bt.n1r = (int)p1r.size();
  bt.n2 = (int)p2.size();
  bt.nchunks = coff[C];
  bt.d_ubase = upload(ub, st);
  bt.d_vbase = upload(vb, st);
  bt.d_blk = upload(blk, st);
  bt.d_p1 = upload(p1, st);
  bt.d_p1r = upload(p1r, st);
  bt.d_p2 = upload(p2, st);

  // scratch; x is read in 16-byte vectors that may run up to 7 elements past a box
  bt.d_xp = dalloc<double>(bt.N + 64);
  bt.d_xf = dalloc<float>(bt.N + 64);
  H3D_HIP(hipMemsetAsync(bt.d_xp, 0, sizeof(double) * (bt.N + 64), st));
  H3D_HIP(hipMemsetAsync(bt.d_xf, 0, sizeof(float) * (bt.N + 64), st));
  bt.d_t = dalloc<double>(std::max<int64_t>(bt.T, 1));
  bt.d_part = dalloc<double>(std::max<int64_t>(bt.npart, 1));
  bt.d_ylev = dalloc<double>(std::max<int64_t>(bt.N * bt.nL, 1));
  bt.d_ynear = dalloc<double>(std::max<int64_t>(bt.N, 1));
  bt.d_queue = dalloc<unsigned>(2);
  for (int i = 0; i < C; ++i) bt.factor_bytes += es[i]->fac_bytes[prec];
  H3D_HIP(hipStreamSynchronize(st));
}

NearTables batch_near_tables(const Batch& bt) {
  NearTables nt;
  nt.items = bt.d_near;
  nt.n = bt.nn;
  nt.nbr = bt.d_nbr;
  nt.bstart = bt.d_lstart;
  nt.bcount = bt.d_lcount;
  nt.px = bt.d_px, nt.py = bt.d_py, nt.pz = bt.d_pz;
  nt.lx = bt.d_lx, nt.ly = bt.d_ly, nt.lz = bt.d_lz;
  nt.leafc = bt.d_leafc;
  nt.perm = bt.d_perm;
  nt.kd = bt.kd;
  nt.xp = bt.d_xp;
  nt.xf = bt.d_xf;
  nt.ynear = bt.d_ynear;
  return nt;
}

void batch_product(Batch& bt, const double* b, double* y, hipStream_t st, hipStream_t side,
                   hipEvent_t fork, hipEvent_t join) {
  const bool acc64 = bt.acc64;
  far_gather(b, bt.d_perm, bt.N, bt.d_xp, acc64 ? nullptr : bt.d_xf, st);
  far_zero(bt.d_ylev, bt.N * bt.nL, st);
  const NearTables nt = batch_near_tables(bt);
  const bool overlap = bt.nn > 0 && side;
  if (overlap) {
    H3D_HIP(hipEventRecord(fork, st));
    H3D_HIP(hipStreamWaitEvent(side, fork, 0));
    far_near(nt, bt.prec, acc64, false, side);
    H3D_HIP(hipEventRecord(join, side));
  }
  far_phase1(bt.prec, acc64, bt.tune_p1 & 0xff, bt.d_p1, bt.n1, bt.d_vbase, bt.d_blk, bt.d_xp,
             bt.d_xf, bt.d_t, bt.d_part, st,
             (bt.tune_p1 & H3D_PERSISTENT) ? bt.d_queue : nullptr);
  far_reduce(acc64, bt.d_p1r, bt.n1r, bt.d_part, bt.d_t, st);
  far_phase2(bt.prec, acc64, bt.tune_p2 & 0xff, bt.d_p2, bt.n2, bt.d_ubase, bt.d_t, bt.d_ylev,
             bt.N, st, (bt.tune_p2 & H3D_PERSISTENT) ? bt.d_queue + 1 : nullptr);
  if (overlap) H3D_HIP(hipStreamWaitEvent(st, join, 0));
  else far_near(nt, bt.prec, acc64, false, st);
  bcombine_kernel<<<(unsigned)((bt.N + 255) / 256), 256, 0, st>>>(
      bt.d_ylev, bt.nL, bt.N, bt.nn ? bt.d_ynear : nullptr, bt.d_iperm, bt.d_alpha, bt.d_diag, b,
      y);
  H3D_LAUNCH_CHECK();
}

void batch_matvec(Batch& bt, const double* b, double* y) {
  batch_product(bt, b, y, bt.stream, bt.side, bt.ev_fork, bt.ev_join);
}

void batch_autotune(Batch& bt, int reps, double* ms) {
  if (reps < 1) throw Error(H3D_ERR_INVALID, "reps must be >= 1");
  const int widths[4] = {16, 8, 4, 2};
  const int es = elem_size(bt.prec);
  hipStream_t st = bt.stream;
  hipEvent_t a = nullptr, z = nullptr;
  H3D_HIP(hipEventCreate(&a));
  try {
    H3D_HIP(hipEventCreate(&z));
    // timing does not depend on the values: x is whatever the last product left
    for (int which = 1; which <= 2; ++which) {
      auto run = [&](int w, bool persist) {
        if (which == 1)
          far_phase1(bt.prec, bt.acc64, w, bt.d_p1, bt.n1, bt.d_vbase, bt.d_blk, bt.d_xp, bt.d_xf,
                     bt.d_t, bt.d_part, st, persist ? bt.d_queue : nullptr);
        else
          far_phase2(bt.prec, bt.acc64, w, bt.d_p2, bt.n2, bt.d_ubase, bt.d_t, bt.d_ylev, bt.N, st,
                     persist ? bt.d_queue + 1 : nullptr);
      };
      auto time = [&](int w, bool persist) {
        run(w, persist);  // untimed
        H3D_HIP(hipEventRecord(a, st));
        for (int r = 0; r < reps; ++r) run(w, persist);
        H3D_HIP(hipEventRecord(z, st));
        H3D_HIP(hipEventSynchronize(z));
        float t = 0;
        H3D_HIP(hipEventElapsedTime(&t, a, z));
        return (double)t / reps;
      };
      double best = 1e300;
      int bw = 16;
      for (int q = 0; q < 4; ++q) {
        const int w = widths[q];
        double v = -1;
        if (w >= es) {
          v = time(w, false);
          if (v < best) best = v, bw = w;
        }
        if (ms) ms[(which - 1) * 4 + q] = v;
      }
      const double vp = time(bw, true);
      if (ms) ms[8 + which - 1] = vp;
      if (vp < best) bw |= H3D_PERSISTENT;
      (which == 1 ? bt.tune_p1 : bt.tune_p2) = bw;
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(a);
    if (z) (void)hipEventDestroy(z);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
}

}  // namespace h3d
