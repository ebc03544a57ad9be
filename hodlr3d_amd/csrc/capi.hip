// The C ABI (include/hodlr3d.h): argument checks, handle lifetime, host/device
// staging and error translation.  Every entry point returns a status code;
// exceptions never cross the boundary.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch.hpp"
#include "engine.hpp"
#include "kernels.hpp"

struct h3d_handle {
  h3d::Engine e;
};

namespace h3d {

std::atomic<int64_t> g_alloc_countdown{-1};

Engine::~Engine() { free_all(); }

void Engine::free_all() {
  dfree(d_px);
  dfree(d_py);
  dfree(d_pz);
  dfree(d_lx);
  dfree(d_ly);
  dfree(d_lz);
  dfree(d_perm);
  dfree(d_iperm);
  dfree(d_diag);
  dfree(d_leafc);
  dfree(d_bstart);
  dfree(d_bcount);
  for (Level& lv : lev) {
    dfree(lv.d_iloff);
    dfree(lv.d_ilidx);
    dfree(lv.d_nbr);
  }
  lev.clear();
  for (Chunk& c : chunks) {
    for (int p = 0; p < 3; ++p)
      if (c.data[p]) (void)hipFree(c.data[p]);
    dfree(c.piv);
  }
  chunks.clear();
  for (int p = 0; p < 3; ++p) {
    dfree(d_ubase[p]);
    dfree(d_vbase[p]);
    fac_bytes[p] = 0;
  }
  mix.free_all();
  dfree(d_blk);
  dfree(d_by_src);
  dfree(d_btoff);
  plan.free_device();
  for (int q = 0; q < 2; ++q) {
    dfree(d_p2_narrow[q]);
    p2_narrow[q].clear();
  }
  dfree(d_near);
  dfree(d_xp);
  dfree(d_xf);
  dfree(d_t);
  dfree(d_part);
  dfree(d_ylev);
  dfree(d_ynear);
  if (side) (void)hipStreamSynchronize(side);
  if (side2) (void)hipStreamSynchronize(side2);
  for (hipEvent_t* x : {&ev_fork, &ev_join, &ev_join2, &ev_switch})
    if (*x) (void)hipEventDestroy(*x), *x = nullptr;
  if (side) (void)hipStreamDestroy(side);
  if (side2) (void)hipStreamDestroy(side2);
  side = side2 = nullptr;
  dfree(d_io);
  dfree(d_xw);
  dfree(d_tw);
  dfree(d_ylevw);
  dfree(d_ynw);
  dfree(d_iow);
  for (GroupTables& g : grp) g.free_all();
  release_mixed_groups(*this);
  w_cap = io_w = tw_w = 0;
  for (auto& ev : this->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (own_stream) (void)hipStreamDestroy(own_stream);
  own_stream = nullptr;
}

// Gauss-Legendre rule on [-1,1] by Newton's iteration on P_n from the
// Chebyshev-like initial guesses (textbook; see Abramowitz & Stegun 25.4.29).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0);
  w.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    long double z = std::cos(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double dp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) {
        long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      long double dz = p1 / dp;
      z -= dz;
      if (std::fabs((double)dz) < 1e-20) break;
    }
    x[i] = -(double)z;
    w[i] = (double)(2 / ((1 - z * z) * dp * dp));
  }
}

// IE_CUBE_3D self term (include/kernel.hpp:45-60, src/Integral3D.cpp:9-43):
// 1 + 8 * integral over [0,h/2]^3 of 1/(4 pi r), h = N^(-1/3), 16^3 points.
static double ie_cube_kii(int64_t N) {
  std::vector<double> x, w;
  gauss_legendre(16, x, w);
  const double h = 1.0 / std::cbrt(double(N));
  const double c = 0.5 * (h * 0.5), l = 0.5 * (h * 0.5);
  double acc = 0.0;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j)
      for (int k = 0; k < 16; ++k) {
        double tx = c * x[i] + l, ty = c * x[j] + l, tz = c * x[k] + l;
        acc += w[i] * w[j] * w[k] / (4 * 3.1415926535897932384 * std::sqrt(tx * tx + ty * ty + tz * tz));
      }
  return acc * (c * c * c) * 8 + 1;
}

}  // namespace h3d

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return H3D_OK;
  } catch (const h3d::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return H3D_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return H3D_ERR_INTERNAL;
  }
}

int nlevels_formula(int64_t N, int leaf1d) {
  // src/HODLR3D.cpp:6-7
  double cubeRootN = std::pow((double)N, 1.0 / 3.0);
  return (int)std::ceil(3 * std::log(double(cubeRootN) / leaf1d) / std::log(8));
}

void check_level(h3d_handle* h, int level) {
  if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
  if (level < 1 || level > h->e.nL) throw h3d::Error(H3D_ERR_INVALID, "level out of range");
}

}  // namespace

extern "C" {

int h3d_nlevels(int64_t N, int leaf1d) {
  if (N < 1 || leaf1d < 1) return -1;
  return nlevels_formula(N, leaf1d);
}

int h3d_create(const double* xyz, int64_t N, int leaf1d, double L, int tol_pow,
               const h3d_kernel* kernel, const h3d_options* opt, h3d_handle** out) {
  if (out) *out = nullptr;
  return guarded([&] {
    if (!xyz || !out) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (N < 1 || N > INT32_MAX) throw h3d::Error(H3D_ERR_INVALID, "N out of range");
    if (leaf1d < 1) throw h3d::Error(H3D_ERR_INVALID, "leaf1d must be >= 1");
    if (!(L > 0) || !std::isfinite(L)) throw h3d::Error(H3D_ERR_INVALID, "L must be positive");
    for (int64_t i = 0; i < 3 * N; ++i)
      if (!std::isfinite(xyz[i])) throw h3d::Error(H3D_ERR_NONFINITE, "non-finite coordinate");
    int nL = (opt && opt->nlevels >= 0) ? opt->nlevels : nlevels_formula(N, leaf1d);
    if (nL < 1 || nL > 7)
      throw h3d::Error(H3D_ERR_INVALID, "nLevels must be in [1,7] (got " + std::to_string(nL) + ")");
    h3d_kernel k{7, 1.0, 0.0, nullptr};
    if (kernel) k = *kernel;
    if ((k.qchoice < 0 || k.qchoice > 16) && k.qchoice != h3d::kQPlugin)
      throw h3d::Error(H3D_ERR_INVALID, "qchoice out of range");
    auto* h = new h3d_handle();
    try {
      h3d::Engine& e = h->e;
      e.N = N;
      e.nL = nL;
      e.leaf1d = leaf1d;
      e.L = L;
      if (opt) {
        e.ctr[0] = opt->center[0];
        e.ctr[1] = opt->center[1];
        e.ctr[2] = opt->center[2];
        e.ws_budget = opt->workspace_bytes;
        if (opt->store_mask & ~7) throw h3d::Error(H3D_ERR_INVALID, "bad store_mask");
        e.store_mask = opt->store_mask ? opt->store_mask : 1;
      }
      e.tol_pow = tol_pow;
      e.tol = std::pow(10, -1.0 * tol_pow);  // src/ACA.cpp:5
      e.kern.q = k.qchoice;
      e.kern.scale = k.scale;
      e.kern.diag = k.diag;
      if (k.qchoice == 16 && k.scale == 0.0 && k.diag == 0.0) {
        e.kern.scale = 1.0 / double(N);  // include/kernel.hpp:59
        e.kern.diag = h3d::ie_cube_kii(N);
      }
      H3D_HIP(hipStreamCreateWithFlags(&e.own_stream, hipStreamNonBlocking));
      e.stream = e.own_stream;
      if (k.diag_vec) {
        e.d_diag = h3d::dalloc<double>(N);
        H3D_HIP(hipMemcpyAsync(e.d_diag, k.diag_vec, sizeof(double) * N, hipMemcpyHostToDevice,
                               e.stream));
      }
      h3d::build_tree(e, xyz);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int h3d_assemble(h3d_handle* h) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    h3d::Engine& e = h->e;
    if (e.assembled) throw h3d::Error(H3D_ERR_STATE, "already assembled");
    if (e.factors_lost)
      throw h3d::Error(H3D_ERR_STATE, "factors lost in a failed move: create the handle again");
    h3d::assemble_all(e);
    h3d::build_plan(e);
    h3d::ensure_scratch(e, 1);
  });
}

int h3d_save(h3d_handle* h, const char* path) {
  return guarded([&] {
    if (!h || !path) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (h->e.kern.q == h3d::kQPlugin)
      throw h3d::Error(H3D_ERR_INVALID, "plugin kernels are not saved (a file cannot name the "
                                        "plugin built into the library that reads it)");
    h3d::save_engine(h->e, path);
  });
}

int h3d_load(const char* path, const h3d_options* opt, h3d_handle** out) {
  if (out) *out = nullptr;
  return guarded([&] {
    if (!path || !out) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    auto* h = new h3d_handle();
    try {
      h3d::load_engine(h->e, path, opt ? opt->workspace_bytes : 0);
      h3d::build_plan(h->e);
      h3d::ensure_scratch(h->e, 1);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int h3d_prepare_precision(h3d_handle* h, int prec) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (prec < 0 || prec > H3D_PREC_MIXED) throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
    if (!h->e.assembled) throw h3d::Error(H3D_ERR_STATE, "not assembled");
    if (prec == H3D_PREC_MIXED) {
      if (!h->e.mix.ready) h3d::build_mixed(h->e, 5.0 * h->e.tol);
      return;
    }
    h3d::prepare_precision(h->e, prec);
  });
}

int h3d_move_precision(h3d_handle* h, int prec) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (prec < 0 || prec > 2) throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
    if (!h->e.assembled) throw h3d::Error(H3D_ERR_STATE, "not assembled");
    h3d::prepare_precision(h->e, prec, true);
  });
}

int h3d_set_operator(h3d_handle* h, double alpha, double diag) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (!std::isfinite(alpha) || !std::isfinite(diag)) throw h3d::Error(H3D_ERR_INVALID, "non-finite");
    h->e.alpha = alpha;
    h->e.kern.diag = diag;
  });
}

int h3d_autotune_rhs(h3d_handle* h, int prec, int nrhs, int reps, double* ms, int* best) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (!h->e.assembled) throw h3d::Error(H3D_ERR_STATE, "autotune before assemble");
    h3d::autotune_rhs(h->e, prec, nrhs, reps, ms, best);
  });
}

int h3d_set_rhs_block(h3d_handle* h, int width) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (width != 0 && width != 16 && width != 32)
      throw h3d::Error(H3D_ERR_INVALID, "width must be 0, 16 or 32");
    h->e.rhs_block = width;
  });
}

int h3d_set_tuning(h3d_handle* h, int prec, int p1_bytes, int p2_bytes) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    const int p = prec & 0xff;
    if (p < 0 || p > H3D_PREC_MIXED || (prec & ~(0xff | H3D_ACC_FP64)))
      throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
    // the mixed copy's fp32 class bounds its widths from below
    const int es = p == H3D_PREC_MIXED ? 4 : h3d::elem_size(p);
    auto ok = [&](int b) { return (b == 16 || b == 8 || b == 4 || b == 2) && b >= es; };
    if (!ok(p1_bytes) || !ok(p2_bytes)) throw h3d::Error(H3D_ERR_INVALID, "bad load width");
    if (p == H3D_PREC_MIXED) {  // both classes: fp32 (slot 6) and fp16 (slot 7)
      if (p1_bytes != 16)
        throw h3d::Error(H3D_ERR_INVALID, "the mixed phase 1 takes 16-byte loads");
      for (int key : {6, 7}) {
        h->e.tune_p1[key] = p1_bytes;
        h->e.tune_p2[key] = p2_bytes;
      }
      return;
    }
    const int key = p + ((prec & H3D_ACC_FP64) && p != 0 ? 3 : 0);
    h->e.tune_p1[key] = p1_bytes;
    h->e.tune_p2[key] = p2_bytes;
  });
}

namespace {
// Tuning slots of a mode: {phase-1/phase-2 slot, second slot or -1}.
void tuning_slots(int prec, int& a, int& b) {
  const int p = prec & 0xff;
  if (p < 0 || p > H3D_PREC_MIXED || (prec & ~(0xff | H3D_ACC_FP64)))
    throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
  if (p == H3D_PREC_MIXED) {
    a = 6;
    b = 7;
    return;
  }
  a = p + ((prec & H3D_ACC_FP64) && p != 0 ? 3 : 0);
  b = -1;
}
}  // namespace

int h3d_get_tuning(h3d_handle* h, int prec, int* v) {
  return guarded([&] {
    if (!h || !v) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    int a, b;
    tuning_slots(prec, a, b);
    const h3d::Engine& e = h->e;
    v[0] = e.tune_p1[a];
    v[1] = e.tune_p2[a];
    v[2] = b >= 0 ? e.tune_p1[b] : 0;
    v[3] = b >= 0 ? e.tune_p2[b] : 0;
    v[4] = b >= 0 ? e.tune_k[a] : 0;
    v[5] = b >= 0 ? e.tune_k[b] : 0;
  });
}

int h3d_set_tuning_state(h3d_handle* h, int prec, const int* v) {
  return guarded([&] {
    if (!h || !v) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    int a, b;
    tuning_slots(prec, a, b);
    const int p = prec & 0xff;
    const int es = p == H3D_PREC_MIXED ? 4 : h3d::elem_size(p);
    auto ok = [&](int w, int lo) { return (w == 16 || w == 8 || w == 4 || w == 2) && w >= lo; };
    if (!ok(v[0], es) || !ok(v[1], es)) throw h3d::Error(H3D_ERR_INVALID, "bad load width");
    if (b >= 0) {
      if (v[0] != 16 || v[2] != 16)
        throw h3d::Error(H3D_ERR_INVALID, "the mixed phase 1 takes 16-byte loads");
      if (!ok(v[3], 2)) throw h3d::Error(H3D_ERR_INVALID, "bad load width");
      for (int q : {4, 5})
        if (v[q] != 0 && v[q] != 1 && v[q] != 4)
          throw h3d::Error(H3D_ERR_INVALID, "rows per lane must be 0, 1 or 4");
    }
    h3d::Engine& e = h->e;
    e.tune_p1[a] = v[0];
    e.tune_p2[a] = v[1];
    if (b >= 0) {
      e.tune_p1[b] = v[2];
      e.tune_p2[b] = v[3];
      e.tune_k[a] = v[4];
      e.tune_k[b] = v[5];
    }
  });
}

int h3d_compress(h3d_handle* h, double tol, h3d_mixed_stats* out) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (!h->e.assembled) throw h3d::Error(H3D_ERR_STATE, "not assembled");
    h3d::build_mixed(h->e, tol);
    if (out) h3d_get_mixed_stats(h, out);
  });
}

int h3d_get_mixed_stats(h3d_handle* h, h3d_mixed_stats* s) {
  return guarded([&] {
    if (!h || !s) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    const h3d::MixedCopy& m = h->e.mix;
    std::memset(s, 0, sizeof(*s));
    if (!m.ready) return;
    s->tol = m.tol;
    const double a2 = m.fro2_far + m.fro2_near;
    s->predicted_rel_err =
        a2 > 0 ? std::sqrt((m.err2_drop + m.err2_fp8 + m.err2_fp16 + m.err2_fp24 + m.err2_fp32) / a2)
               : 0;
    s->err_drop = a2 > 0 ? std::sqrt(m.err2_drop / a2) : 0;
    s->err_fp16 = a2 > 0 ? std::sqrt(m.err2_fp16 / a2) : 0;
    s->err_fp32 = a2 > 0 ? std::sqrt(m.err2_fp32 / a2) : 0;
    s->fro_far = std::sqrt(m.fro2_far);
    s->fro_near = std::sqrt(m.fro2_near);
    s->aca_rank_sum = m.aca_rank_sum;
    s->rank_fp32 = m.cls[0].rank_sum;
    s->rank_fp16 = m.cls[1].rank_sum;
    s->dropped = m.dropped;
    s->raw_blocks = m.raw_blocks;
    s->raw_rank = m.raw_rank;
    s->bytes = m.cls[0].bytes + m.cls[1].bytes + m.cls[2].bytes + m.cls[3].bytes;
    s->u_elems[0] = m.cls[0].u_elems;
    s->u_elems[1] = m.cls[1].u_elems;
    s->v_elems[0] = m.cls[0].v_elems;
    s->v_elems[1] = m.cls[1].v_elems;
    s->build_ms = m.ms;
    s->single_jacobi = m.single_jacobi ? 1 : 0;
    s->formats = m.built_formats;
    s->err_fp8 = a2 > 0 ? std::sqrt(m.err2_fp8 / a2) : 0;
    s->rank_fp8 = m.cls[2].rank_sum;
    s->u_elems_fp8 = m.cls[2].u_elems;
    s->v_elems_fp8 = m.cls[2].v_elems;
    s->err_fp24 = a2 > 0 ? std::sqrt(m.err2_fp24 / a2) : 0;
    s->rank_fp24 = m.cls[3].rank_sum;
    s->u_elems_fp24 = m.cls[3].u_elems;
    s->v_elems_fp24 = m.cls[3].v_elems;
    s->err_fp8_model = a2 > 0 ? std::sqrt(m.err2_fp8_model / a2) : 0;
  });
}

static_assert(H3D_MIXED_FP16 == h3d::kMixFp16 && H3D_MIXED_FP8 == h3d::kMixFp8 &&
                  H3D_MIXED_FP24 == h3d::kMixFp24 && H3D_MIXED_INT8 == h3d::kMixInt8,
              "mixed-copy format bits");

int h3d_set_mixed_formats(h3d_handle* h, int formats) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (formats & ~(H3D_MIXED_FP16 | H3D_MIXED_FP8 | H3D_MIXED_FP24 | H3D_MIXED_INT8))
      throw h3d::Error(H3D_ERR_INVALID, "unknown mixed-copy format");
    h->e.mix.formats = formats;
  });
}

int h3d_autotune(h3d_handle* h, int prec, int reps, double* ms) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    h3d::autotune(h->e, prec, reps, ms);
  });
}

int h3d_kernel_timing(h3d_handle* h, int enable, double* ms, int* count) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    h3d::Engine& e = h->e;
    if (ms)
      for (int q = 0; q < 15; ++q) ms[q] = e.ktime[q];
    if (count) *count = e.kcount;
    for (int q = 0; q < 15; ++q) e.ktime[q] = 0;
    e.kcount = 0;
    e.timing = enable != 0;
  });
}

int h3d_release_precision(h3d_handle* h, int prec) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    if (prec < 0 || prec > H3D_PREC_MIXED) throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
    if (prec == H3D_PREC_MIXED) {
      h3d::release_mixed(h->e);
      return;
    }
    h3d::release_precision(h->e, prec);
  });
}

static int do_matvec(h3d_handle* h, const double* b, double* y, int nrhs, int prec, int where,
                     bool transpose) {
  return guarded([&] {
    if (!h || !b || !y) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (nrhs < 0) throw h3d::Error(H3D_ERR_INVALID, "nrhs < 0");
    if (where != H3D_HOST && where != H3D_DEVICE) throw h3d::Error(H3D_ERR_INVALID, "bad where");
    h3d::Engine& e = h->e;
    if (!e.assembled) throw h3d::Error(H3D_ERR_STATE, "matvec before assemble");
    if (nrhs == 0) return;
    const int64_t N = e.N;
    // Build what the mode needs (a factor copy, the mixed copy, scratch sized
    // for it) before x is staged: doing it inside the product would free and
    // reallocate the staging buffer that already holds x.
    h3d::prepare_mode(e, prec);
    if (where == H3D_DEVICE) {
      h3d::matvec_device(e, b, y, nrhs, prec, transpose);
      return;
    }
    const int G = h3d::block_width(e, nrhs, prec & 0xff, transpose, (prec & H3D_ACC_FP64) != 0);
    if (G > 1) {  // stage whole groups of columns
      const int Wg = e.rhs_block == 16 ? 16 : 32;
      if (e.io_w < Wg) {
        h3d::dfree(e.d_iow);
        e.d_iow = h3d::dalloc<double>(2 * N * Wg);
        e.io_w = Wg;
      }
      for (int r0 = 0; r0 < nrhs; r0 += Wg) {
        const int cnt = std::min(Wg, nrhs - r0);
        H3D_HIP(hipMemcpyAsync(e.d_iow, b + (int64_t)r0 * N, sizeof(double) * N * cnt,
                               hipMemcpyHostToDevice, e.stream));
        h3d::matvec_device(e, e.d_iow, e.d_iow + N * Wg, cnt, prec, transpose);
        H3D_HIP(hipMemcpyAsync(y + (int64_t)r0 * N, e.d_iow + N * Wg, sizeof(double) * N * cnt,
                               hipMemcpyDeviceToHost, e.stream));
      }
      H3D_HIP(hipStreamSynchronize(e.stream));
      return;
    }
    for (int r = 0; r < nrhs; ++r) {
      H3D_HIP(hipMemcpyAsync(e.d_io, b + (int64_t)r * N, sizeof(double) * N, hipMemcpyHostToDevice,
                             e.stream));
      h3d::matvec_device(e, e.d_io, e.d_io + N, 1, prec, transpose);
      H3D_HIP(hipMemcpyAsync(y + (int64_t)r * N, e.d_io + N, sizeof(double) * N,
                             hipMemcpyDeviceToHost, e.stream));
    }
    H3D_HIP(hipStreamSynchronize(e.stream));
  });
}

int h3d_matvec(h3d_handle* h, const double* b, double* y, int nrhs, int prec, int where) {
  return do_matvec(h, b, y, nrhs, prec, where, false);
}

int h3d_matvec_transpose(h3d_handle* h, const double* b, double* y, int nrhs, int prec,
                         int where) {
  return do_matvec(h, b, y, nrhs, prec, where, true);
}

int h3d_matvec_noalloc(h3d_handle* h, const double* b, double* y, int prec) {
  return guarded([&] {
    if (!h || !b || !y) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    h3d::matvec_noalloc(h->e, b, y, prec);
  });
}

int h3d_set_diag(h3d_handle* h, const double* diag_vec, int where) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    h3d::Engine& e = h->e;
    if (!diag_vec) {
      h3d::dfree(e.d_diag);
      return;
    }
    if (!e.d_diag) e.d_diag = h3d::dalloc<double>(e.N);
    H3D_HIP(hipMemcpyAsync(e.d_diag, diag_vec, sizeof(double) * e.N,
                           where == H3D_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           e.stream));
    H3D_HIP(hipStreamSynchronize(e.stream));
  });
}

int h3d_gmres(h3d_handle* h, const double* b, double* x, int maxit, double threshold,
              double* error, int* iterations, double* hist, int where) {
  return guarded([&] {
    if (!h || !b || !x || !error || !iterations || !hist)
      throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (maxit < 1) throw h3d::Error(H3D_ERR_INVALID, "maxit must be >= 1");
    h3d::Engine& e = h->e;
    if (!e.assembled) throw h3d::Error(H3D_ERR_STATE, "gmres before assemble");
    const int64_t N = e.N;
    if (where == H3D_DEVICE) {
      h3d::gmres_reference(e, b, x, maxit, threshold, error, iterations, hist);
      return;
    }
    double* db = h3d::dalloc<double>(2 * N);
    try {
      H3D_HIP(hipMemcpyAsync(db, b, sizeof(double) * N, hipMemcpyHostToDevice, e.stream));
      h3d::gmres_reference(e, db, db + N, maxit, threshold, error, iterations, hist);
      H3D_HIP(hipMemcpyAsync(x, db + N, sizeof(double) * N, hipMemcpyDeviceToHost, e.stream));
      H3D_HIP(hipStreamSynchronize(e.stream));
    } catch (...) {
      h3d::dfree(db);
      throw;
    }
    h3d::dfree(db);
  });
}

int h3d_solve_ir(h3d_handle* h, const double* b, double* x, int restart, int max_inner,
                 double inner_tol, int inner_prec, double tol, int max_outer, int transpose,
                 h3d_solve_info* info, int where) {
  int st = H3D_OK;
  int rc = guarded([&] {
    if (!h || !b || !x) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if ((inner_prec & 0xff) > H3D_PREC_MIXED || (inner_prec & ~(0xff | H3D_ACC_FP64)))
      throw h3d::Error(H3D_ERR_INVALID, "unknown precision");
    h3d::Engine& e = h->e;
    if (!e.assembled) throw h3d::Error(H3D_ERR_STATE, "solve before assemble");
    const int64_t N = e.N;
    if (where == H3D_DEVICE) {
      st = h3d::solve_ir(e, b, x, restart, max_inner, inner_tol, inner_prec, tol, max_outer,
                         transpose != 0, info);
      return;
    }
    double* db = h3d::dalloc<double>(2 * N);
    try {
      H3D_HIP(hipMemcpyAsync(db, b, sizeof(double) * N, hipMemcpyHostToDevice, e.stream));
      st = h3d::solve_ir(e, db, db + N, restart, max_inner, inner_tol, inner_prec, tol, max_outer,
                         transpose != 0, info);
      H3D_HIP(hipMemcpyAsync(x, db + N, sizeof(double) * N, hipMemcpyDeviceToHost, e.stream));
      H3D_HIP(hipStreamSynchronize(e.stream));
    } catch (...) {
      h3d::dfree(db);
      throw;
    }
    h3d::dfree(db);
  });
  if (rc == H3D_OK && st != H3D_OK) g_last_error = "refinement did not reach the tolerance";
  return rc != H3D_OK ? rc : st;
}

int h3d_direct(h3d_handle* h, const double* b, const int64_t* rows, int64_t nrows, double* out,
               int where) {
  return guarded([&] {
    if (!h || !b || (!rows && nrows > 0) || (!out && nrows > 0))
      throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (nrows < 0) throw h3d::Error(H3D_ERR_INVALID, "nrows < 0");
    h3d::Engine& e = h->e;
    if (nrows == 0) return;
    if (where == H3D_DEVICE) {
      h3d::direct_rows(e, b, rows, nrows, out);
      return;
    }
    double* db = h3d::dalloc<double>(e.N + nrows);
    try {
      H3D_HIP(hipMemcpyAsync(db, b, sizeof(double) * e.N, hipMemcpyHostToDevice, e.stream));
      h3d::direct_rows(e, db, rows, nrows, db + e.N);
      H3D_HIP(hipMemcpyAsync(out, db + e.N, sizeof(double) * nrows, hipMemcpyDeviceToHost, e.stream));
      H3D_HIP(hipStreamSynchronize(e.stream));
    } catch (...) {
      h3d::dfree(db);
      throw;
    }
    h3d::dfree(db);
  });
}

int h3d_get_matrix(h3d_handle* h, const int64_t* rows, int64_t nrows, const int64_t* cols,
                   int64_t ncols, double* out, int where) {
  return guarded([&] {
    if (!h || (nrows > 0 && ncols > 0 && (!rows || !cols || !out)))
      throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (nrows < 0 || ncols < 0) throw h3d::Error(H3D_ERR_INVALID, "negative size");
    if (where != H3D_HOST && where != H3D_DEVICE) throw h3d::Error(H3D_ERR_INVALID, "bad where");
    if (nrows == 0 || ncols == 0) return;
    if (nrows > INT64_MAX / ncols) throw h3d::Error(H3D_ERR_INVALID, "block too large");
    h3d::Engine& e = h->e;
    if (where == H3D_DEVICE) {
      h3d::get_entries(e, rows, nrows, cols, ncols, out);
      return;
    }
    h3d::DevBuf<double> d(nrows * ncols);
    h3d::get_entries(e, rows, nrows, cols, ncols, d.p);
    H3D_HIP(hipMemcpyAsync(out, d.p, sizeof(double) * nrows * ncols, hipMemcpyDeviceToHost,
                           e.stream));
    H3D_HIP(hipStreamSynchronize(e.stream));
  });
}

const char* h3d_plugin_name(void) { return h3d_user_kernel::name; }

int h3d_debug_fail_alloc(int64_t after) {
  h3d::g_alloc_countdown.store(after < 0 ? -1 : after);
  return H3D_OK;
}

int h3d_export_perm(h3d_handle* h, int32_t* perm) {
  return guarded([&] {
    if (!h || !perm) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    std::memcpy(perm, h->e.perm.data(), sizeof(int32_t) * h->e.N);
  });
}

int h3d_export_boxes(h3d_handle* h, int level, int32_t* start, int32_t* count) {
  return guarded([&] {
    if (!h || !start || !count) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    if (level < 0 || level > h->e.nL) throw h3d::Error(H3D_ERR_INVALID, "level out of range");
    const int64_t lo = h3d::level_offset(level), nb = int64_t(1) << (3 * level);
    std::memcpy(start, h->e.bstart.data() + lo, sizeof(int32_t) * nb);
    std::memcpy(count, h->e.bcount.data() + lo, sizeof(int32_t) * nb);
  });
}

int h3d_export_neighbors(h3d_handle* h, int level, int32_t* nbr26) {
  return guarded([&] {
    check_level(h, level);
    if (!nbr26) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    const h3d::Level& lv = h->e.lev[level];
    H3D_HIP(hipMemcpy(nbr26, lv.d_nbr, sizeof(int32_t) * lv.nboxes * 26, hipMemcpyDeviceToHost));
  });
}

int64_t h3d_il_size(h3d_handle* h, int level) {
  if (!h || level < 1 || level > h->e.nL) return -1;
  return (int64_t)h->e.lev[level].ilidx.size();
}

int h3d_export_il(h3d_handle* h, int level, int32_t* offsets, int32_t* boxes) {
  return guarded([&] {
    check_level(h, level);
    if (!offsets || !boxes) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    const h3d::Level& lv = h->e.lev[level];
    std::memcpy(offsets, lv.iloff.data(), sizeof(int32_t) * lv.iloff.size());
    std::memcpy(boxes, lv.ilidx.data(), sizeof(int32_t) * lv.ilidx.size());
  });
}

int h3d_export_ranks(h3d_handle* h, int level, int32_t* rank, int32_t* status) {
  return guarded([&] {
    check_level(h, level);
    if (!h->e.assembled) throw h3d::Error(H3D_ERR_STATE, "not assembled");
    const h3d::Level& lv = h->e.lev[level];
    if (rank) std::memcpy(rank, lv.rank.data(), sizeof(int32_t) * lv.rank.size());
    if (status) std::memcpy(status, lv.status.data(), sizeof(int32_t) * lv.status.size());
  });
}

int h3d_export_block(h3d_handle* h, int level, int64_t b, int32_t* I, int32_t* J, double* U,
                     double* V) {
  return guarded([&] {
    check_level(h, level);
    h3d::Engine& e = h->e;
    if (!e.assembled) throw h3d::Error(H3D_ERR_STATE, "not assembled");
    const h3d::Level& lv = e.lev[level];
    if (b < 0 || b >= (int64_t)lv.rank.size()) throw h3d::Error(H3D_ERR_INVALID, "block out of range");
    const int r = lv.rank[b];
    if (r == 0) return;
    // locate the target box of block b
    int64_t k = std::upper_bound(lv.iloff.begin(), lv.iloff.end(), (int32_t)b) - lv.iloff.begin() - 1;
    const int64_t lo = h3d::level_offset(level);
    const int32_t ki = lv.ilidx[b];
    const int n = e.bcount[lo + k], m = e.bcount[lo + ki];
    const h3d::Chunk& ch = e.chunks[lv.chunk[k]];
    if (I) H3D_HIP(hipMemcpy(I, ch.piv + lv.pivoff[b], sizeof(int32_t) * r, hipMemcpyDeviceToHost));
    if (J) H3D_HIP(hipMemcpy(J, ch.piv + lv.pivoff[b] + r, sizeof(int32_t) * r, hipMemcpyDeviceToHost));
    if ((U || V) && !ch.data[0]) throw h3d::Error(H3D_ERR_STATE, "fp64 factors not resident");
    const int ld = h3d::aca_ldu(n), ldv = m + (m & 1);
    if (U) {
      const double* src = (const double*)ch.data[0] + lv.uoff[k] + (int64_t)lv.col[b] * ld;
      H3D_HIP(hipMemcpy2D(U, sizeof(double) * n, src, sizeof(double) * ld, sizeof(double) * n, r,
                          hipMemcpyDeviceToHost));
    }
    if (V) {
      const double* src = (const double*)ch.data[0] + ch.nU + lv.voff[b];
      H3D_HIP(hipMemcpy2D(V, sizeof(double) * m, src, sizeof(double) * ldv, sizeof(double) * m, r,
                          hipMemcpyDeviceToHost));
    }
  });
}

int h3d_get_stats(h3d_handle* h, h3d_stats* s) {
  return guarded([&] {
    if (!h || !s) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    const h3d::Engine& e = h->e;
    std::memset(s, 0, sizeof(*s));
    s->N = e.N;
    s->nlevels = e.nL;
    s->max_rank = e.max_rank;
    s->n_blocks = e.n_blocks;
    s->rank_sum = e.rank_sum;
    s->factor_elems = e.factor_elems;
    s->find_memory = e.find_memory;
    s->near_pairs = e.near_pairs;
    s->t_tree_ms = e.t_tree_ms;
    s->t_assemble_ms = e.t_assemble_ms;
    for (int p = 0; p < 3; ++p) s->factor_bytes[p] = e.fac_bytes[p];
    s->u_elems = e.u_elems;
    s->v_elems = e.v_elems;
    s->ref_ass_time = e.ref_regen_ms * 1e-3;
    s->ref_matvec_time = e.ref_blas2_ms * 1e-3;
  });
}

int h3d_get_config(h3d_handle* h, h3d_config* c) {
  return guarded([&] {
    if (!h || !c) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    const h3d::Engine& e = h->e;
    std::memset(c, 0, sizeof(*c));
    c->N = e.N;
    c->nlevels = e.nL;
    c->leaf1d = e.leaf1d;
    c->L = e.L;
    for (int i = 0; i < 3; ++i) c->center[i] = e.ctr[i];
    c->tol_pow = e.tol_pow;
    c->qchoice = e.kern.q;
    c->kernel_scale = e.kern.scale;
    c->kernel_diag = e.kern.diag;
    c->alpha = e.alpha;
    c->has_diag_vec = e.d_diag ? 1 : 0;
  });
}

struct h3d_graph {
  std::vector<h3d_handle*> hs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int h3d_graph_create(h3d_handle* const* hs, int count, const double* const* b, double* const* y,
                     int prec, int transpose, h3d_graph** out) {
  return guarded([&] {
    if (!hs || !b || !y || !out || count < 1) throw h3d::Error(H3D_ERR_INVALID, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
      if (!hs[i] || !b[i] || !y[i]) throw h3d::Error(H3D_ERR_INVALID, "null handle or vector");
      h3d::prepare_mode(hs[i]->e, prec);  // copies and scratch exist before the capture
      H3D_HIP(hipStreamSynchronize(hs[i]->e.stream));
    }
    // Instances go to kBranches streams forked from the capture stream and
    // joined back, so the graph holds independent branches that can run
    // concurrently (small instances alone leave most of the GPU idle).
    constexpr int kBranches = 8;
    const int nb = std::min(count, kBranches);
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> br(nb, nullptr);
    std::vector<hipEvent_t> ev(nb + 1, nullptr);
    auto release_streams = [&] {
      for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
      for (hipStream_t x : br)
        if (x) (void)hipStreamDestroy(x);
      if (cap) (void)hipStreamDestroy(cap);
    };
    try {
      H3D_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      for (int k = 0; k < nb; ++k) H3D_HIP(hipStreamCreateWithFlags(&br[k], hipStreamNonBlocking));
      for (auto& x : ev) H3D_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    } catch (...) {
      release_streams();
      throw;
    }
    // every product of one handle goes to the same branch (a handle's scratch
    // is shared by its products, so they must stay ordered)
    std::vector<int> branch(count);
    for (int i = 0, next = 0; i < count; ++i) {
      branch[i] = -1;
      for (int j = 0; j < i; ++j)
        if (hs[j] == hs[i]) branch[i] = branch[j];
      if (branch[i] < 0) branch[i] = next++ % nb;
    }
    std::vector<hipStream_t> saved(count);
    std::vector<bool> timing(count), overlap(count);
    for (int i = 0; i < count; ++i) {
      saved[i] = hs[i]->e.stream;
      timing[i] = hs[i]->e.timing;
      overlap[i] = hs[i]->e.overlap_near;
    }
    // captured products keep every kernel on their branch (the near field in
    // line, not forked to the handle's second stream)
    for (int i = 0; i < count; ++i) hs[i]->e.timing = false, hs[i]->e.overlap_near = false;
    auto restore = [&] {
      for (int i = 0; i < count; ++i) {
        hs[i]->e.stream = saved[i];
        hs[i]->e.timing = timing[i];
        hs[i]->e.overlap_near = overlap[i];
      }
    };
    auto* g = new h3d_graph;
    try {
      H3D_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      try {
        H3D_HIP(hipEventRecord(ev[nb], cap));  // fork
        for (int k = 0; k < nb; ++k) H3D_HIP(hipStreamWaitEvent(br[k], ev[nb], 0));
        for (int i = 0; i < count; ++i) {
          hs[i]->e.stream = br[branch[i]];
          h3d::matvec_device(hs[i]->e, b[i], y[i], 1, prec, transpose != 0);
        }
        for (int k = 0; k < nb; ++k) {  // join
          H3D_HIP(hipEventRecord(ev[k], br[k]));
          H3D_HIP(hipStreamWaitEvent(cap, ev[k], 0));
        }
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cap, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      H3D_HIP(hipStreamEndCapture(cap, &g->graph));
      H3D_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      restore();
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
      release_streams();
      throw;
    }
    restore();
    release_streams();
    g->hs.assign(hs, hs + count);
    *out = g;
  });
}

int h3d_graph_launch(h3d_graph* g, void* stream) {
  return guarded([&] {
    if (!g || !g->exec) throw h3d::Error(H3D_ERR_INVALID, "null graph");
    H3D_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  });
}

void h3d_graph_destroy(h3d_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

struct h3d_batch {
  h3d::Batch b;
};

int h3d_batch_create(h3d_handle* const* hs, int count, int mode, h3d_batch** out) {
  return guarded([&] {
    if (!hs || !out || count < 1) throw h3d::Error(H3D_ERR_INVALID, "bad arguments");
    *out = nullptr;
    std::vector<h3d::Engine*> es(count);
    for (int i = 0; i < count; ++i) {
      if (!hs[i]) throw h3d::Error(H3D_ERR_INVALID, "null handle");
      es[i] = &hs[i]->e;
    }
    auto* bt = new h3d_batch;
    try {
      h3d::build_batch(bt->b, es.data(), count, mode);
    } catch (...) {
      delete bt;
      throw;
    }
    *out = bt;
  });
}

int h3d_batch_layout(h3d_batch* bt, int64_t* offsets) {
  return guarded([&] {
    if (!bt || !offsets) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    std::memcpy(offsets, bt->b.off.data(), sizeof(int64_t) * bt->b.off.size());
  });
}

int h3d_batch_matvec(h3d_batch* bt, const double* d_b, double* d_y) {
  return guarded([&] {
    if (!bt || !d_b || !d_y) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    h3d::batch_matvec(bt->b, d_b, d_y);
  });
}

int h3d_batch_set_stream(h3d_batch* bt, void* stream) {
  return guarded([&] {
    if (!bt) throw h3d::Error(H3D_ERR_INVALID, "null batch");
    hipStream_t to = (hipStream_t)stream;  // NULL: the legacy default stream
    if (to == bt->b.stream) return;
    // the batch's scratch is shared by its products: the new stream waits for
    // what the old one holds
    H3D_HIP(hipEventRecord(bt->b.ev_fork, bt->b.stream));
    H3D_HIP(hipStreamWaitEvent(to, bt->b.ev_fork, 0));
    bt->b.stream = to;
  });
}

int h3d_batch_synchronize(h3d_batch* bt) {
  return guarded([&] {
    if (!bt) throw h3d::Error(H3D_ERR_INVALID, "null batch");
    H3D_HIP(hipStreamSynchronize(bt->b.stream));
  });
}

int h3d_batch_autotune(h3d_batch* bt, int reps, double* ms) {
  return guarded([&] {
    if (!bt) throw h3d::Error(H3D_ERR_INVALID, "null batch");
    h3d::batch_autotune(bt->b, reps, ms);
    H3D_HIP(hipStreamSynchronize(bt->b.stream));
  });
}

int h3d_batch_set_tuning(h3d_batch* bt, int p1_bytes, int p2_bytes) {
  return guarded([&] {
    if (!bt) throw h3d::Error(H3D_ERR_INVALID, "null batch");
    const int es = h3d::elem_size(bt->b.prec);
    for (int v : {p1_bytes, p2_bytes}) {
      const int w = v & ~H3D_PERSISTENT;
      if ((w != 2 && w != 4 && w != 8 && w != 16) || w < es)
        throw h3d::Error(H3D_ERR_INVALID,
                         "load width must be 2, 4, 8 or 16 bytes (| H3D_PERSISTENT) and hold an element");
    }
    bt->b.tune_p1 = p1_bytes;
    bt->b.tune_p2 = p2_bytes;
  });
}

int h3d_batch_get_tuning(h3d_batch* bt, int* p1_bytes, int* p2_bytes) {
  return guarded([&] {
    if (!bt || !p1_bytes || !p2_bytes) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    *p1_bytes = bt->b.tune_p1;
    *p2_bytes = bt->b.tune_p2;
  });
}

int h3d_batch_graph_create(h3d_batch* bt, const double* d_b, double* d_y, h3d_graph** out) {
  return guarded([&] {
    if (!bt || !d_b || !d_y || !out) throw h3d::Error(H3D_ERR_INVALID, "null pointer");
    *out = nullptr;
    H3D_HIP(hipStreamSynchronize(bt->b.stream));
    // the capture runs on a stream of its own, the near field on a branch
    // forked from it (streams and events made for this capture only)
    hipStream_t cap = nullptr, br = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto release = [&] {
      for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
      if (br) (void)hipStreamDestroy(br);
      if (cap) (void)hipStreamDestroy(cap);
    };
    auto* g = new h3d_graph;
    try {
      H3D_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      H3D_HIP(hipStreamCreateWithFlags(&br, hipStreamNonBlocking));
      for (auto& x : ev) H3D_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
      H3D_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      try {
        h3d::batch_product(bt->b, d_b, d_y, cap, br, ev[0], ev[1]);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cap, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      H3D_HIP(hipStreamEndCapture(cap, &g->graph));
      H3D_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
      release();
      throw;
    }
    release();
    *out = g;
  });
}

void h3d_batch_destroy(h3d_batch* bt) {
  if (!bt) return;
  delete bt;
}

namespace {
// Moves a handle to stream `to` behind everything already queued on its
// current stream: all of a handle's products share its scratch, so a product
// on the new stream must not start before the old stream's ones have ended.
void switch_stream(h3d::Engine& e, hipStream_t to) {
  if (to == e.stream) return;
  if (!e.ev_switch) H3D_HIP(hipEventCreateWithFlags(&e.ev_switch, hipEventDisableTiming));
  H3D_HIP(hipEventRecord(e.ev_switch, e.stream));
  H3D_HIP(hipStreamWaitEvent(to, e.ev_switch, 0));
  e.stream = to;
}
}  // namespace

int h3d_set_stream(h3d_handle* h, void* stream) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    switch_stream(h->e, (hipStream_t)stream);
  });
}

int h3d_use_own_stream(h3d_handle* h) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    switch_stream(h->e, h->e.own_stream);
  });
}

int h3d_synchronize(h3d_handle* h) {
  return guarded([&] {
    if (!h) throw h3d::Error(H3D_ERR_INVALID, "null handle");
    H3D_HIP(hipStreamSynchronize(h->e.stream));
  });
}

void h3d_destroy(h3d_handle* h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->e.stream);
  delete h;
}

const char* h3d_status_string(int status) {
  switch (status) {
    case H3D_OK: return "ok";
    case H3D_ERR_INVALID: return "invalid argument";
    case H3D_ERR_OOM: return "out of memory";
    case H3D_ERR_NONFINITE: return "non-finite kernel entry";
    case H3D_ERR_HIP: return "HIP runtime error";
    case H3D_ERR_STATE: return "call out of order";
    case H3D_ERR_NOT_CONVERGED: return "not converged";
    case H3D_ERR_INTERNAL: return "internal error";
    case H3D_ERR_IO: return "file error";
    default: return "unknown status";
  }
}

const char* h3d_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
