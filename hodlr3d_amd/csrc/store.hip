// Saved assembled representation: one file holding everything that defines the
// factorised operator (points, kernel, tree permutation, per-level factor
// bookkeeping, factor chunks and ACA pivots), so the ACA is paid once.  The
// reference keeps nothing on disk; the format is this library's own:
//
//   header   "H3DFACT1", u32 version, u32 byte-order mark, then the fixed
//            fields of FileHeader below
//   points   3N doubles in the caller's original order (x0 y0 z0 x1 ...)
//   perm     N int32 (position -> original index; checked after the tree rebuild)
//   diag     N doubles when has_diag
//   levels   per level 1..nL: nblk, then rank, status, col (int32), voff,
//            pivoff (int64) per block and Rt, toff (int32), uoff (int64),
//            chunk (int32) per box
//   chunks   per chunk: nU, nV, npiv (int64), the (nU + nV) stored elements
//            of precision `prec`, then npiv int32 pivots
//   trailer  u64 checksum of everything before it (64-bit words, see Hasher)
//
// Loading rebuilds the octree from the points (a deterministic radix sort), so
// the lists and neighbour slots need not be stored, and checks every offset
// against its chunk before any of it reaches the device.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

namespace h3d {
namespace {

constexpr char kMagic[8] = {'H', '3', 'D', 'F', 'A', 'C', 'T', '1'};
constexpr uint32_t kVersion = 2;  // 2: U columns of > 256 rows on 32-element strides
constexpr uint32_t kByteOrder = 0x01020304u;
constexpr size_t kStage = size_t(64) << 20;  // host staging for factor transfers

struct FileHeader {
  int64_t N;
  int32_t nL, leaf1d;
  double L, ctr[3];
  int32_t tol_pow, q;
  double kscale, kdiag, alpha;
  int32_t has_diag, prec;  // prec: stored factor precision (0 fp64, 1 fp32)
  int64_t nblk_all;
  int32_t T, nchunks;
};

// Order-dependent 64-bit hash over 8-byte words (the tail zero-padded).
struct Hasher {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  uint64_t carry = 0;
  int nc = 0;
  void word(uint64_t w) {
    h ^= w + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
    h *= 0xFF51AFD7ED558CCDull;
  }
  void bytes(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    while (n && nc) {
      carry |= (uint64_t)*c++ << (8 * nc);
      --n;
      if (++nc == 8) {
        word(carry);
        carry = 0;
        nc = 0;
      }
    }
    for (; n >= 8; n -= 8, c += 8) {
      uint64_t w;
      std::memcpy(&w, c, 8);
      word(w);
    }
    while (n--) {
      carry |= (uint64_t)*c++ << (8 * nc);
      ++nc;
    }
  }
  uint64_t digest() const {
    Hasher t = *this;
    if (t.nc) t.word(t.carry);
    t.word(0x5bd1e995ull);
    return t.h;
  }
};

struct Writer {
  FILE* f = nullptr;
  Hasher hs;
  std::string path;
  explicit Writer(const char* p) : path(p) {
    f = std::fopen(p, "wb");
    if (!f) throw Error(H3D_ERR_IO, "cannot open " + path + " for writing");
  }
  ~Writer() {
    if (f) std::fclose(f);
  }
  void raw(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f) != n) throw Error(H3D_ERR_IO, "write failed: " + path);
    hs.bytes(p, n);
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    raw(v.data(), sizeof(T) * v.size());
  }
  template <typename T>
  void val(const T& v) {
    raw(&v, sizeof(T));
  }
  void close() {
    const uint64_t d = hs.digest();
    if (std::fwrite(&d, 8, 1, f) != 1 || std::fclose(f) != 0) {
      f = nullptr;
      throw Error(H3D_ERR_IO, "write failed: " + path);
    }
    f = nullptr;
  }
};

struct Reader {
  FILE* f = nullptr;
  Hasher hs;
  std::string path;
  int64_t left = 0;  // bytes before the trailer
  explicit Reader(const char* p) : path(p) {
    f = std::fopen(p, "rb");
    if (!f) throw Error(H3D_ERR_IO, "cannot open " + path);
    if (std::fseek(f, 0, SEEK_END) != 0) throw Error(H3D_ERR_IO, "cannot seek " + path);
    const long sz = std::ftell(f);
    std::rewind(f);
    if (sz < 8) throw Error(H3D_ERR_IO, path + ": not a saved representation");
    left = (int64_t)sz - 8;
  }
  ~Reader() {
    if (f) std::fclose(f);
  }
  void raw(void* p, size_t n) {
    if ((int64_t)n > left) throw Error(H3D_ERR_IO, path + ": truncated or corrupt");
    if (n && std::fread(p, 1, n, f) != n) throw Error(H3D_ERR_IO, "read failed: " + path);
    left -= (int64_t)n;
    hs.bytes(p, n);
  }
  template <typename T>
  void vec(std::vector<T>& v, int64_t n) {
    if (n < 0 || n > left / (int64_t)sizeof(T))
      throw Error(H3D_ERR_IO, path + ": truncated or corrupt");
    v.resize(n);
    raw(v.data(), sizeof(T) * n);
  }
  template <typename T>
  T val() {
    T v;
    raw(&v, sizeof(T));
    return v;
  }
  void finish() {
    if (left != 0) throw Error(H3D_ERR_IO, path + ": trailing data");
    uint64_t d = 0;
    if (std::fread(&d, 8, 1, f) != 1) throw Error(H3D_ERR_IO, "read failed: " + path);
    if (d != hs.digest()) throw Error(H3D_ERR_IO, path + ": checksum mismatch");
  }
};

[[noreturn]] void corrupt(const std::string& path, const char* what) {
  throw Error(H3D_ERR_IO, path + ": corrupt (" + what + ")");
}

}  // namespace

void save_engine(Engine& e, const char* path) {
  if (!e.assembled) throw Error(H3D_ERR_STATE, "not assembled");
  const int prec = e.fac_bytes[0] > 0 ? 0 : (e.fac_bytes[1] > 0 ? 1 : -1);
  if (prec < 0) throw Error(H3D_ERR_STATE, "no fp64 or fp32 factor copy to save");
  hipStream_t st = e.stream;
  H3D_HIP(hipStreamSynchronize(st));
  const int64_t N = e.N;
  // points back in the original order
  std::vector<double> px(N), py(N), pz(N), xyz(3 * N);
  H3D_HIP(hipMemcpy(px.data(), e.d_px, sizeof(double) * N, hipMemcpyDeviceToHost));
  H3D_HIP(hipMemcpy(py.data(), e.d_py, sizeof(double) * N, hipMemcpyDeviceToHost));
  H3D_HIP(hipMemcpy(pz.data(), e.d_pz, sizeof(double) * N, hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < N; ++s) {
    const int64_t o = e.perm[s];
    xyz[3 * o] = px[s];
    xyz[3 * o + 1] = py[s];
    xyz[3 * o + 2] = pz[s];
  }
  std::vector<double>().swap(px);
  std::vector<double>().swap(py);
  std::vector<double>().swap(pz);
  std::vector<double> diag;
  if (e.d_diag) {
    diag.resize(N);
    H3D_HIP(hipMemcpy(diag.data(), e.d_diag, sizeof(double) * N, hipMemcpyDeviceToHost));
  }

  Writer w(path);
  w.raw(kMagic, 8);
  w.val(kVersion);
  w.val(kByteOrder);
  FileHeader h{};
  h.N = N;
  h.nL = e.nL;
  h.leaf1d = e.leaf1d;
  h.L = e.L;
  for (int i = 0; i < 3; ++i) h.ctr[i] = e.ctr[i];
  h.tol_pow = e.tol_pow;
  h.q = e.kern.q;
  h.kscale = e.kern.scale;
  h.kdiag = e.kern.diag;
  h.alpha = e.alpha;
  h.has_diag = e.d_diag ? 1 : 0;
  h.prec = prec;
  h.nblk_all = e.nblk_all;
  h.T = e.T;
  h.nchunks = (int32_t)e.chunks.size();
  w.val(h);
  w.vec(xyz);
  std::vector<double>().swap(xyz);
  w.vec(e.perm);
  if (h.has_diag) w.vec(diag);
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    w.val((int64_t)lv.rank.size());
    w.vec(lv.rank);
    w.vec(lv.status);
    w.vec(lv.col);
    w.vec(lv.voff);
    w.vec(lv.pivoff);
    w.vec(lv.Rt);
    w.vec(lv.toff);
    w.vec(lv.uoff);
    w.vec(lv.chunk);
  }
  const size_t es = elem_size(prec);
  std::vector<char> stage(kStage);
  std::vector<int32_t> piv;
  for (const Chunk& c : e.chunks) {
    w.val(c.nU);
    w.val(c.nV);
    w.val(c.npiv);
    const size_t total = (size_t)(c.nU + c.nV) * es;
    for (size_t off = 0; off < total; off += kStage) {
      const size_t n = std::min(kStage, total - off);
      H3D_HIP(hipMemcpy(stage.data(), (const char*)c.data[prec] + off, n, hipMemcpyDeviceToHost));
      w.raw(stage.data(), n);
    }
    piv.resize(c.npiv);
    if (c.npiv)
      H3D_HIP(hipMemcpy(piv.data(), c.piv, sizeof(int32_t) * c.npiv, hipMemcpyDeviceToHost));
    w.vec(piv);
  }
  w.close();
}

void load_engine(Engine& e, const char* path, int64_t ws_budget) {
  const auto t0 = std::chrono::steady_clock::now();
  Reader r(path);
  char magic[8];
  r.raw(magic, 8);
  if (std::memcmp(magic, kMagic, 8) != 0) throw Error(H3D_ERR_IO, r.path + ": not a saved representation");
  if (r.val<uint32_t>() != kVersion) throw Error(H3D_ERR_IO, r.path + ": unsupported version");
  if (r.val<uint32_t>() != kByteOrder) throw Error(H3D_ERR_IO, r.path + ": foreign byte order");
  const FileHeader h = r.val<FileHeader>();
  if (h.N < 1 || h.N > INT32_MAX || h.nL < 1 || h.nL > 7 || h.leaf1d < 1 || !(h.L > 0) ||
      !std::isfinite(h.L) || h.q < 0 || h.q > 16 || (h.prec != 0 && h.prec != 1) ||
      h.nchunks < 0 || h.T < 0 || h.nblk_all < 0 || (h.has_diag != 0 && h.has_diag != 1))
    corrupt(r.path, "header");
  const int64_t N = h.N;
  e.N = N;
  e.nL = h.nL;
  e.leaf1d = h.leaf1d;
  e.L = h.L;
  for (int i = 0; i < 3; ++i) e.ctr[i] = h.ctr[i];
  e.ws_budget = ws_budget;
  e.store_mask = 1 << h.prec;
  e.tol_pow = h.tol_pow;
  e.tol = std::pow(10, -1.0 * h.tol_pow);
  e.kern.q = h.q;
  e.kern.scale = h.kscale;
  e.kern.diag = h.kdiag;
  e.alpha = h.alpha;
  H3D_HIP(hipStreamCreateWithFlags(&e.own_stream, hipStreamNonBlocking));
  e.stream = e.own_stream;
  hipStream_t st = e.stream;
  {
    std::vector<double> xyz;
    r.vec(xyz, 3 * N);
    for (double v : xyz)
      if (!std::isfinite(v)) corrupt(r.path, "points");
    build_tree(e, xyz.data());
  }
  {
    std::vector<int32_t> perm;
    r.vec(perm, N);
    if (perm != e.perm) throw Error(H3D_ERR_IO, r.path + ": the rebuilt tree does not match the saved permutation");
  }
  if (h.has_diag) {
    std::vector<double> diag;
    r.vec(diag, N);
    e.d_diag = dalloc<double>(N);
    H3D_HIP(hipMemcpy(e.d_diag, diag.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  }
  // per-level bookkeeping; offsets are checked once the chunk sizes are known
  e.nblk_all = 0;
  for (int j = 1; j <= e.nL; ++j) {
    Level& lv = e.lev[j];
    const int64_t nblk = r.val<int64_t>();
    if (nblk != (int64_t)lv.ilidx.size()) corrupt(r.path, "interaction lists");
    lv.bbase = e.nblk_all;
    e.nblk_all += nblk;
    r.vec(lv.rank, nblk);
    r.vec(lv.status, nblk);
    r.vec(lv.col, nblk);
    r.vec(lv.voff, nblk);
    r.vec(lv.pivoff, nblk);
    r.vec(lv.Rt, lv.nboxes);
    r.vec(lv.toff, lv.nboxes);
    r.vec(lv.uoff, lv.nboxes);
    r.vec(lv.chunk, lv.nboxes);
  }
  if (e.nblk_all != h.nblk_all) corrupt(r.path, "block count");
  e.T = h.T;
  const size_t es = elem_size(h.prec);
  std::vector<char> stage(kStage);
  e.chunks.reserve(h.nchunks);
  for (int32_t ci = 0; ci < h.nchunks; ++ci) {
    Chunk c;
    c.nU = r.val<int64_t>();
    c.nV = r.val<int64_t>();
    c.npiv = r.val<int64_t>();
    const int64_t lim = r.left / (int64_t)es;
    if (c.nU < 0 || c.nV < 0 || c.npiv < 0 || c.nU > lim || c.nV > lim - c.nU ||
        c.npiv > r.left / 4)
      corrupt(r.path, "chunk size");
    e.chunks.push_back(c);  // owned by the engine from here on (freed on any error)
    Chunk& ch = e.chunks.back();
    const size_t bytes = (size_t)(ch.nU + ch.nV + kChunkTail) * es;
    ch.data[h.prec] = dalloc<char>((int64_t)bytes);
    H3D_HIP(hipMemsetAsync(ch.data[h.prec], 0, bytes, st));
    H3D_HIP(hipStreamSynchronize(st));
    const size_t total = (size_t)(ch.nU + ch.nV) * es;
    for (size_t off = 0; off < total; off += kStage) {
      const size_t n = std::min(kStage, total - off);
      r.raw(stage.data(), n);
      H3D_HIP(hipMemcpy((char*)ch.data[h.prec] + off, stage.data(), n, hipMemcpyHostToDevice));
    }
    std::vector<int32_t> piv;
    r.vec(piv, ch.npiv);
    ch.piv = dalloc<int32_t>(std::max<int64_t>(ch.npiv, 1));
    if (ch.npiv)
      H3D_HIP(hipMemcpy(ch.piv, piv.data(), sizeof(int32_t) * ch.npiv, hipMemcpyHostToDevice));
    e.fac_bytes[h.prec] += (ch.nU + ch.nV) * (int64_t)es;
  }
  r.finish();
  // every offset inside its chunk, every t index inside T (nothing unchecked
  // reaches a kernel)
  int64_t tsum = 0;
  for (int j = 1; j <= e.nL; ++j) {
    const Level& lv = e.lev[j];
    const int64_t lo = level_offset(j);
    for (int64_t k = 0; k < lv.nboxes; ++k) {
      const int64_t n = e.bcount[lo + k];
      const int32_t c = lv.chunk[k];
      const int64_t Rt = lv.Rt[k];
      if (Rt < 0 || lv.toff[k] != tsum) corrupt(r.path, "t offsets");
      tsum += Rt;
      if (lv.iloff[k + 1] == lv.iloff[k] && Rt == 0) continue;
      if (c < 0 || c >= h.nchunks) corrupt(r.path, "chunk id");
      const Chunk& ch = e.chunks[c];
      if (lv.uoff[k] < 0 || lv.uoff[k] + (int64_t)aca_ldu(n) * Rt > ch.nU) corrupt(r.path, "U offset");
      int64_t R = 0;
      for (int32_t b = lv.iloff[k]; b < lv.iloff[k + 1]; ++b) {
        const int64_t rk = lv.rank[b], m = e.bcount[lo + lv.ilidx[b]];
        if (rk < 0 || rk > std::min(n, m) || lv.col[b] != R) corrupt(r.path, "block ranks");
        R += rk;
        if (lv.voff[b] < 0 || lv.voff[b] + rk * (m + (m & 1)) > ch.nV) corrupt(r.path, "V offset");
        if (lv.pivoff[b] < 0 || lv.pivoff[b] + 2 * rk > ch.npiv) corrupt(r.path, "pivot offset");
      }
      if (R != Rt) corrupt(r.path, "box columns");
    }
  }
  if (tsum != e.T) corrupt(r.path, "t length");
  finish_assembly(e);
  e.t_assemble_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace h3d
