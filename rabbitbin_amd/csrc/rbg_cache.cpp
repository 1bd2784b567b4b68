// librbg: RabbitBin's graph cache file (RBCACHE1), read and written on the
// host.  The layout is the reference's (src/impl/rb_cache.cpp:94-175 writes,
// :178-322 reads), so a graph built here can be handed to `rabbitbin bin
// --load-cache`, and a cache the reference wrote can be read here:
//
//   "RBCACHE1"  u32 version (3)  u32 sizeof(StoredDistance) = 4  u32 sizeof(size_t) = 8
//   u64 nobs, nobs1, num_depth_samples   u64 seed, totalSize, totalSize1
//   f64 simCutoff, b0   u64 minContig, min_small_contig   u8 has_depth
//   strvec contig_names, small_contig_names    (u64 count, then per string u32 length + bytes)
//   vec<u64> seq_lens, small_seq_lens          (u64 count, then the elements)
//   mat depth_matrix (ranked), small_depth_matrix   (u64 rows, u64 cols, f32 row-major)
//   vec<f32> g_large_means, g_depth_raw   vec<f64> g_depth_colnorm
//   vec<u32> from, to (version >= 3; vec<u64> before)   vec<f32> sComp
//   u8 has_snv [vec<f32> pi, vec<u32> nsnv, vec<u32> cov]   (version >= 2)
//
// All little-endian, as the reference's fwrite of its in-memory values on x86.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/rbg.h"

namespace {

const char kMagic[8] = {'R', 'B', 'C', 'A', 'C', 'H', 'E', '1'};
constexpr uint32_t kVersion = 3;

struct Writer {
  FILE* f;
  bool ok = true;
  void raw(const void* p, size_t bytes) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) ok = false;
  }
  template <class T>
  void pod(const T& v) {
    raw(&v, sizeof(T));
  }
  template <class T>
  void vec(const T* p, uint64_t n) {
    pod(n);
    if (n) raw(p, n * sizeof(T));
  }
  void strvec(const char* bytes, const uint64_t* off, uint64_t n) {
    pod(n);
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t len = (uint32_t)(off[i + 1] - off[i]);
      pod(len);
      raw(bytes + off[i], len);
    }
  }
  void mat(const float* p, uint64_t r, uint64_t c) {
    pod(r);
    pod(c);
    if (r && c) raw(p, r * c * sizeof(float));
  }
};

struct Reader {
  FILE* f;
  bool ok = true;
  void raw(void* p, size_t bytes) {
    if (ok && bytes && std::fread(p, 1, bytes, f) != bytes) ok = false;
  }
  template <class T>
  void pod(T& v) {
    raw(&v, sizeof(T));
  }
  // A length followed by its elements, into malloc'd storage; lengths are
  // checked against what the rest of the file can hold.
  template <class T>
  T* vec(uint64_t& n, uint64_t left) {
    n = 0;
    pod(n);
    if (!ok) return nullptr;
    if (n > left / sizeof(T)) {
      ok = false;
      return nullptr;
    }
    T* p = (T*)std::malloc(std::max<uint64_t>(n, 1) * sizeof(T));
    if (!p) {
      ok = false;
      return nullptr;
    }
    raw(p, n * sizeof(T));
    return p;
  }
  void strvec(char*& bytes, uint64_t*& off, uint64_t& n, uint64_t left) {
    n = 0;
    pod(n);
    if (!ok || n > left / 4) {
      ok = false;
      return;
    }
    std::vector<char> buf;
    off = (uint64_t*)std::malloc((n + 1) * sizeof(uint64_t));
    if (!off) {
      ok = false;
      return;
    }
    off[0] = 0;
    for (uint64_t i = 0; i < n && ok; ++i) {
      uint32_t len = 0;
      pod(len);
      if (!ok || len > left) {
        ok = false;
        break;
      }
      const size_t at = buf.size();
      buf.resize(at + len);
      raw(buf.data() + at, len);
      off[i + 1] = buf.size();
    }
    bytes = (char*)std::malloc(std::max<size_t>(buf.size(), 1));
    if (!bytes) {
      ok = false;
      return;
    }
    if (!buf.empty()) std::memcpy(bytes, buf.data(), buf.size());
  }
  float* mat(uint64_t& r, uint64_t& c, uint64_t left) {
    r = c = 0;
    pod(r);
    pod(c);
    if (!ok) return nullptr;
    if (r && c > left / 4 / r) {
      ok = false;
      return nullptr;
    }
    float* p = (float*)std::malloc(std::max<uint64_t>(r * c, 1) * sizeof(float));
    if (!p) {
      ok = false;
      return nullptr;
    }
    raw(p, r * c * sizeof(float));
    return p;
  }
};

uint64_t file_left(FILE* f) {
  const long at = std::ftell(f);
  if (at < 0 || std::fseek(f, 0, SEEK_END) != 0) return 0;
  const long end = std::ftell(f);
  std::fseek(f, at, SEEK_SET);
  return end > at ? (uint64_t)(end - at) : 0;
}

}  // namespace

extern "C" {

rbg_status rbg_cache_write(const char* path, const rbg_cache* c) {
  if (!path || !c) return RBG_ERR_ARG;
  if ((c->nobs && !c->name_off && c->names) || (c->n_edges && (!c->from || !c->to || !c->scomp)))
    return RBG_ERR_ARG;
  for (uint64_t e = 0; e < c->n_edges; ++e)
    if (c->from[e] >= c->nobs || c->to[e] >= c->nobs) return RBG_ERR_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return RBG_ERR_ARG;
  Writer w{f};
  w.raw(kMagic, 8);
  w.pod(kVersion);
  w.pod((uint32_t)sizeof(float));     // StoredDistance
  w.pod((uint32_t)sizeof(uint64_t));  // size_t
  w.pod(c->nobs);
  w.pod(c->nobs1);
  w.pod(c->samples);
  w.pod(c->seed);
  w.pod(c->total_size);
  w.pod(c->total_size1);
  w.pod(c->sim_cutoff);
  w.pod(c->b0);
  w.pod(c->min_contig);
  w.pod(c->min_small_contig);
  w.pod((uint8_t)(c->has_depth ? 1 : 0));
  w.strvec(c->names, c->name_off, c->names ? c->nobs : 0);
  w.strvec(c->small_names, c->small_name_off, c->small_names ? c->nobs1 : 0);
  w.vec(c->seq_lens, c->seq_lens ? c->nobs : 0);
  w.vec(c->small_seq_lens, c->small_seq_lens ? c->nobs1 : 0);
  w.mat(c->depth, c->depth ? c->depth_rows : 0, c->depth ? c->depth_cols : 0);
  w.mat(c->small_depth, c->small_depth ? c->small_depth_rows : 0, c->small_depth ? c->small_depth_cols : 0);
  w.vec(c->large_means, c->large_means ? c->n_large_means : 0);
  w.vec(c->depth_raw, c->depth_raw ? c->n_depth_raw : 0);
  w.vec(c->colnorm, c->colnorm ? c->n_colnorm : 0);
  w.vec(c->from, c->n_edges);
  w.vec(c->to, c->n_edges);
  w.vec(c->scomp, c->n_edges);
  const uint8_t snv = c->has_snv && c->n_snv ? 1 : 0;
  w.pod(snv);
  if (snv) {
    w.vec(c->snv_pi, c->n_snv);
    w.vec(c->snv_nsnv, c->n_snv);
    w.vec(c->snv_cov, c->n_snv);
  }
  const bool ok = w.ok && std::ferror(f) == 0;
  return std::fclose(f) == 0 && ok ? RBG_OK : RBG_ERR_ARG;
}

void rbg_cache_free(rbg_cache* c) {
  if (!c) return;
  void* ps[] = {c->names, c->name_off, c->small_names, c->small_name_off, c->seq_lens, c->small_seq_lens,
                c->depth, c->small_depth, c->large_means, c->depth_raw, c->colnorm, c->from, c->to,
                c->scomp, c->snv_pi, c->snv_nsnv, c->snv_cov};
  for (void* p : ps) std::free(p);
  std::memset(c, 0, sizeof(*c));
}

rbg_status rbg_cache_read(const char* path, rbg_cache* c) {
  if (!path || !c) return RBG_ERR_ARG;
  std::memset(c, 0, sizeof(*c));
  FILE* f = std::fopen(path, "rb");
  if (!f) return RBG_ERR_ARG;
  Reader r{f};
  char magic[8];
  r.raw(magic, 8);
  uint32_t ver = 0, sz_sd = 0, sz_st = 0;
  r.pod(ver);
  r.pod(sz_sd);
  r.pod(sz_st);
  if (!r.ok || std::memcmp(magic, kMagic, 8) != 0 || ver < 1 || ver > kVersion || sz_sd != 4 || sz_st != 8) {
    std::fclose(f);
    return RBG_ERR_UNSUPPORTED;
  }
  c->version = ver;
  uint8_t hd = 0;
  r.pod(c->nobs);
  r.pod(c->nobs1);
  r.pod(c->samples);
  r.pod(c->seed);
  r.pod(c->total_size);
  r.pod(c->total_size1);
  r.pod(c->sim_cutoff);
  r.pod(c->b0);
  r.pod(c->min_contig);
  r.pod(c->min_small_contig);
  r.pod(hd);
  c->has_depth = hd != 0;
  if (r.ok && c->nobs > std::numeric_limits<uint32_t>::max()) r.ok = false;  // 32-bit node ids
  uint64_t left = file_left(f), k = 0;
  uint64_t nn = 0, ns = 0;
  r.strvec(c->names, c->name_off, nn, left);
  r.strvec(c->small_names, c->small_name_off, ns, left);
  if (r.ok && ((nn && nn != c->nobs) || (ns && ns != c->nobs1))) r.ok = false;
  if (!nn) {
    std::free(c->names);
    std::free(c->name_off);
    c->names = nullptr;
    c->name_off = nullptr;
  }
  if (!ns) {
    std::free(c->small_names);
    std::free(c->small_name_off);
    c->small_names = nullptr;
    c->small_name_off = nullptr;
  }
  c->seq_lens = r.vec<uint64_t>(k, left);
  if (r.ok && k != c->nobs) r.ok = k == 0;
  if (k == 0) {
    std::free(c->seq_lens);
    c->seq_lens = nullptr;
  }
  c->small_seq_lens = r.vec<uint64_t>(k, left);
  if (r.ok && k != c->nobs1) r.ok = k == 0;
  if (k == 0) {
    std::free(c->small_seq_lens);
    c->small_seq_lens = nullptr;
  }
  c->depth = r.mat(c->depth_rows, c->depth_cols, left);
  c->small_depth = r.mat(c->small_depth_rows, c->small_depth_cols, left);
  c->large_means = r.vec<float>(c->n_large_means, left);
  c->depth_raw = r.vec<float>(c->n_depth_raw, left);
  c->colnorm = r.vec<double>(c->n_colnorm, left);
  uint64_t nf = 0, nt = 0, nsc = 0;
  if (ver >= 3) {
    c->from = r.vec<uint32_t>(nf, left);
    c->to = r.vec<uint32_t>(nt, left);
  } else {  // size_t endpoints, range-checked into 32-bit ids
    uint64_t* lf = r.vec<uint64_t>(nf, left);
    uint64_t* lt = r.vec<uint64_t>(nt, left);
    if (r.ok) {
      c->from = (uint32_t*)std::malloc(std::max<uint64_t>(nf, 1) * 4);
      c->to = (uint32_t*)std::malloc(std::max<uint64_t>(nt, 1) * 4);
      if (!c->from || !c->to) r.ok = false;
      for (uint64_t e = 0; r.ok && e < nf; ++e) {
        if (lf[e] > std::numeric_limits<uint32_t>::max()) r.ok = false;
        c->from[e] = (uint32_t)lf[e];
      }
      for (uint64_t e = 0; r.ok && e < nt; ++e) {
        if (lt[e] > std::numeric_limits<uint32_t>::max()) r.ok = false;
        c->to[e] = (uint32_t)lt[e];
      }
    }
    std::free(lf);
    std::free(lt);
  }
  c->scomp = r.vec<float>(nsc, left);
  if (r.ok && (nf != nt || nf != nsc)) r.ok = false;
  c->n_edges = nf;
  for (uint64_t e = 0; r.ok && e < nf; ++e)
    if (c->from[e] >= c->nobs || c->to[e] >= c->nobs) r.ok = false;
  if (r.ok && ver >= 2) {
    uint8_t snv = 0;
    r.pod(snv);
    if (r.ok && snv) {
      uint64_t a = 0, b = 0, d = 0;
      c->snv_pi = r.vec<float>(a, left);
      c->snv_nsnv = r.vec<uint32_t>(b, left);
      c->snv_cov = r.vec<uint32_t>(d, left);
      c->n_snv = a;
      c->has_snv = r.ok && a == b && a == d;
      if (!c->has_snv) r.ok = false;
    }
  }
  std::fclose(f);
  if (!r.ok) {
    rbg_cache_free(c);
    return RBG_ERR_ARG;
  }
  return RBG_OK;
}

}  // extern "C"
