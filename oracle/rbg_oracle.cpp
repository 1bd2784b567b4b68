// CPU oracle, part 1: RNG primitives, streaming ProbMinHash4, stage (a) and
// the depth features.  TEST INFRASTRUCTURE ONLY (see rbg_oracle.h).
#include "rbg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <unordered_map>
#include <vector>

#include <omp.h>

#include "../rabbitbin_amd/csrc/zig_tables.h"  // Marsaglia-Tsang tables

namespace {

// ---------------------------------------------------------------------------
// RNG primitives.  src/probmh.cpp:60-87 (wyrand), 319-338 (ziggurat),
// 344-357 (truncated exponential), src/hash_int.h:79-87 (seeded fmix64).
// ---------------------------------------------------------------------------
inline uint64_t wymix(uint64_t a, uint64_t b) {
  __uint128_t r = (__uint128_t)a * b;
  return (uint64_t)(r >> 64) ^ (uint64_t)r;
}
inline uint64_t wyrand(uint64_t& s) {
  s += 0xa0761d6478bd642fULL;
  return wymix(s, s ^ 0xe7037ed1a0b428dbULL);
}
inline double unit53(uint64_t bits) { return (double)(bits >> 11) * 0x1p-53; }
inline double wy_unit(uint64_t& s) { return unit53(wyrand(s)); }

// Lemire bounded draw on the high 32 bits (probmh.cpp:72-86).
inline uint32_t wy_bounded(uint32_t n, uint64_t& s) {
  uint64_t x = wyrand(s) >> 32;
  uint64_t prod = x * (uint64_t)n;
  uint32_t lo = (uint32_t)prod;
  if (lo < n) {
    const uint32_t thr = (uint32_t)(0u - n) % n;  // 2^32 mod n
    while (lo < thr) {
      x = wyrand(s) >> 32;
      prod = x * (uint64_t)n;
      lo = (uint32_t)prod;
    }
  }
  return (uint32_t)(prod >> 32);
}

inline uint64_t fmix64(uint64_t x, uint64_t seed) {
  x ^= seed;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// Exp(1) ziggurat.  The slow path's y, upper and lower tests are the fused
// forms GCC emits for probmh.cpp:329-333.
inline double zig_exp(uint64_t& s) {
  double shift = 0.0;
  for (;;) {
    const uint64_t bits = wyrand(s);
    const int i = (int)(bits & 0xFF);
    const double u = unit53(bits);
    const double x = u * kZigX[i];
    if (x < kZigX[i + 1]) return shift + x;
    if (i == 0) {
      shift += kZigX[1];
      continue;
    }
    const double y01 = wy_unit(s);
    const double y = std::fma(y01, kZigY[i + 1] - kZigY[i], kZigY[i]);
    const double ub = std::fma(kZigX[i] - kZigX[i + 1], y01, -(kZigX[i] - x));
    const double lb = y - std::fma(kZigX[i + 1] - x, kZigY[i + 1], kZigY[i + 1]);
    if (ub < 0 && (lb < 0 || y < std::exp(-x))) return x + shift;
  }
}

inline double ted_sample(double c1, double c2, double c3, uint64_t& s) {
  double x = wy_unit(s) * c1;
  if (x < 1.0) return x;
  for (;;) {
    x = wy_unit(s);
    if (x <= c2) return x;
    double y = wy_unit(s) * 0.5;
    if (y > 1.0 - x) {
      x = 1.0 - x;
      y = 1.0 - y;
    }
    if (x <= c3 * (1.0 - y)) return x;
    const double c1y = c1 * y;
    if (c1y <= 1.0 - x) return x;
  }
}

struct Ted {
  double boundary, gap, c1, c2, c3;
};

// TED table for m registers (probmh.cpp:506-572).
void build_ted(uint32_t m, std::vector<Ted>& t, double& fb_inv) {
  t.assign(m - 1, Ted{});
  const double fb = std::log1p(1.0 / (double)(m - 1));
  fb_inv = 1.0 / fb;
  auto coef = [](double rate, Ted& e) {
    e.c1 = (rate != 0.0) ? std::expm1(rate) / rate : 1.0;
    e.c2 = (rate != 0.0) ? -std::log1p(std::expm1(-rate) * 0.5) / rate : 0.5;
    e.c3 = (rate != 0.0) ? -std::expm1(-rate) / rate : 1.0;
  };
  t[0].boundary = 1.0;
  t[0].gap = 1.0;
  coef(fb, t[0]);
  double prev = fb;
  for (uint32_t i = 1; i < m - 1; ++i) {
    const double b = std::log1p((double)(i + 1) / (double)(m - i - 1));
    const double rate = b - prev;
    const double bn = b / fb;
    t[i].boundary = bn;
    t[i].gap = bn - t[i - 1].boundary;
    coef(rate, t[i]);
    prev = b;
  }
}

struct TedCache {
  std::mutex mu;
  std::unordered_map<uint32_t, std::pair<std::vector<Ted>, double>> map;
  const std::pair<std::vector<Ted>, double>& get(uint32_t m) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = map.find(m);
    if (it != map.end()) return it->second;
    std::pair<std::vector<Ted>, double> v;
    build_ted(m, v.first, v.second);
    return map.emplace(m, std::move(v)).first->second;
  }
};
TedCache g_ted;

// ---------------------------------------------------------------------------
// Streaming ProbMinHash4 (probmh.cpp:410-479 tracker and permutation,
// 494-504 reset, 628-686 addHash / addHashFromRng).  max_L = m.
// ---------------------------------------------------------------------------
struct Pmh {
  uint32_t m;
  uint64_t seed;
  const Ted* ted;
  double fb_inv;
  std::vector<double> tree;     // [0,m) leaves = registers, then internal
  std::vector<uint64_t> win;    // winners, 0 = unset
  std::vector<uint32_t> pval, pver;
  uint32_t pidx = 0, ver = 0;

  Pmh(uint32_t m_, uint64_t seed_) : m(m_), seed(seed_) {
    const auto& t = g_ted.get(m);
    ted = t.first.data();
    fb_inv = t.second;
    tree.assign(2 * m - 1, std::numeric_limits<double>::infinity());
    win.assign(m, 0);
    pval.assign(m, 0);
    pver.assign(m, 0);
  }
  void reset() {
    std::fill(tree.begin(), tree.end(), std::numeric_limits<double>::infinity());
    std::fill(win.begin(), win.end(), 0);
  }
  double top() const { return tree[2 * m - 2]; }
  bool update(uint32_t idx, double v) {
    if (v >= tree[idx]) return false;
    const uint32_t last = 2 * m - 2;
    for (;;) {
      tree[idx] = v;
      const uint32_t par = m + (idx >> 1);
      if (par > last) break;
      const double sib = tree[idx ^ 1u];
      if (!(sib < tree[par])) break;
      if (v < sib) v = sib;
      idx = par;
    }
    return true;
  }
  void perm_reset() {
    pidx = 0;
    if (ver == 0 || ver == UINT32_MAX) {
      for (uint32_t i = 0; i < m; ++i) {
        pval[i] = i;
        pver[i] = 0;
      }
      ver = 1;
    } else {
      ++ver;
    }
  }
  uint32_t perm_next(uint64_t& rng) {
    const uint32_t k = pidx + wy_bounded(m - pidx, rng);
    const uint32_t res = (pver[k] != ver) ? k : pval[k];
    const uint32_t cur = (pver[pidx] != ver) ? pidx : pval[pidx];
    pval[k] = cur;
    pver[k] = ver;
    ++pidx;
    return res;
  }
  void propose(uint64_t& rng, double hv) {
    const uint32_t reg = perm_next(rng);
    if (update(reg, hv)) win[reg] = rng;
  }
  void add_from_rng(uint64_t rng, double w) {
    if (!(w > 0.0)) return;
    const double winv = 1.0 / w;
    const double hv0 = unit53(rng) * ted[0].c1;
    const double first = (hv0 < 1.0) ? hv0 : ted_sample(ted[0].c1, ted[0].c2, ted[0].c3, rng);
    double hv = winv * first;
    double cur = top();
    if (!(hv < cur)) return;
    perm_reset();
    uint32_t updates = 0;
    uint32_t i = 1;
    while (hv < cur) {
      propose(rng, hv);
      cur = top();
      if (++updates >= m) break;
      if (!(winv * ted[i - 1].boundary < cur)) break;
      if (i < m - 1) {
        const Ted& tp = ted[i];
        hv = winv * std::fma(tp.gap, ted_sample(tp.c1, tp.c2, tp.c3, rng), ted[i - 1].boundary);
      } else {
        hv = winv * std::fma(fb_inv, zig_exp(rng), ted[m - 2].boundary);
        if (updates < m && hv < cur) propose(rng, hv);
        break;
      }
      ++i;
    }
  }
  void add_hash(uint64_t h, double w) {
    if (!(w > 0.0)) return;
    add_from_rng(fmix64(h, seed), w);
  }
};

// 2-bit base code: A/a 0, C/c 1, G/g 2, T/t 3, else 255
// (src/rabbit_sketch.h:56-72).
inline uint8_t base_code(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 255;
  }
}

}  // namespace

extern "C" {

uint64_t orc_fmix64(uint64_t x, uint64_t seed) { return fmix64(x, seed); }
void orc_wyrand_seq(uint64_t* state, int n, uint64_t* out) {
  for (int i = 0; i < n; ++i) out[i] = wyrand(*state);
}
void orc_wy_uniform_int_seq(uint64_t* state, uint32_t bound, int n, uint32_t* out) {
  for (int i = 0; i < n; ++i) out[i] = wy_bounded(bound, *state);
}
void orc_zig_seq(uint64_t* state, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = zig_exp(*state);
}
void orc_ted_seq(double c1, double c2, double c3, uint64_t* state, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = ted_sample(c1, c2, c3, *state);
}
void orc_ted_params(uint32_t m, double* out, double* fbinv) {
  const auto& t = g_ted.get(m);
  for (uint32_t i = 0; i + 1 < m; ++i) {
    out[5 * i + 0] = t.first[i].boundary;
    out[5 * i + 1] = t.first[i].gap;
    out[5 * i + 2] = t.first[i].c1;
    out[5 * i + 3] = t.first[i].c2;
    out[5 * i + 4] = t.first[i].c3;
  }
  *fbinv = t.second;
}

void* orc_pmh_new(uint32_t m, uint64_t seed) { return new Pmh(m, seed); }
void orc_pmh_free(void* p) { delete static_cast<Pmh*>(p); }
void orc_pmh_reset(void* p) { static_cast<Pmh*>(p)->reset(); }
void orc_pmh_add_hash(void* p, uint64_t h, double w) { static_cast<Pmh*>(p)->add_hash(h, w); }
void orc_pmh_get(void* p, uint64_t* winners, double* regs) {
  Pmh* s = static_cast<Pmh*>(p);
  for (uint32_t r = 0; r < s->m; ++r) {
    if (winners) winners[r] = s->win[r];
    if (regs) regs[r] = s->tree[r];
  }
}

// Canonical 4-mer histogram: a non-ACGT byte resets the window
// (src/rabbitbin.cpp:1467-1484).
void orc_kmer4_hist(const char* seq, size_t len, uint32_t cnt[256],
                    uint64_t base_cnt[4], uint64_t* total) {
  std::memset(cnt, 0, 256 * sizeof(uint32_t));
  base_cnt[0] = base_cnt[1] = base_cnt[2] = base_cnt[3] = 0;
  uint64_t tot = 0;
  uint32_t fwd = 0, rev = 0;
  int valid = 0;
  for (size_t i = 0; i < len; ++i) {
    const uint8_t b = base_code(seq[i]);
    if (b == 255) {
      valid = 0;
      fwd = rev = 0;
      continue;
    }
    base_cnt[b]++;
    fwd = ((fwd << 2) | b) & 0xFFu;
    rev = (rev >> 2) | ((uint32_t)(3u ^ b) << 6);
    if (++valid >= 4) {
      ++cnt[fwd <= rev ? fwd : rev];
      ++tot;
    }
  }
  *total = tot;
}

// Per-base independence (GC) normalisation, cap 20 (src/rabbitbin.cpp:1491-1516).
// The expected frequency p_fwd + p_rev is a plain add of two loop-carried
// products (no fused form in the reference build).
void orc_kmer4_weights(const uint32_t cnt[256], const uint64_t base_cnt[4],
                       uint64_t total, double cap, double w[256]) {
  for (int v = 0; v < 256; ++v) w[v] = 0.0;
  if (!total) return;
  const uint64_t vb = base_cnt[0] + base_cnt[1] + base_cnt[2] + base_cnt[3];
  double p[4];
  for (int b = 0; b < 4; ++b) p[b] = vb > 0 ? (double)base_cnt[b] / (double)vb : 0.25;
  const double inv = 1.0 / (double)total;
  for (int v = 0; v < 256; ++v) {
    if (!cnt[v]) continue;
    double pf = 1.0, pr = 1.0;
    uint32_t code = (uint32_t)v;
    for (int pos = 0; pos < 4; ++pos) {
      const uint32_t b = code & 3u;
      pf *= p[b];
      pr *= p[3u - b];
      code >>= 2;
    }
    const double obs = (double)cnt[v] * inv;
    const double ex = pf + pr;
    w[v] = (ex > 1e-14) ? std::min(obs / ex, cap) : 0.0;
  }
}

// build_pmh_winners, k = 4 (src/rabbitbin.cpp:1436-1589): elements inserted
// in ascending canonical code, winners folded 64 -> 32 bits.
void orc_build_pmh_winners(const char* seq, size_t len, uint32_t m,
                           uint64_t seed, uint32_t* out32) {
  thread_local Pmh* tl = nullptr;
  thread_local uint32_t tl_m = 0;
  thread_local uint64_t tl_seed = 0;
  if (!tl || tl_m != m || tl_seed != seed) {
    delete tl;
    tl = new Pmh(m, seed);
    tl_m = m;
    tl_seed = seed;
  } else {
    tl->reset();
  }
  if (len >= 4) {
    uint32_t cnt[256];
    uint64_t bc[4], total;
    orc_kmer4_hist(seq, len, cnt, bc, &total);
    if (total) {
      double w[256];
      orc_kmer4_weights(cnt, bc, total, 20.0, w);
      for (int v = 0; v < 256; ++v)
        if (w[v] > 0.0) tl->add_hash((uint64_t)v, w[v]);
    }
  }
  for (uint32_t r = 0; r < m; ++r) {
    const uint64_t h = tl->win[r];
    out32[r] = (uint32_t)(h ^ (h >> 32));
  }
}

void orc_sketch_batch(const char* seqs, const uint64_t* offsets, size_t n,
                      uint32_t m, uint64_t seed, uint32_t* out32, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
  for (size_t r = 0; r < n; ++r)
    orc_build_pmh_winners(seqs + offsets[r], (size_t)(offsets[r + 1] - offsets[r]),
                          m, seed, out32 + r * (size_t)m);
}

// fold_win16 (src/rabbitbin.cpp:249-253).
void orc_fold16(const uint32_t* in, size_t n, uint16_t* out) {
  for (size_t t = 0; t < n; ++t) {
    const uint32_t w = in[t];
    if (w == 0) {
      out[t] = 0;
      continue;
    }
    const uint16_t h = (uint16_t)(w ^ (w >> 16));
    out[t] = h ? h : (uint16_t)1;
  }
}

// Depth features.  Average ranks per row (src/ranker.h:55-66, 128-133,
// applied in place at src/rabbitbin.cpp:4867-4874); raw snapshot and column
// normalisation 1/mean (src/rabbitbin.cpp:4792-4806); centred, L2-normalised
// rank rows (src/rabbitbin.cpp:4972-4988, where ss += d*d is fused);
// anynz after the rank transform (src/impl/rb_abundance.cpp:94-104).
void orc_depth_features(const float* depth, size_t n, size_t S, double min_cv,
                        float* ranked, float* unit, float* raw,
                        double* colnorm, uint8_t* anynz) {
  std::vector<float> rk((size_t)n * S);
  std::memcpy(rk.data(), depth, rk.size() * sizeof(float));
  if (S > 1) {
    if (raw) std::memcpy(raw, depth, rk.size() * sizeof(float));
    if (colnorm) {
      for (size_t i = 0; i < S; ++i) {
        double s = 0.0;
        for (size_t r = 0; r < n; ++r) s += depth[r * S + i];
        const double mean = s / (double)std::max<size_t>(n, 1);
        colnorm[i] = (mean > 1e-12) ? 1.0 / mean : 0.0;
      }
    }
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < n; ++r) {
      float* row = rk.data() + r * S;
      std::vector<uint32_t> ord(S);
      std::iota(ord.begin(), ord.end(), 0u);
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return row[a] < row[b]; });
      std::vector<float> out(S);
      for (size_t c = 0, reps; c < S; c += reps) {
        reps = 1;
        while (c + reps < S && row[ord[c]] == row[ord[c + reps]]) ++reps;
        for (size_t k = 0; k < reps; ++k)
          out[ord[c + k]] = (float)(2 * c + reps - 1) / 2 + 1;
      }
      std::memcpy(row, out.data(), S * sizeof(float));
    }
  }
  if (ranked) std::memcpy(ranked, rk.data(), rk.size() * sizeof(float));
  if (unit) {
    std::memset(unit, 0, rk.size() * sizeof(float));
    if (S > 1) {
#pragma omp parallel for schedule(static)
      for (size_t r = 0; r < n; ++r) {
        const float* row = rk.data() + r * S;
        double mean = 0.0;
        for (size_t k = 0; k < S; ++k) mean += row[k];
        mean /= (double)S;
        double ss = 0.0;
        for (size_t k = 0; k < S; ++k) {
          const double d = (double)row[k] - mean;
          ss = std::fma(d, d, ss);
        }
        if (ss > 0.0) {
          const double inv = 1.0 / std::sqrt(ss);
          for (size_t k = 0; k < S; ++k)
            unit[r * S + k] = (float)(((double)row[k] - mean) * inv);
        }
      }
    }
  }
  if (anynz) {
    for (size_t r = 0; r < n; ++r) {
      uint8_t any = 0;
      for (size_t i = 0; i < S; ++i)
        if ((double)rk[r * S + i] > min_cv) {
          any = 1;
          break;
        }
      anynz[r] = any;
    }
  }
}

void orc_rand_seq(unsigned seed, int n, int32_t* out) {
  struct random_data rd;
  std::memset(&rd, 0, sizeof(rd));
  char state[128];
  initstate_r(seed, state, sizeof(state), &rd);
  for (int i = 0; i < n; ++i) random_r(&rd, &out[i]);
}

}  // extern "C"
