// CPU oracle, part 2: stage (b) similarity, stage (c) top-K / calibration /
// mutual filter, stage (d) edge weights.  TEST INFRASTRUCTURE ONLY.
//
// The graph is restated per row: every row scans all partners and keeps a
// top-K heap under the reference's total order (sv desc, id asc,
// src/rabbitbin.h:418-429).  The reference's triangle pass, spin-locked
// heaps, early exits and int8 prefilter are all stated there to leave the
// kept set unchanged (src/impl/rb_graph.cpp:909-924, 981-1022, 755-834), so
// the row-wise statement yields the same graph.
#include "rbg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>
#include <random>
#include <vector>

#include <omp.h>

#include "rbg_oracle_calib.h"

namespace {

using orc_calib::RandStream;
using orc_calib::calib_converge;
using orc_calib::random_unique;

struct Cand {
  uint32_t id;
  float sv;
};
// Heap order of CompareEdge: top() is the weakest kept edge.
struct Weaker {
  bool operator()(const Cand& a, const Cand& b) const {
    if (a.sv != b.sv) return a.sv > b.sv;
    return a.id < b.id;
  }
};

inline bool is_nz(const uint8_t* anynz, size_t i, size_t j) {
  return !anynz || anynz[i] || anynz[j];
}

}  // namespace

extern "C" {

// pmh_match_count16 (src/rabbitbin.cpp:1261-1293).
uint32_t orc_match_count16(const uint16_t* a, const uint16_t* b, uint32_t m) {
  // the same count, written without branches so that the compiler vectorises it
  uint16_t c16[16] = {0};
  uint32_t c = 0, k = 0;
  for (; k + 16 <= m; k += 16)
    for (int l = 0; l < 16; ++l)
      c16[l] = (uint16_t)(c16[l] + ((a[k + l] == b[k + l]) & (a[k + l] != 0)));
  for (; k < m; ++k) c += (uint32_t)((a[k] == b[k]) & (a[k] != 0));
  for (int l = 0; l < 16; ++l) c += c16[l];
  return c;
}

// pmh_sv_from_count (src/rabbitbin.cpp:1297-1305).
double orc_sv_from_count(uint32_t cnt, uint32_t m, double b0) {
  const double inv_m = 1.0 / (double)m;
  double j = (double)cnt * inv_m;
  if (b0 > 0.0 && b0 < 1.0) {
    const double inv1mb0 = 1.0 / (1.0 - b0);
    j = (j - b0) * inv1mb0;
    if (j < 0.0) j = 0.0;
  }
  if (j > 1.0 - 1e-6) j = 1.0 - 1e-6;
  return j;
}

// estimate_pmh_baseline (src/rabbitbin.cpp:1340-1369), 16-bit winners.
double orc_pmh_baseline(const uint16_t* w16, size_t n, uint32_t m) {
  if (n < 4) return 0.0;
  const double inv_m = 1.0 / (double)m;
  std::vector<float> vals;
  vals.reserve(30000);
  std::mt19937_64 rng(0x9E3779B97F4A7C15ULL);
  std::uniform_int_distribution<size_t> uni(0, n - 1);
  for (size_t s = 0; s < 30000; ++s) {
    const size_t i = uni(rng), j = uni(rng);
    if (i == j) continue;
    vals.push_back((float)((double)orc_match_count16(w16 + i * m, w16 + j * m, m) * inv_m));
  }
  if (vals.empty()) return 0.0;
  std::nth_element(vals.begin(), vals.begin() + vals.size() / 2, vals.end());
  return (double)vals[vals.size() / 2];
}

// unit_dot_f, AVX-512 build (src/rabbitbin.cpp:754-776): 16-lane FMA
// accumulators acc0/acc1 over 32-wide chunks, one 16-wide chunk, a masked
// tail into acc0, then GCC's _mm512_reduce_add_ps tree over acc0+acc1
// (lanes l+8, then l+4, then (0,2),(1,3), then the final pair).
float orc_unit_dot_f(const float* a, const float* b, uint32_t S) {
  float acc0[16] = {0}, acc1[16] = {0};
  uint32_t k = 0;
  for (; k + 32 <= S; k += 32)
    for (int l = 0; l < 16; ++l) {
      acc0[l] = std::fmaf(a[k + l], b[k + l], acc0[l]);
      acc1[l] = std::fmaf(a[k + 16 + l], b[k + 16 + l], acc1[l]);
    }
  for (; k + 16 <= S; k += 16)
    for (int l = 0; l < 16; ++l) acc0[l] = std::fmaf(a[k + l], b[k + l], acc0[l]);
  if (k < S) {
    const uint32_t rem = S - k;
    for (uint32_t l = 0; l < 16; ++l) {
      const float x = l < rem ? a[k + l] : 0.0f;
      const float y = l < rem ? b[k + l] : 0.0f;
      acc0[l] = std::fmaf(x, y, acc0[l]);
    }
    k = S;
  }
  if (k == 0) return 0.0f;
  float v[16];
  for (int l = 0; l < 16; ++l) v[l] = acc0[l] + acc1[l];
  float t8[8], t4[4];
  for (int l = 0; l < 8; ++l) t8[l] = v[l + 8] + v[l];
  for (int l = 0; l < 4; ++l) t4[l] = t8[l + 4] + t8[l];
  const float s0 = t4[0] + t4[2];
  const float s1 = t4[1] + t4[3];
  return s0 + s1;
}

int orc_build_graph(const OrcGraphIn* in, OrcGraphOut* out) {
  std::memset(out, 0, sizeof(*out));
  const size_t n = in->n;
  const uint32_t m = in->m, S = in->S, K = in->K;
  if (m < 2 || K == 0) return -1;
  int regime = in->force_regime ? in->force_regime : (n > 25000 ? 2 : 1);
  if (regime == 2 && n < 2500) return -2;  // 10 x 2500-contig samples need n >= 2500
  out->regime = regime;
  const int threads = in->threads > 0 ? in->threads : omp_get_max_threads();

  std::vector<float> svf(m + 1);
  std::vector<double> svd(m + 1);
  for (uint32_t c = 0; c <= m; ++c) {
    svd[c] = orc_sv_from_count(c, m, in->b0);
    svf[c] = (float)svd[c];
  }
  // Abundance-first gate (src/impl/rb_graph.cpp:693-706).
  const bool gate = regime == 2 && in->unit && S > 1 && in->w_comp < 1.0 &&
                    in->w_comp < in->min_edge_weight;
  const float gate_thr =
      gate ? (float)((in->min_edge_weight - in->w_comp) / (1.0 - in->w_comp) - 1e-4) : 0.0f;

  out->topk_id = (uint32_t*)std::calloc(n * K, sizeof(uint32_t));
  out->topk_sv = (float*)std::calloc(n * K, sizeof(float));
  out->topk_len = (uint32_t*)std::calloc(n, sizeof(uint32_t));
  out->rowmax_cnt = (uint32_t*)std::calloc(n, sizeof(uint32_t));
  uint64_t gated = 0;

#pragma omp parallel for num_threads(threads) schedule(dynamic, 8) reduction(+ : gated)
  for (size_t i = 0; i < n; ++i) {
    std::priority_queue<Cand, std::vector<Cand>, Weaker> h;
    uint32_t rmax = 0;
    const uint16_t* wi = in->win16 + i * m;
    for (size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      if (regime == 2) {
        if (!is_nz(in->anynz, i, j)) continue;
        if (gate && orc_unit_dot_f(in->unit + i * S, in->unit + j * S, S) < gate_thr) continue;
      }
      const uint32_t c = orc_match_count16(wi, in->win16 + j * m, m);
      if (regime == 2 && j > i) ++gated;  // unordered pairs
      if (c > rmax) rmax = c;
      if (regime == 1 && !is_nz(in->anynz, i, j)) continue;
      const Cand cd{(uint32_t)j, svf[c]};
      if (h.size() < K) {
        h.push(cd);
      } else if (Weaker()(cd, h.top())) {
        h.pop();
        h.push(cd);
      }
    }
    out->rowmax_cnt[i] = rmax;
    const uint32_t len = (uint32_t)h.size();
    out->topk_len[i] = len;
    for (uint32_t t = len; t-- > 0;) {  // weakest first -> fill from the back
      out->topk_id[i * K + t] = h.top().id;
      out->topk_sv[i * K + t] = h.top().sv;
      h.pop();
    }
  }
  out->gated_pairs = gated;

  // Calibration.
  RandStream rs(in->rand_seed);
  double sim_cutoff;
  if (regime == 1) {
    // calibrate_sim_cutoff(coverage, full=true) (src/impl/rb_graph.cpp:1362-1384):
    // n rand() draws for the permutation, maxsim over every partner in FP64.
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    random_unique(idx.begin(), idx.end(), n, rs);
    std::vector<double> maxsim(n);
    for (size_t t = 0; t < n; ++t) maxsim[t] = n > 1 ? svd[out->rowmax_cnt[idx[t]]] : 0.0;
    sim_cutoff = (double)calib_converge(maxsim, in->coverage, rs);
  } else {
    // 10 rounds x 2500 samples, lowest round wins (src/impl/rb_graph.cpp:502-525,
    // 1243-1254).
    const size_t R = 10, SAMP = 2500;
    std::vector<std::vector<size_t>> idxs(R, std::vector<size_t>(n));
    for (size_t r = 0; r < R; ++r) {
      std::iota(idxs[r].begin(), idxs[r].end(), 0);
      random_unique(idxs[r].begin(), idxs[r].end(), SAMP, rs);
    }
    std::vector<int> sround(n, -1);
    std::vector<size_t> slocal(n, 0);
    for (int r = (int)R - 1; r >= 0; --r)
      for (size_t li = 0; li < SAMP; ++li) {
        sround[idxs[r][li]] = r;
        slocal[idxs[r][li]] = li;
      }
    std::vector<std::vector<float>> maxsim(R, std::vector<float>(SAMP, 0.f));
    for (size_t c = 0; c < n; ++c)
      if (sround[c] >= 0) {
        const float v = svf[out->rowmax_cnt[c]];
        float& ms = maxsim[sround[c]][slocal[c]];
        if (v > ms) ms = v;
      }
    double sum_p = 0;
    for (size_t r = 0; r < R; ++r) {
      std::vector<double> sub(SAMP);
      for (size_t li = 0; li < SAMP; ++li) sub[li] = (double)maxsim[r][li];
      double mp = (double)calib_converge(sub, in->coverage, rs);
      if (mp < 51.0) mp = 50.0;
      sum_p += mp;
      if (r == 1 && sum_p / 2 < 51.0) {
        sum_p = 50.0 * (double)R;
        break;
      }
    }
    sim_cutoff = sum_p / (double)R;
  }
  out->sim_cutoff = sim_cutoff;
  const double cutoff = sim_cutoff / 1000.;

  // Edge emission, ordered by (i, j).
  std::vector<uint32_t> ef, et;
  std::vector<float> es;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t len = out->topk_len[i];
    std::vector<std::pair<uint32_t, float>> row;
    for (uint32_t t = 0; t < len; ++t) {
      const uint32_t j = out->topk_id[i * K + t];
      const float sv = out->topk_sv[i * K + t];
      if (j > i && (double)sv > cutoff) row.emplace_back(j, sv);
    }
    std::sort(row.begin(), row.end());
    for (auto& [j, sv] : row) {
      if (regime == 2) {
        // mutual: i must be among j's kept neighbours (src/impl/rb_graph.cpp:1283-1348)
        bool found = false;
        const uint32_t lj = out->topk_len[j];
        for (uint32_t t = 0; t < lj; ++t)
          if (out->topk_id[(size_t)j * K + t] == i) {
            found = (double)out->topk_sv[(size_t)j * K + t] > cutoff;
            break;
          }
        if (!found) continue;
      }
      ef.push_back((uint32_t)i);
      et.push_back(j);
      es.push_back(sv);
    }
  }
  out->n_edges = ef.size();
  out->from = (uint32_t*)std::malloc(std::max<size_t>(1, ef.size()) * sizeof(uint32_t));
  out->to = (uint32_t*)std::malloc(std::max<size_t>(1, ef.size()) * sizeof(uint32_t));
  out->sv = (float*)std::malloc(std::max<size_t>(1, ef.size()) * sizeof(float));
  std::memcpy(out->from, ef.data(), ef.size() * sizeof(uint32_t));
  std::memcpy(out->to, et.data(), et.size() * sizeof(uint32_t));
  std::memcpy(out->sv, es.data(), es.size() * sizeof(float));
  return 0;
}

void orc_graph_free(OrcGraphOut* out) {
  std::free(out->from);
  std::free(out->to);
  std::free(out->sv);
  std::free(out->topk_id);
  std::free(out->topk_sv);
  std::free(out->topk_len);
  std::free(out->rowmax_cnt);
  std::memset(out, 0, sizeof(*out));
}

// depth_corr_fast, AVX-512 build (src/rabbitbin.cpp:705-746): FP32 FMA over
// 32/16-wide chunks reduced once, then an FP64 tail for the remainder.
static double corr_ref(const float* a, const float* b, uint32_t S) {
  float acc0[16] = {0}, acc1[16] = {0};
  uint32_t k = 0;
  double s = 0.0;
  for (; k + 32 <= S; k += 32)
    for (int l = 0; l < 16; ++l) {
      acc0[l] = std::fmaf(a[k + l], b[k + l], acc0[l]);
      acc1[l] = std::fmaf(a[k + 16 + l], b[k + 16 + l], acc1[l]);
    }
  for (; k + 16 <= S; k += 16)
    for (int l = 0; l < 16; ++l) acc0[l] = std::fmaf(a[k + l], b[k + l], acc0[l]);
  if (k > 0) {
    float v[16], t8[8], t4[4];
    for (int l = 0; l < 16; ++l) v[l] = acc0[l] + acc1[l];
    for (int l = 0; l < 8; ++l) t8[l] = v[l + 8] + v[l];
    for (int l = 0; l < 4; ++l) t4[l] = t8[l + 4] + t8[l];
    s = (double)((t4[0] + t4[2]) + (t4[1] + t4[3]));
  }
  for (; k < S; ++k) s += (double)a[k] * (double)b[k];
  return s;
}

static double corr_fp64(const float* a, const float* b, uint32_t S) {
  double s = 0.0;
  for (uint32_t k = 0; k < S; ++k) s += (double)a[k] * (double)b[k];
  return s;
}

// cal_depth_wjac (src/rabbitbin.cpp:685-698).
static double wjac(const float* a, const float* b, const double* colnorm, uint32_t S) {
  double mn = 0.0, mx = 0.0;
  for (uint32_t i = 0; i < S; ++i) {
    double x = a[i], y = b[i];
    if (colnorm) {
      x *= colnorm[i];
      y *= colnorm[i];
    }
    if (x < y) {
      mn += x;
      mx += y;
    } else {
      mn += y;
      mx += x;
    }
  }
  if (mx <= 0.0) return 0.0;
  return mn / mx;
}

void orc_edge_weights(const uint32_t* from, const uint32_t* to, const float* scomp,
                      uint64_t ne, const float* unit, const float* raw,
                      const double* colnorm, uint32_t S, double w_comp,
                      double min_edge_weight, double neg_thr, int mode, float* out_w,
                      double* out_w64) {
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < ne; ++e) {
    const size_t i = from[e], j = to[e];
    double w;
    if (S == 0) {
      w = scomp[e] < min_edge_weight ? 0.0 : (double)scomp[e];
    } else if (S == 1) {
      w = scomp[e];
    } else {
      // depth_edge_term, default fuse = min(corr+, wjac) (src/rabbitbin.cpp:930-970)
      const double corr = mode == 0 ? corr_ref(unit + i * S, unit + j * S, S)
                                    : corr_fp64(unit + i * S, unit + j * S, S);
      bool ok = std::isfinite(corr) && !(neg_thr > -0.99 && corr < neg_thr);
      double d = 0.0;
      if (ok) {
        const double cp = corr < 0.0 ? 0.0 : corr;
        double wj = wjac(raw + i * S, raw + j * S, colnorm, S);
        if (!std::isfinite(wj)) wj = 0.0;
        wj = wj < 0.0 ? 0.0 : (wj > 1.0 ? 1.0 : wj);
        d = cp < wj ? cp : wj;
        w = w_comp * (double)scomp[e] + (1.0 - w_comp) * d;
        if (!std::isfinite(w) || w < min_edge_weight) w = 0.0;
      } else {
        w = 0.0;
      }
    }
    out_w[e] = (float)w;
    if (out_w64) out_w64[e] = w;
  }
}

// Edges whose FP64 arithmetic lies within tol of a cut (the edges the
// reference's mixed FP32/FP64 correlation could decide differently from FP64):
// the FP64 correlation within tol of the negative cut, or the FP64 weight
// before the min-edge cut within tol of that cut.  S <= 1 weights are exact.
void orc_edge_near(const uint32_t* from, const uint32_t* to, const float* scomp, uint64_t ne,
                   const float* unit, const float* raw, const double* colnorm, uint32_t S,
                   double w_comp, double min_edge_weight, double neg_thr, double tol,
                   uint8_t* near) {
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < ne; ++e) {
    uint8_t f = 0;
    if (S > 1) {
      const size_t i = from[e], j = to[e];
      const double corr = corr_fp64(unit + i * S, unit + j * S, S);
      if (neg_thr > -0.99 && std::fabs(corr - neg_thr) <= tol) f = 1;
      if (std::isfinite(corr) && !(neg_thr > -0.99 && corr < neg_thr)) {
        const double cp = corr < 0.0 ? 0.0 : corr;
        double wj = wjac(raw + i * S, raw + j * S, colnorm, S);
        if (!std::isfinite(wj)) wj = 0.0;
        wj = wj < 0.0 ? 0.0 : (wj > 1.0 ? 1.0 : wj);
        const double w = w_comp * (double)scomp[e] + (1.0 - w_comp) * (cp < wj ? cp : wj);
        if (std::fabs(w - min_edge_weight) <= tol) f = 1;
      }
    }
    near[e] = f;
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Checks of the restatement against the hardware and the libraries.
// ---------------------------------------------------------------------------
#include <immintrin.h>

// unit_dot_f with AVX-512 intrinsics, the loop shape of src/rabbitbin.cpp:
// 754-776.  Used only to pin orc_unit_dot_f (scalar emulation of the same
// lane/accumulator order) on an AVX-512 host.
__attribute__((target("avx512f"))) static float unit_dot_avx512(const float* a, const float* b,
                                                                 uint32_t S) {
  uint32_t k = 0;
  float r = 0.0f;
  __m512 acc0 = _mm512_setzero_ps(), acc1 = _mm512_setzero_ps();
  for (; k + 32 <= S; k += 32) {
    acc0 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k), _mm512_loadu_ps(b + k), acc0);
    acc1 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k + 16), _mm512_loadu_ps(b + k + 16), acc1);
  }
  for (; k + 16 <= S; k += 16)
    acc0 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k), _mm512_loadu_ps(b + k), acc0);
  if (k < S) {
    const __mmask16 mask = (__mmask16)((1u << (S - k)) - 1u);
    acc0 = _mm512_fmadd_ps(_mm512_maskz_loadu_ps(mask, a + k), _mm512_maskz_loadu_ps(mask, b + k), acc0);
    k = S;
  }
  if (k > 0) r = _mm512_reduce_add_ps(_mm512_add_ps(acc0, acc1));
  return r;
}

// depth_corr_fast with AVX-512 intrinsics, the loop shape of
// src/rabbitbin.cpp:705-746 (FP32 chunks of 32 and 16, one reduce, FP64 tail).
// Used only to pin corr_ref on an AVX-512 host.
__attribute__((target("avx512f"))) static double depth_corr_avx512(const float* a, const float* b,
                                                                   uint32_t S) {
  uint32_t k = 0;
  double s = 0.0;
  __m512 acc0 = _mm512_setzero_ps(), acc1 = _mm512_setzero_ps();
  for (; k + 32 <= S; k += 32) {
    acc0 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k), _mm512_loadu_ps(b + k), acc0);
    acc1 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k + 16), _mm512_loadu_ps(b + k + 16), acc1);
  }
  for (; k + 16 <= S; k += 16) acc0 = _mm512_fmadd_ps(_mm512_loadu_ps(a + k), _mm512_loadu_ps(b + k), acc0);
  if (k > 0) s = (double)_mm512_reduce_add_ps(_mm512_add_ps(acc0, acc1));
  for (; k < S; ++k) s += (double)a[k] * (double)b[k];
  return s;
}

extern "C" {

int orc_depth_corr_avx512(const float* a, const float* b, uint32_t S, double* out) {
  if (!__builtin_cpu_supports("avx512f")) return -1;
  *out = depth_corr_avx512(a, b, S);
  return 0;
}

// The two correlation forms of the edge weight: mode 0 the reference's
// depth_corr_fast order, mode 1 FP64 throughout.
double orc_depth_corr(const float* a, const float* b, uint32_t S, int mode) {
  return mode == 0 ? corr_ref(a, b, S) : corr_fp64(a, b, S);
}

int orc_unit_dot_f_avx512(const float* a, const float* b, uint32_t S, float* out) {
  if (!__builtin_cpu_supports("avx512f")) return -1;
  *out = unit_dot_avx512(a, b, S);
  return 0;
}

// k-th output (1-based) of std::mt19937_64 with the default seed.
uint64_t orc_mt19937_64_nth(uint64_t k) {
  std::mt19937_64 g;
  uint64_t v = 0;
  for (uint64_t i = 0; i < k; ++i) v = g();
  return v;
}

// The (i, j) draws of estimate_pmh_baseline (src/rabbitbin.cpp:1346-1349),
// i == j kept; returns the number written (30000).
uint32_t orc_baseline_draws(uint64_t n, uint64_t* out_i, uint64_t* out_j) {
  std::mt19937_64 rng(0x9E3779B97F4A7C15ULL);
  std::uniform_int_distribution<size_t> uni(0, n - 1);
  for (uint32_t s = 0; s < 30000; ++s) {
    out_i[s] = uni(rng);
    out_j[s] = uni(rng);
  }
  return 30000;
}

// Calibration alone, from per-row maximum counts (same code path as
// orc_build_graph).
double orc_calibrate(const uint32_t* rowmax, uint64_t n, uint32_t m, double b0, double coverage,
                     unsigned rand_seed, int regime) {
  std::vector<float> svf(m + 1);
  std::vector<double> svd(m + 1);
  for (uint32_t c = 0; c <= m; ++c) {
    svd[c] = orc_sv_from_count(c, m, b0);
    svf[c] = (float)svd[c];
  }
  RandStream rs(rand_seed);
  if (regime == 1) {
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    random_unique(idx.begin(), idx.end(), n, rs);
    std::vector<double> maxsim(n);
    for (size_t t = 0; t < n; ++t) maxsim[t] = n > 1 ? svd[rowmax[idx[t]]] : 0.0;
    return (double)calib_converge(maxsim, coverage, rs);
  }
  const size_t R = 10, SAMP = 2500;
  std::vector<std::vector<size_t>> idxs(R, std::vector<size_t>(n));
  for (size_t r = 0; r < R; ++r) {
    std::iota(idxs[r].begin(), idxs[r].end(), 0);
    random_unique(idxs[r].begin(), idxs[r].end(), SAMP, rs);
  }
  std::vector<int> sround(n, -1);
  std::vector<size_t> slocal(n, 0);
  for (int r = (int)R - 1; r >= 0; --r)
    for (size_t li = 0; li < SAMP; ++li) {
      sround[idxs[r][li]] = r;
      slocal[idxs[r][li]] = li;
    }
  std::vector<std::vector<float>> maxsim(R, std::vector<float>(SAMP, 0.f));
  for (size_t c = 0; c < n; ++c)
    if (sround[c] >= 0) maxsim[sround[c]][slocal[c]] = svf[rowmax[c]];
  double sum_p = 0;
  for (size_t r = 0; r < R; ++r) {
    std::vector<double> sub(maxsim[r].begin(), maxsim[r].end());
    double mp = (double)calib_converge(sub, coverage, rs);
    if (mp < 51.0) mp = 50.0;
    sum_p += mp;
    if (r == 1 && sum_p / 2 < 51.0) {
      sum_p = 50.0 * (double)R;
      break;
    }
  }
  return sum_p / (double)R;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Timed CPU restatement of the fused pair pass for a row sample (the CPU
// baseline of bench.py): rows [row0, row1) against every partner, the gate
// and 16-bit register compare in AVX-512 (src/rabbitbin.cpp:754-776,
// 1261-1293) and a top-K heap per row.  OpenMP over rows.
// ---------------------------------------------------------------------------
__attribute__((target("avx512f,avx512bw"))) static uint32_t match16_avx512(const uint16_t* a,
                                                                           const uint16_t* b,
                                                                           uint32_t m) {
  uint32_t cnt = 0, k = 0;
  const __m512i vz = _mm512_setzero_si512();
  for (; k + 32 <= m; k += 32) {
    const __m512i va = _mm512_loadu_si512((const void*)(a + k));
    const __m512i vb = _mm512_loadu_si512((const void*)(b + k));
    cnt += (uint32_t)__builtin_popcount(
        (unsigned)(_mm512_cmpeq_epi16_mask(va, vb) & _mm512_cmpneq_epi16_mask(va, vz)));
  }
  for (; k < m; ++k) cnt += (a[k] != 0 && a[k] == b[k]) ? 1u : 0u;
  return cnt;
}

extern "C" double orc_pair_pass_rows(const uint16_t* win16, const float* unit, const uint8_t* anynz,
                                     uint64_t n, uint32_t m, uint32_t S, double b0, uint32_t K,
                                     uint64_t row0, uint64_t row1, int threads, uint64_t* pairs,
                                     uint64_t* gated_out) {
  if (threads <= 0) threads = omp_get_max_threads();
  const bool avx = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
  std::vector<float> svf(m + 1);
  for (uint32_t c = 0; c <= m; ++c) svf[c] = (float)orc_sv_from_count(c, m, b0);
  const bool gate = unit && S > 2;  // w_comp = 0 for S >= 3
  const float thr = (float)(0.7 - 1e-4);
  uint64_t gated = 0;
  const double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : gated)
  for (uint64_t i = row0; i < row1; ++i) {
    std::priority_queue<Cand, std::vector<Cand>, Weaker> h;
    const uint16_t* wi = win16 + i * m;
    for (uint64_t j = 0; j < n; ++j) {
      if (j == i || !is_nz(anynz, i, j)) continue;
      if (gate) {
        float c;
        if (avx)
          c = unit_dot_avx512(unit + i * S, unit + j * S, S);
        else
          c = orc_unit_dot_f(unit + i * S, unit + j * S, S);
        if (c < thr) continue;
      }
      ++gated;
      const uint32_t cnt = avx ? match16_avx512(wi, win16 + j * m, m)
                               : orc_match_count16(wi, win16 + j * m, m);
      const Cand cd{(uint32_t)j, svf[cnt]};
      if (h.size() < K) {
        h.push(cd);
      } else if (Weaker()(cd, h.top())) {
        h.pop();
        h.push(cd);
      }
    }
  }
  const double secs = omp_get_wtime() - t0;
  *pairs = (row1 - row0) * (n - 1);
  *gated_out = gated;
  return secs;
}

// Fused-regime top-K of selected rows against all partners
// (src/impl/rb_graph.cpp:931-975 heap order, gate of :1108-1116): for each
// row r in rows[], the gated partners j != r (is_nz, !(unit_dot_f < thr))
// ranked by (sv desc, id asc), the first K ids written to out_ids[t*K ..]
// sorted by id, their number to out_len[t], the row's maximum count over
// its gated partners to out_rowmax[t] and the number of gated partners to
// out_gated[t].  thr <= -inf disables the gate.  A size-independent check of
// the selection for sampled rows of a large assembly.
extern "C" void orc_rows_topk(const uint16_t* win16, const float* unit, const uint8_t* anynz,
                              uint64_t n, uint32_t m, uint32_t S, double b0, uint32_t K, float thr,
                              const uint32_t* rows, uint32_t nrows, uint32_t* out_ids,
                              uint32_t* out_len, uint32_t* out_rowmax, uint64_t* out_gated,
                              int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  std::vector<float> svf(m + 1);
  for (uint32_t c = 0; c <= m; ++c) svf[c] = (float)orc_sv_from_count(c, m, b0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (uint32_t t = 0; t < nrows; ++t) {
    const size_t i = rows[t];
    std::priority_queue<Cand, std::vector<Cand>, Weaker> h;
    uint32_t rmax = 0;
    uint64_t gated = 0;
    for (size_t j = 0; j < n; ++j) {
      if (j == i || !is_nz(anynz, i, j)) continue;
      if (unit && orc_unit_dot_f(unit + i * S, unit + j * S, S) < thr) continue;
      ++gated;
      const uint32_t c = orc_match_count16(win16 + i * m, win16 + j * m, m);
      if (c > rmax) rmax = c;
      const Cand cd{(uint32_t)j, svf[c]};
      if (h.size() < K) {
        h.push(cd);
      } else if (Weaker()(cd, h.top())) {
        h.pop();
        h.push(cd);
      }
    }
    std::vector<uint32_t> ids;
    while (!h.empty()) {
      ids.push_back(h.top().id);
      h.pop();
    }
    std::sort(ids.begin(), ids.end());
    std::copy(ids.begin(), ids.end(), out_ids + (size_t)t * K);
    out_len[t] = (uint32_t)ids.size();
    out_rowmax[t] = rmax;
    out_gated[t] = gated;
  }
}
