// Performance port of the reference's fused-regime graph pass, for timing
// the reference's CPU path on the host cores.  TEST INFRASTRUCTURE ONLY:
// bench.py's cpu_baseline times it and tests/ check that its graph equals
// orc_build_graph's; the product library never links or calls it.
//
// It follows gen_fused_calib_graph (src/impl/rb_graph.cpp:501-1356) with the
// default knobs, in the steps that decide its speed:
//  * 10 calibration samples of 2,500 contigs drawn before the pass
//    (:502-525), sample maxima kept per thread and merged after it
//    (:588-592, 1237-1242);
//  * the upper triangle cut into TILE x TILE tile pairs handed out one at a
//    time (OpenMP dynamic, 1), TILE sized from the L2 (:605-665);
//  * the abundance gate on every non-zero pair: the exact-safe int8
//    prefilter (VNNI dot, src/rabbitbin.cpp:885-898) where S >= 16, then the
//    FP32 unit dot (src/rabbitbin.cpp:754-776) (:681-834, 1088-1116);
//  * the composition early exit: the first 64 registers bound the
//    similarity from above, and a pair whose bound is below both rows'
//    thresholds stops there; sampled rows always take the full compare
//    (:836-857, 1024-1081);
//  * every evaluated pair offered to both rows' top-K heaps, each row behind
//    a test-and-test-and-set spinlock with a lock-free threshold on the
//    packed (similarity, ~id) key (:545-610, 905-976);
//  * calibration, then the mutual drain over per-row lists sorted by
//    neighbour id, emitted per thread and concatenated (:1243-1348).
// With the host lacking AVX-512 BW / VNNI the same arithmetic runs in scalar
// form (same results, slower); the stats say which ran.
#include "rbg_oracle.h"

#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <vector>

#include "rbg_oracle_calib.h"

namespace {

using Clock = std::chrono::steady_clock;
double secs_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// ---- pair kernels ----
// This file is compiled twice: with REFPORT_AVX512=1 and AVX-512 F/BW/VNNI
// enabled for the whole unit into its own library (the reference is built
// -march=native, so its kernels inline into the pass), and with
// REFPORT_AVX512=0 into liboracle.so as a scalar form of the same
// arithmetic.  The caller loads the AVX-512 library only on a host with those
// instruction sets.
#if REFPORT_AVX512
inline uint32_t match16(const uint16_t* a, const uint16_t* b, uint32_t m) {
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

inline float unit_dot(const float* a, const float* b, uint32_t S) {
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

// Signed int8 dot over `stride` bytes (a multiple of 64): the left row is
// biased by 128 into unsigned bytes for dpbusd and 128 * sum(right) is taken
// back off, which is exact.
inline int32_t q8_dot(const int8_t* a, const int8_t* b, size_t stride, int32_t bsum) {
  __m512i acc = _mm512_setzero_si512();
  const __m512i bias = _mm512_set1_epi8((char)0x80);
  for (size_t k = 0; k < stride; k += 64) {
    const __m512i av = _mm512_add_epi8(_mm512_loadu_si512((const void*)(a + k)), bias);
    acc = _mm512_dpbusd_epi32(acc, av, _mm512_loadu_si512((const void*)(b + k)));
  }
  return _mm512_reduce_add_epi32(acc) - 128 * bsum;
}
#else
inline uint32_t match16(const uint16_t* a, const uint16_t* b, uint32_t m) { return orc_match_count16(a, b, m); }
inline float unit_dot(const float* a, const float* b, uint32_t S) { return orc_unit_dot_f(a, b, S); }
inline int32_t q8_dot(const int8_t* a, const int8_t* b, size_t stride, int32_t) {
  int32_t s = 0;
  for (size_t k = 0; k < stride; ++k) s += (int32_t)a[k] * (int32_t)b[k];
  return s;
}
#endif

inline void spin_pause() { __builtin_ia32_pause(); }

// ---- row heaps ----

struct Edge32 {
  uint32_t id;
  float sv;
};
// top() is the weakest kept edge: lower sv, then larger id.
struct KeepOrder {
  bool operator()(const Edge32& a, const Edge32& b) const {
    if (a.sv != b.sv) return a.sv > b.sv;
    return a.id < b.id;
  }
};
using RowHeap = std::priority_queue<Edge32, std::vector<Edge32>, KeepOrder>;

// One cache line per row: threshold of the weakest kept edge (as a float and
// as the packed key that also orders equal similarities by id) and a lock.
struct alignas(64) RowGuard {
  std::atomic<float> thr_sv;
  std::atomic<uint64_t> thr_key;
  std::atomic<uint8_t> busy;
};

inline uint64_t pack_key(float sv, uint32_t id) {
  uint32_t b;
  std::memcpy(&b, &sv, 4);
  return ((uint64_t)b << 32) | (uint32_t)~id;
}

size_t l2_kib() {
  FILE* f = std::fopen("/sys/devices/system/cpu/cpu0/cache/index2/size", "r");
  unsigned v = 0;
  if (f) {
    if (std::fscanf(f, "%u", &v) != 1) v = 0;
    std::fclose(f);
  }
  return v;
}

}  // namespace

extern "C" {
struct OrcPortStats {
  double prep_s;     // sample draws, int8 rows, tile list
  double pass_s;     // the triangle pass
  double calib_s;    // merge of sample maxima + calibration
  double mutual_s;   // drain + mutual emit
  uint64_t tile;
  uint64_t tilepairs;
  uint64_t q8_pruned;      // pairs the int8 bound rejected
  uint64_t float_checked;  // pairs that reached the FP32 dot
  uint64_t ub_pruned;      // pairs stopped by the composition early exit
  uint64_t compares;       // pairs that took the full register compare
  int avx512;              // 1 when the AVX-512 (F/BW/VNNI) build ran
  int q8;                  // 1 when the int8 prefilter ran
};
}

namespace {

int port_fused(const OrcGraphIn* in, OrcGraphOut* out, OrcPortStats* st) {
  const size_t n = in->n;
  const uint32_t m = in->m, S = in->S, K = in->K;
  const int nth = in->threads > 0 ? in->threads : omp_get_max_threads();
  out->regime = 2;
  st->avx512 = REFPORT_AVX512;

  const auto t_prep = Clock::now();
  // Calibration samples: 10 rounds of 2,500, the lowest round wins.
  constexpr size_t R = 10, SAMP = 2500;
  orc_calib::RandStream rs(in->rand_seed);
  std::vector<int8_t> sround(n, -1);
  std::vector<uint32_t> slocal(n, 0);
  {
    std::vector<std::vector<size_t>> idxs(R, std::vector<size_t>(n));
    for (size_t r = 0; r < R; ++r) {
      std::iota(idxs[r].begin(), idxs[r].end(), 0);
      orc_calib::random_unique(idxs[r].begin(), idxs[r].end(), SAMP, rs);
    }
    for (int r = (int)R - 1; r >= 0; --r)
      for (size_t li = 0; li < SAMP; ++li) {
        sround[idxs[r][li]] = (int8_t)r;
        slocal[idxs[r][li]] = (uint32_t)li;
      }
  }
  std::vector<double> svd(m + 1);
  std::vector<float> svf(m + 1);
  for (uint32_t c = 0; c <= m; ++c) {
    svd[c] = orc_sv_from_count(c, m, in->b0);
    svf[c] = (float)svd[c];
  }

  // Abundance gate and its int8 prefilter.
  const bool gate = in->unit && S > 1 && in->w_comp < 1.0 && in->w_comp < in->min_edge_weight;
  const double corr_min = gate ? (in->min_edge_weight - in->w_comp) / (1.0 - in->w_comp) : -2.0;
  const float thr = (float)(corr_min - 1e-4);
  bool use_q8 = false;
  size_t qstride = 0;
  int32_t qmax = std::numeric_limits<int32_t>::min();
  std::vector<int8_t> q8;
  std::vector<int32_t> qsum;
  constexpr int QS = 127;
  if (REFPORT_AVX512 && gate && S >= 16 && (size_t)S <= (size_t)std::numeric_limits<int32_t>::max() / (255u * QS)) {
    qstride = ((size_t)S + 63) & ~(size_t)63;
    q8.assign(n * qstride, 0);
    qsum.assign(n, 0);
    double max_x = 0.0, max_e = 0.0;
#pragma omp parallel for num_threads(nth) schedule(static) reduction(max : max_x, max_e)
    for (size_t r = 0; r < n; ++r) {
      const float* u = in->unit + r * S;
      int8_t* q = q8.data() + r * qstride;
      int32_t s = 0;
      double xx = 0.0, ee = 0.0;
      for (uint32_t k = 0; k < S; ++k) {
        long v = std::lround((double)u[k] * QS);
        v = std::max(-(long)QS, std::min((long)QS, v));
        q[k] = (int8_t)v;
        s += (int32_t)v;
        const double x = (double)v / QS, e = (double)u[k] - x;
        xx += x * x;
        ee += e * e;
      }
      qsum[r] = s;
      max_x = std::max(max_x, std::sqrt(xx));
      max_e = std::max(max_e, std::sqrt(ee));
    }
    // |u.v - x.y| <= 2|x||e| + |e|^2 over the quantised rows, plus the FP32
    // dot's forward error over ceil(S/16) + 4 rounding steps.
    const double eps = std::numeric_limits<float>::epsilon();
    const double ops = (double)((S + 15u) / 16u + 4u);
    const double gamma = ops * eps < 0.5 ? ops * eps / (1.0 - ops * eps) : 1.0;
    const double un = max_x + max_e;
    const double slack = 2.0 * max_x * max_e + max_e * max_e + gamma * un * un + 1e-7;
    const double cut = std::ceil(((double)thr - slack) * QS * QS);
    if (cut > (double)std::numeric_limits<int32_t>::min() && cut <= (double)std::numeric_limits<int32_t>::max()) {
      qmax = (int32_t)cut - 1;
      use_q8 = true;
    }
  }
  st->q8 = use_q8 ? 1 : 0;

  // Tile size: both rows of a tile pair (winners and depth row) in 85% of L2.
  size_t tile = 10;
  {
    const size_t row = 2 * ((size_t)m * sizeof(uint16_t) + (S > 1 ? (size_t)S * sizeof(float) : 0));
    tile = (size_t)((double)l2_kib() * 1024. * 0.85) / row;
    tile = std::max((size_t)10, std::min((size_t)1024, tile));
    if (S > 1 && (size_t)S * sizeof(float) > (size_t)m * sizeof(uint16_t) / 2) tile = std::min((size_t)256, tile);
    if (tile >= 16) tile = tile / 16 * 16;
  }
  std::vector<std::pair<uint32_t, uint32_t>> tiles;
  for (size_t ii = 0; ii < n; ii += tile)
    for (size_t jj = ii; jj < n; jj += tile) tiles.emplace_back((uint32_t)ii, (uint32_t)jj);
  st->tile = tile;
  st->tilepairs = tiles.size();

  std::vector<RowHeap> heaps(n);
  std::vector<RowGuard> guard(n);
  for (size_t r = 0; r < n; ++r) {
    guard[r].thr_sv.store(std::numeric_limits<float>::lowest(), std::memory_order_relaxed);
    guard[r].thr_key.store(0, std::memory_order_relaxed);
    guard[r].busy.store(0, std::memory_order_relaxed);
  }
  std::vector<std::vector<float>> tmax((size_t)nth, std::vector<float>(R * SAMP, 0.f));
  const bool early = m >= 16;
  const uint32_t EE = m >= 64 ? 64u : m;
  st->prep_s = secs_since(t_prep);

  const auto t_pass = Clock::now();
  uint64_t n_q8 = 0, n_fl = 0, n_ub = 0, n_cmp = 0, n_gated = 0;
#pragma omp parallel num_threads(nth) reduction(+ : n_q8, n_fl, n_ub, n_cmp, n_gated)
  {
    float* my_max = tmax[(size_t)omp_get_thread_num()].data();
    auto offer = [&](size_t r, uint32_t other, float sv) {
      RowGuard& g = guard[r];
      if (sv < g.thr_sv.load(std::memory_order_relaxed)) return;
      if (pack_key(sv, other) <= g.thr_key.load(std::memory_order_relaxed)) return;
      while (g.busy.exchange(1, std::memory_order_acquire))
        while (g.busy.load(std::memory_order_relaxed)) spin_pause();
      RowHeap& h = heaps[r];
      const Edge32 c{other, sv};
      bool moved = false;
      if (h.size() < K) {
        h.push(c);
        moved = h.size() == K;
      } else if (KeepOrder()(c, h.top())) {
        h.pop();
        h.push(c);
        moved = true;
      }
      if (moved) {
        g.thr_sv.store(h.top().sv, std::memory_order_relaxed);
        g.thr_key.store(pack_key(h.top().sv, h.top().id), std::memory_order_relaxed);
      }
      g.busy.store(0, std::memory_order_release);
    };
    auto pair = [&](size_t i, size_t j, int ri) {
      ++n_gated;
      const uint16_t* a = in->win16 + i * m;
      const uint16_t* b = in->win16 + j * m;
      const int rj = sround[j];
      float sv;
      if (early && ri < 0 && rj < 0) {
        const uint32_t cp = match16(a, b, EE);
        const float up = (float)svd[cp + (m - EE)];
        const float ti = guard[i].thr_sv.load(std::memory_order_relaxed);
        const float tj = guard[j].thr_sv.load(std::memory_order_relaxed);
        if (up < (ti < tj ? ti : tj)) {
          ++n_ub;
          return;
        }
        sv = svf[cp + match16(a + EE, b + EE, m - EE)];
      } else {
        sv = svf[match16(a, b, m)];
      }
      ++n_cmp;
      offer(i, (uint32_t)j, sv);
      offer(j, (uint32_t)i, sv);
      if (ri >= 0) {
        float& x = my_max[(size_t)ri * SAMP + slocal[i]];
        if (sv > x) x = sv;
      }
      if (rj >= 0) {
        float& x = my_max[(size_t)rj * SAMP + slocal[j]];
        if (sv > x) x = sv;
      }
    };
#pragma omp for schedule(dynamic, 1)
    for (size_t p = 0; p < tiles.size(); ++p) {
      const size_t ii = tiles[p].first, jj = tiles[p].second;
      const size_t iend = std::min(ii + tile, n), jend = std::min(jj + tile, n);
      for (size_t i = ii; i < iend; ++i) {
        const int ri = sround[i];
        const size_t j0 = ii == jj ? i + 1 : jj;
        const bool nzi = !in->anynz || in->anynz[i];
        if (use_q8) {
          const int8_t* qi = q8.data() + i * qstride;
          for (size_t j = j0; j < jend; ++j) {
            if (!nzi && !in->anynz[j]) continue;
            if (q8_dot(qi, q8.data() + j * qstride, qstride, qsum[j]) <= qmax) {
              ++n_q8;
              continue;
            }
            ++n_fl;
            if (unit_dot(in->unit + i * S, in->unit + j * S, S) < thr) continue;
            pair(i, j, ri);
          }
        } else {
          for (size_t j = j0; j < jend; ++j) {
            if (!nzi && !in->anynz[j]) continue;
            if (gate) {
              ++n_fl;
              if (unit_dot(in->unit + i * S, in->unit + j * S, S) < thr) continue;
            }
            pair(i, j, ri);
          }
        }
      }
    }
  }
  st->pass_s = secs_since(t_pass);
  st->q8_pruned = n_q8;
  st->float_checked = n_fl;
  st->ub_pruned = n_ub;
  st->compares = n_cmp;
  out->gated_pairs = n_gated;

  const auto t_cal = Clock::now();
  std::vector<float> smax(R * SAMP, 0.f);
  for (size_t x = 0; x < R * SAMP; ++x)
    for (int t = 0; t < nth; ++t) smax[x] = std::max(smax[x], tmax[(size_t)t][x]);
  tmax.clear();
  double sum_p = 0;
  for (size_t r = 0; r < R; ++r) {
    std::vector<double> sub(smax.begin() + r * SAMP, smax.begin() + (r + 1) * SAMP);
    double mp = (double)orc_calib::calib_converge(sub, in->coverage, rs);
    if (mp < 51.0) mp = 50.0;
    sum_p += mp;
    if (r == 1 && sum_p / 2 < 51.0) {
      sum_p = 50.0 * (double)R;
      break;
    }
  }
  out->sim_cutoff = sum_p / (double)R;
  const float cutoff = (float)(out->sim_cutoff / 1000.);
  st->calib_s = secs_since(t_cal);

  // Mutual drain: per-row lists above the cutoff sorted by neighbour id,
  // then (i, j), i < j, kept when i is in j's list.
  const auto t_mut = Clock::now();
  std::vector<std::vector<Edge32>> nbr(n);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
  for (size_t i = 0; i < n; ++i) {
    RowHeap& h = heaps[i];
    std::vector<Edge32>& row = nbr[i];
    row.reserve(h.size());
    for (; !h.empty(); h.pop())
      if (h.top().sv > cutoff) row.push_back(h.top());
    RowHeap().swap(h);
    std::sort(row.begin(), row.end(), [](const Edge32& a, const Edge32& b) { return a.id < b.id; });
  }
  std::vector<std::vector<uint32_t>> tf((size_t)nth), tt((size_t)nth);
  std::vector<std::vector<float>> ts((size_t)nth);
#pragma omp parallel num_threads(nth)
  {
    const size_t t = (size_t)omp_get_thread_num();
#pragma omp for schedule(dynamic, 64)
    for (size_t i = 0; i < n; ++i)
      for (const Edge32& e : nbr[i]) {
        if (e.id <= i) continue;
        const std::vector<Edge32>& rj = nbr[e.id];
        auto it = std::lower_bound(rj.begin(), rj.end(), (uint32_t)i,
                                   [](const Edge32& p, uint32_t v) { return p.id < v; });
        if (it != rj.end() && it->id == (uint32_t)i) {
          tf[t].push_back((uint32_t)i);
          tt[t].push_back(e.id);
          ts[t].push_back(e.sv);
        }
      }
  }
  size_t ne = 0;
  for (int t = 0; t < nth; ++t) ne += tf[(size_t)t].size();
  out->n_edges = ne;
  out->from = (uint32_t*)std::malloc(std::max<size_t>(1, ne) * sizeof(uint32_t));
  out->to = (uint32_t*)std::malloc(std::max<size_t>(1, ne) * sizeof(uint32_t));
  out->sv = (float*)std::malloc(std::max<size_t>(1, ne) * sizeof(float));
  size_t o = 0;
  for (int t = 0; t < nth; ++t) {
    const size_t c = tf[(size_t)t].size();
    std::memcpy(out->from + o, tf[(size_t)t].data(), c * sizeof(uint32_t));
    std::memcpy(out->to + o, tt[(size_t)t].data(), c * sizeof(uint32_t));
    std::memcpy(out->sv + o, ts[(size_t)t].data(), c * sizeof(float));
    o += c;
  }
  st->mutual_s = secs_since(t_mut);
  return 0;
}

}  // namespace

extern "C" {

#if REFPORT_AVX512
// AVX-512 build (liboracle_port512.so): call only on a host with AVX-512
// F/BW/VNNI.
__attribute__((visibility("default"))) int orc_port_fused_graph_avx512(const OrcGraphIn* in, OrcGraphOut* out,
                                                                       OrcPortStats* st) {
#else
// Scalar build (liboracle.so).
int orc_port_fused_graph(const OrcGraphIn* in, OrcGraphOut* out, OrcPortStats* st) {
#endif
  // The fused regime of orc_build_graph (same inputs; in->force_regime is
  // ignored) by the reference's fused pass.  Returns 0, -1 on bad arguments,
  // -2 when n < 2500 (the 10 x 2,500-contig samples).
  std::memset(out, 0, sizeof(*out));
  std::memset(st, 0, sizeof(*st));
  if (in->m < 2 || in->K == 0) return -1;
  if (in->n < 2500) return -2;
  return port_fused(in, out, st);
}

#if !REFPORT_AVX512
// OpenMP placement the library runs under: proc_bind (omp_proc_bind_t), the
// number of places and the default thread count.
void orc_omp_info(int* proc_bind, int* places, int* max_threads) {
  *proc_bind = (int)omp_get_proc_bind();
  *places = omp_get_num_places();
  *max_threads = omp_get_max_threads();
}
#endif

}  // extern "C"
