// librbg: Fisher label propagation over the weighted graph (host), the direct
// consumer of the edge weights: cluster_by_propagation with its default
// options (src/impl/rb_cluster.cpp:376-717) on the graph the reference builds
// for it (src/rabbitbin.cpp:5597-5651).
//
// Input contract (src/rabbitbin.cpp:5597-5624): edges with weight > 0 only,
// weights clamped to 1 - 1e-6 (float); every node's incident edges in
// ascending neighbour id; the node order is iota(n) shuffled by std::shuffle
// with std::default_random_engine(seed) (:5643-5645).
//
// The propagation is the reference's sequential (Gauss-Seidel) sweep: nodes
// in that order, each re-labelled from its neighbours' current labels, so
// the result depends on the order in which labels change; it runs on the host
// for that reason.  Per candidate label k of node v: m = number of v's arcs
// into k, acc = sum of log(1 - w) over them (FP64, in arc order); score = the
// chi-squared CDF with 2m degrees of freedom at -2 acc, 1 - e^-u sum_{i<m}
// u^i / i! with u = -acc (src/impl/rb_cluster.cpp:3-22, :52-66).  The best
// score wins, the first such label in arc order on exact ties; v keeps its
// label when that label's score is within 1e-12 of the best.  Only nodes
// whose neighbourhood changed since they were last visited are visited; a
// node that returns to a label it held before is frozen; the sweep repeats
// until a full pass changes nothing (:590-717).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/rbg.h"

namespace {

// 1 - e^-u sum_{i<m} u^i/i!, u = x/2, clamped at 0 (the Poisson-sum form of
// the chi-squared CDF with 2m degrees of freedom, in the reference's order).
double chi2_cdf_2m(int m, double x) {
  const double u = 0.5 * x;
  double p = std::exp(-u);
  double q = p;
  for (int i = 1; i < m; ++i) {
    p *= u / (double)i;
    q += p;
  }
  const double v = 1.0 - q;
  return v < 0.0 ? 0.0 : v;
}

constexpr double kTieEps = 1e-12;

}  // namespace

extern "C" {

rbg_status rbg_lpa_node_order(uint32_t n, uint64_t seed, uint32_t* order) {
  if (n && !order) return RBG_ERR_ARG;
  std::vector<size_t> o(n);
  std::iota(o.begin(), o.end(), 0);
  std::shuffle(o.begin(), o.end(), std::default_random_engine(seed));
  for (uint32_t i = 0; i < n; ++i) order[i] = (uint32_t)o[i];
  return RBG_OK;
}

rbg_status rbg_label_propagation(const uint32_t* from, const uint32_t* to, const float* score, uint64_t ne,
                                 uint32_t n, uint64_t seed, uint32_t* membership, uint32_t* rounds_out) {
  if (!membership || (ne && (!from || !to || !score))) return RBG_ERR_ARG;
  for (uint64_t e = 0; e < ne; ++e) {
    if (from[e] >= n || to[e] >= n || from[e] == to[e]) return RBG_ERR_ARG;
    if (!(score[e] >= 0.0f)) return RBG_ERR_ARG;  // the reference requires non-negative weights
  }
  for (uint32_t v = 0; v < n; ++v) membership[v] = v;
  if (rounds_out) *rounds_out = 0;
  // CSR of the positive edges: per node (neighbour, log(1 - w)), by neighbour
  const float wmax = 1.0f - 1e-6f;
  std::vector<uint64_t> off((size_t)n + 1, 0);
  uint64_t npos = 0;
  for (uint64_t e = 0; e < ne; ++e)
    if (score[e] > 0.0f) {
      ++off[from[e] + 1];
      ++off[to[e] + 1];
      ++npos;
    }
  if (npos == 0) return RBG_OK;  // the reference skips propagation without edges
  for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
  struct Arc {
    uint32_t nb;
    float w;
  };
  std::vector<Arc> arcs(off[n]);
  {
    std::vector<uint64_t> fill(off.begin(), off.end() - 1);
    for (uint64_t e = 0; e < ne; ++e) {
      float s = score[e];
      if (!(s > 0.0f)) continue;
      if (s > wmax) s = wmax;
      const float lw = (float)std::log(1.0 - (double)s);
      arcs[fill[from[e]]++] = Arc{to[e], lw};
      arcs[fill[to[e]]++] = Arc{from[e], lw};
    }
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t v = 0; v < (int64_t)n; ++v)
    std::stable_sort(arcs.begin() + off[v], arcs.begin() + off[v + 1],
                     [](const Arc& a, const Arc& b) { return a.nb < b.nb; });
  std::vector<uint32_t> order(n);
  rbg_lpa_node_order(n, seed, order.data());

  std::vector<double> acc(n, 0.0);
  std::vector<int32_t> cnt(n, 0);
  std::vector<uint32_t> touched;
  std::vector<uint8_t> active(n, 1), frozen(n, 0);
  std::vector<std::vector<uint32_t>> seen(n);  // labels a node has held, once it changed
  uint32_t rounds = 0;
  for (;;) {
    ++rounds;
    uint64_t changed = 0;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t v = order[i];
      if (!active[v]) continue;
      active[v] = 0;
      touched.clear();
      for (uint64_t a = off[v]; a < off[v + 1]; ++a) {
        const uint32_t k = membership[arcs[a].nb];
        if (cnt[k] == 0) touched.push_back(k);
        acc[k] += (double)arcs[a].w;
        ++cnt[k];
      }
      if (touched.empty()) continue;
      const uint32_t cur = membership[v];
      bool cur_cand = false;
      double cur_val = -std::numeric_limits<double>::infinity();
      double best_val = -std::numeric_limits<double>::infinity();
      uint32_t best = touched[0];
      for (uint32_t k : touched) {
        const double val = chi2_cdf_2m(cnt[k], -2.0 * acc[k]);
        if (k == cur) {
          cur_cand = true;
          cur_val = val;
        }
        if (val > best_val) {
          best_val = val;
          best = k;
        }
        acc[k] = 0.0;
        cnt[k] = 0;
      }
      if (cur_cand && cur_val >= best_val - kTieEps) best = cur;
      if (best == cur || frozen[v]) continue;
      membership[v] = best;
      ++changed;
      for (uint64_t a = off[v]; a < off[v + 1]; ++a) active[arcs[a].nb] = 1;
      std::vector<uint32_t>& s = seen[v];
      if (s.empty()) s.push_back(cur);
      if (std::find(s.begin(), s.end(), best) != s.end())
        frozen[v] = 1;  // back to a label it held: freeze it
      else
        s.push_back(best);
    }
    if (changed == 0) break;
  }
  if (rounds_out) *rounds_out = rounds;
  return RBG_OK;
}

}  // extern "C"
