// Oracle (test infrastructure): Fisher label propagation, restated from the
// reference for the parity tests.  Follows the reference's legacy incidence
// path literally: per-node lists of edge ids sorted by the other endpoint
// (src/rabbitbin.cpp:5597-5624), per-edge log(1 - w) (src/impl/rb_cluster.cpp:
// 436-444), the chi-squared combination (:3-22, :52-66), and the sequential
// active-set sweep with label retention and the revisit freeze (:590-717).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

double chi2_sf(int m, double x) {  // src/impl/rb_cluster.cpp:12-18
  const double u = 0.5 * x;
  double p = std::exp(-u);
  double Q = p;
  for (int i = 1; i < m; ++i) {
    p *= u / (double)i;
    Q += p;
  }
  return Q;
}
double chi2_cdf(int m, double x) {  // :19-22
  double v = 1.0 - chi2_sf(m, x);
  return v < 0.0 ? 0.0 : v;
}

}  // namespace

extern "C" {

double orc_chi2_cdf(int m, double x) { return chi2_cdf(m, x); }

void orc_lpa_node_order(uint64_t n, uint64_t seed, uint64_t* out) {
  std::vector<size_t> node_order(n);
  std::iota(node_order.begin(), node_order.end(), 0);
  std::shuffle(node_order.begin(), node_order.end(), std::default_random_engine(seed));
  for (uint64_t i = 0; i < n; ++i) out[i] = node_order[i];
}

// membership: n labels; frozen_out (may be NULL): 1 for the nodes frozen by
// the revisit rule.  Returns the number of rounds (0 without edges).
int orc_label_propagation(const uint32_t* from, const uint32_t* to, const float* score_in, uint64_t ne,
                          uint64_t n, uint64_t seed, uint64_t* membership_out, uint8_t* frozen_out) {
  if (frozen_out)
    for (uint64_t i = 0; i < n; ++i) frozen_out[i] = 0;
  std::vector<float> edgeScore(score_in, score_in + ne);
  std::vector<size_t> membership(n);
  std::iota(membership.begin(), membership.end(), 0);
  for (uint64_t i = 0; i < n; ++i) membership_out[i] = i;
  if (n == 0 || ne == 0) return 0;
  // incidence lists (src/rabbitbin.cpp:5600-5624)
  const float SSCR_MAX = 1.0f - 1e-6f;
  std::vector<std::vector<uint32_t>> incs(n);
  for (uint64_t e = 0; e < ne; ++e)
    if (edgeScore[e] > 0) {
      if (edgeScore[e] > SSCR_MAX) edgeScore[e] = SSCR_MAX;
      incs[from[e]].push_back((uint32_t)e);
      incs[to[e]].push_back((uint32_t)e);
    }
  auto other = [&](uint32_t e, size_t v) -> size_t { return from[e] == v ? to[e] : from[e]; };
  for (size_t v = 0; v < n; ++v)
    std::sort(incs[v].begin(), incs[v].end(),
              [&](uint32_t e1, uint32_t e2) { return other(e1, v) < other(e2, v); });
  std::vector<size_t> node_order(n);
  std::iota(node_order.begin(), node_order.end(), 0);
  std::shuffle(node_order.begin(), node_order.end(), std::default_random_engine(seed));
  // per-edge log(1 - w) (src/impl/rb_cluster.cpp:436-444)
  std::vector<float> logsscr(ne);
  for (uint64_t e = 0; e < ne; ++e) logsscr[e] = (float)std::log(1. - edgeScore[e]);
  std::vector<double> nscore(n, 0.0);
  std::vector<int32_t> ncount(n, 0);
  std::vector<size_t> touched;
  std::vector<char> active(n, 1);
  std::unordered_map<size_t, std::unordered_set<size_t>> visited;
  std::unordered_set<size_t> blacklist;
  int round = 0;
  while (true) {
    ++round;
    size_t changed = 0;
    for (size_t i = 0; i < node_order.size(); i++) {
      const size_t v1 = node_order[i];
      if (!active[v1]) continue;
      active[v1] = 0;
      touched.clear();
      for (uint32_t e : incs[v1]) {
        const size_t k = membership[other(e, v1)];
        if (ncount[k] == 0) touched.push_back(k);
        nscore[k] += logsscr[e];
        ncount[k]++;
      }
      if (touched.empty()) continue;
      const size_t current_k = membership[v1];
      bool current_is_candidate = false;
      double current_val = -std::numeric_limits<double>::infinity();
      double best_val = -std::numeric_limits<double>::infinity();
      size_t best_k = touched[0];
      for (size_t k : touched) {
        const double val = chi2_cdf((int)ncount[k], -2.0 * nscore[k]);
        if (k == current_k) {
          current_is_candidate = true;
          current_val = val;
        }
        if (val > best_val) {
          best_val = val;
          best_k = k;
        }
        nscore[k] = 0.0;
        ncount[k] = 0;
      }
      if (current_is_candidate && current_val >= best_val - 1e-12) best_k = current_k;
      const size_t kPrev = membership[v1];
      if (kPrev != best_k && blacklist.find(v1) == blacklist.end()) {
        membership[v1] = best_k;
        ++changed;
        for (uint32_t e : incs[v1]) active[other(e, v1)] = 1;
        auto& seen = visited[v1];
        if (seen.empty()) seen.insert(kPrev);
        if (!seen.insert(best_k).second) blacklist.insert(v1);
      }
    }
    if (changed == 0) break;
  }
  for (uint64_t i = 0; i < n; ++i) membership_out[i] = membership[i];
  if (frozen_out)
    for (size_t v : blacklist) frozen_out[v] = 1;
  return round;
}

}  // extern "C"
