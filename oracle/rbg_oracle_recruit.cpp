// Oracle (test infrastructure): frozen-core recruitment, restated from the
// reference for the parity tests.  Follows recruit_to_frozen_cores
// (src/impl/rb_split.cpp:769-919) and frozen_core_unique_match (:739-762)
// literally: bins visited by ascending id (the reference sorts the BinMap
// keys), the core sums in member order, the closed-form mean Spearman
// threshold, the long batch, then the short batch with each small contig's
// rank transform (src/ranker.h average ranks, mean, ss, 1/sqrt) and the
// first-match / second-match-is-ambiguous scan over the cores.
//
// Every `acc += a * b` of the reference is written as std::fma: the reference
// build (-O3 -march=native, GCC) contracts them, except in the sum of the
// squared core sums, whose first 4 * floor(S / 4) products it forms four at a
// time with vector multiplies (rounded) and adds in order (pinned by
// tests/probes/recruit_probe.cpp, built with the reference's flags).  The
// ranks here come from counting (rank = 1 + #smaller + (#equal - 1)/2), not
// from a sort, so they check the library's sort-based ranks independently.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <vector>

namespace {

struct Core {
  int32_t id;
  double thr;
  std::vector<double> cent;
};

std::vector<Core> build_cores(uint32_t S, uint32_t n_large, const float* unit, const uint64_t* len_large,
                              const uint64_t* len_small, uint32_t n_bins, const int32_t* bin_id,
                              const uint64_t* bin_off, const uint64_t* members, uint64_t core_floor,
                              uint64_t min_cs) {
  std::map<int32_t, uint32_t> by_id;  // ascending id, as the reference sorts the keys
  for (uint32_t b = 0; b < n_bins; ++b) by_id[bin_id[b]] = b;
  std::vector<Core> cores;
  for (const auto& kv : by_id) {
    const uint32_t b = kv.second;
    uint64_t span = 0;
    for (uint64_t t = bin_off[b]; t < bin_off[b + 1]; ++t)
      span += members[t] < n_large ? len_large[members[t]] : len_small[members[t] - n_large];
    if (span < core_floor) continue;
    std::vector<double> sum(S, 0.0);
    double self_sum = 0.0;
    uint64_t valid = 0;
    for (uint64_t t = bin_off[b]; t < bin_off[b + 1]; ++t) {
      const uint64_t c = members[t];
      if (c >= n_large) continue;
      const float* u = unit + c * S;
      double norm2 = 0.0;
      for (uint32_t k = 0; k < S; ++k) norm2 = std::fma((double)u[k], (double)u[k], norm2);
      if (!(norm2 > 0.0)) continue;
      for (uint32_t k = 0; k < S; ++k) sum[k] += u[k];
      self_sum += norm2;
      ++valid;
    }
    if (valid < std::max<uint64_t>(2, min_cs)) continue;
    // the reference build forms the first 4 * floor(S / 4) products four at
    // a time (rounded, added in order) and fuses the last S mod 4
    const uint32_t q = S >= 4 ? (S & ~3u) : 0u;
    double sum_norm2 = 0.0;
    for (uint32_t k = 0; k < q; ++k) {
      const double p = sum[k] * sum[k];
      sum_norm2 += p;
    }
    for (uint32_t k = q; k < S; ++k) sum_norm2 = std::fma(sum[k], sum[k], sum_norm2);
    double tau = (sum_norm2 - self_sum) / ((double)valid * (double)(valid - 1));
    tau = std::max(-1.0, std::min(1.0, tau));
    for (double& v : sum) v /= (double)valid;
    cores.push_back(Core{kv.first, tau, sum});
  }
  return cores;
}

double score_of(const float* u, uint32_t S, const Core& c) {
  double score = 0.0;
  for (uint32_t k = 0; k < S; ++k) score = std::fma((double)u[k], c.cent[k], score);
  return score;
}

int unique_match(const float* u, uint32_t S, const std::vector<Core>& cores, bool& ambiguous) {
  ambiguous = false;
  double norm2 = 0.0;
  for (uint32_t k = 0; k < S; ++k) norm2 = std::fma((double)u[k], (double)u[k], norm2);
  if (!(norm2 > 0.0)) return -1;
  int match = -1;
  for (const Core& c : cores) {
    if (score_of(u, S, c) >= c.thr) {
      if (match >= 0) {
        ambiguous = true;
        return -1;
      }
      match = c.id;
    }
  }
  return match;
}

}  // namespace

extern "C" {

// Cores: ids, thresholds and centroids (n_out x S) of at most cap cores;
// returns the core count.
uint64_t orc_recruit_cores(uint32_t S, uint32_t n_large, const float* unit, const uint64_t* len_large,
                           const uint64_t* len_small, uint32_t n_bins, const int32_t* bin_id, const uint64_t* bin_off,
                           const uint64_t* members, uint64_t core_floor, uint64_t min_cs, uint64_t cap,
                           int32_t* ids, double* thr, double* cent) {
  const std::vector<Core> cores =
      build_cores(S, n_large, unit, len_large, len_small, n_bins, bin_id, bin_off, members, core_floor, min_cs);
  for (size_t i = 0; i < cores.size() && i < cap; ++i) {
    ids[i] = cores[i].id;
    thr[i] = cores[i].thr;
    std::copy(cores[i].cent.begin(), cores[i].cent.end(), cent + i * S);
  }
  return cores.size();
}

// Scores of n unit rows against the cores (ncores x S centroids), n x ncores.
void orc_recruit_scores(uint32_t S, uint64_t n, const float* unit, uint64_t ncores, const double* cent,
                        double* out) {
  for (uint64_t r = 0; r < n; ++r)
    for (uint64_t c = 0; c < ncores; ++c) {
      double score = 0.0;
      for (uint32_t k = 0; k < S; ++k) score = std::fma((double)unit[r * S + k], cent[c * S + k], score);
      out[r * ncores + c] = score;
    }
}

// The whole recruitment.  stats: cores, long assigned / ambiguous /
// unmatched, short assigned / ambiguous / unmatched.
void orc_recruit_frozen_cores(uint32_t S, uint32_t n_large, const float* unit_large, const uint64_t* len_large,
                              uint32_t n_small, const float* depth_small, const uint64_t* len_small, uint32_t n_bins,
                              const int32_t* bin_id, const uint64_t* bin_off, const uint64_t* members,
                              uint64_t core_floor, uint64_t min_cs, int32_t* assign_large, int32_t* assign_small,
                              uint64_t* stats) {
  std::fill(assign_large, assign_large + n_large, -1);
  std::fill(assign_small, assign_small + n_small, -1);
  std::fill(stats, stats + 7, 0ull);
  if (S < 2 || n_bins == 0) return;
  const std::vector<Core> cores = build_cores(S, n_large, unit_large, len_large, len_small, n_bins, bin_id, bin_off,
                                              members, core_floor, min_cs);
  stats[0] = cores.size();
  if (cores.empty()) return;
  std::vector<char> binned(n_large, 0);
  for (uint64_t t = bin_off[0]; t < bin_off[n_bins]; ++t)
    if (members[t] < n_large) binned[members[t]] = 1;
  for (uint32_t c = 0; c < n_large; ++c) {
    if (binned[c]) continue;
    bool amb = false;
    const int id = unique_match(unit_large + (size_t)c * S, S, cores, amb);
    assign_large[c] = id;
    if (id >= 0) ++stats[1];
    else if (amb) ++stats[2];
    else ++stats[3];
  }
  std::vector<double> raw(S), ranked(S);
  std::vector<float> unit(S);
  for (uint32_t s = 0; s < n_small; ++s) {
    for (uint32_t k = 0; k < S; ++k) raw[k] = (double)depth_small[(size_t)s * S + k];
    for (uint32_t k = 0; k < S; ++k) {  // average ranks by counting
      uint32_t lt = 0, eq = 0;
      for (uint32_t q = 0; q < S; ++q) {
        lt += raw[q] < raw[k];
        eq += raw[q] == raw[k];
      }
      ranked[k] = 1.0 + (double)lt + ((double)eq - 1.0) / 2.0;
    }
    double mean = 0.0;
    for (double v : ranked) mean += v;
    mean /= (double)S;
    double ss = 0.0;
    for (double v : ranked) {
      const double d = v - mean;
      ss = std::fma(d, d, ss);
    }
    if (!(ss > 0.0)) {
      ++stats[6];
      continue;
    }
    const double inv = 1.0 / std::sqrt(ss);
    for (uint32_t k = 0; k < S; ++k) unit[k] = (float)((ranked[k] - mean) * inv);
    bool amb = false;
    const int id = unique_match(unit.data(), S, cores, amb);
    assign_small[s] = id;
    if (id >= 0) ++stats[4];
    else if (amb) ++stats[5];
    else ++stats[6];
  }
}

}  // extern "C"
