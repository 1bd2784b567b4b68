// Compiler probe for the frozen-core recruitment arithmetic of RabbitBin.
//
// recruit_to_frozen_cores and frozen_core_unique_match (src/impl/rb_split.cpp:
// 739-862) accumulate `norm2 += (double)u[k] * u[k]`, `sum_norm2 += v * v` and
// `score += (double)unit[k] * core.centroid[k]`; whether the compiler fuses
// each multiply-add decides the last bits of the thresholds and scores.  This
// program restates those loops in the same form and types, so that built with
// the reference's flags (-std=c++17 -O3 -march=native -funroll-loops, no
// FP-contraction option) it yields the bits the reference build yields on this
// host.  A test compares them with the oracle (which writes every multiply-add
// as std::fma) and with librbg's assignments.
//
// I/O (binary, native endianness): stdin u64 S, u64 n_large, u64 n_bins,
// n_large*S f32 unit rows, then per bin i64 id, u64 count, count u64 members
// (large contigs), then u64 min_cs, u64 n_cand, n_cand*S f32 rows; stdout
// u64 cores, per core i64 id, f64 threshold, S f64 centroid, then
// n_cand*cores f64 scores.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <vector>

static void rd(void* p, size_t n) {
  if (fread(p, 1, n, stdin) != n) std::exit(2);
}
template <class T>
static T rd1() {
  T v;
  rd(&v, sizeof v);
  return v;
}

struct Core {
  int64_t id;
  double threshold;
  std::vector<double> centroid;
};

int main() {
  const uint64_t S = rd1<uint64_t>(), n_large = rd1<uint64_t>(), n_bins = rd1<uint64_t>();
  std::vector<float> unit(n_large * S);
  rd(unit.data(), unit.size() * 4);
  std::map<int64_t, std::vector<size_t>> bins;
  for (uint64_t b = 0; b < n_bins; ++b) {
    const int64_t id = rd1<int64_t>();
    const uint64_t cnt = rd1<uint64_t>();
    std::vector<size_t> m(cnt);
    for (auto& x : m) x = (size_t)rd1<uint64_t>();
    bins[id] = m;
  }
  const uint64_t minCS = rd1<uint64_t>(), n_cand = rd1<uint64_t>();
  std::vector<float> cand(n_cand * S);
  rd(cand.data(), cand.size() * 4);

  std::vector<Core> cores;
  for (const auto& kv : bins) {
    std::vector<double> sum(S, 0.0);
    double self_sum = 0.0;
    size_t valid = 0;
    for (size_t c : kv.second) {
      const float* u = unit.data() + c * S;
      double norm2 = 0.0;
      for (size_t k = 0; k < S; ++k) norm2 += (double)u[k] * u[k];
      if (!(norm2 > 0.0)) continue;
      for (size_t k = 0; k < S; ++k) sum[k] += u[k];
      self_sum += norm2;
      ++valid;
    }
    if (valid < std::max<size_t>(2, minCS)) continue;
    double sum_norm2 = 0.0;
    for (double v : sum) sum_norm2 += v * v;
    double tau = (sum_norm2 - self_sum) / ((double)valid * (double)(valid - 1));
    tau = std::max(-1.0, std::min(1.0, tau));
    for (double& v : sum) v /= (double)valid;
    cores.push_back(Core{kv.first, tau, sum});
  }
  const uint64_t nc = cores.size();
  fwrite(&nc, 8, 1, stdout);
  for (const Core& c : cores) {
    fwrite(&c.id, 8, 1, stdout);
    fwrite(&c.threshold, 8, 1, stdout);
    fwrite(c.centroid.data(), 8, S, stdout);
  }
  for (uint64_t r = 0; r < n_cand; ++r) {
    const float* unitp = cand.data() + r * S;
    for (const Core& core : cores) {
      double score = 0.0;
      for (size_t k = 0; k < S; ++k) score += (double)unitp[k] * core.centroid[k];
      fwrite(&score, 8, 1, stdout);
    }
  }
  return 0;
}
