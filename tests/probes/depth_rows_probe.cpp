// Compiler probe for the depth-row arithmetic of RabbitBin's graph path.
//
// The reference derives the abundance gate's inputs and the edge weights' column
// norms with plain C++ expressions whose rounding depends on the compiler and
// its flags (whether `ss += d * d` is contracted into a fused multiply-add).  This
// program restates those expressions in the same form and types, so that built
// with the reference's flags (-std=c++17 -O3 -march=native -funroll-loops, no
// FP-contraction option: CMakeLists.txt:4-5, 49-51, src/CMakeLists.txt:210-253)
// it yields the bits the reference build yields on this host.  A test compares
// them with librbg's rbg_depth_features and the oracle.
//
// Expressions restated:
//   average ranks, ties at the mean position       src/ranker.h:55-66 (float)
//   centred, L2-normalised rank rows                src/rabbitbin.cpp:4972-4988
//   per-sample column norms 1/mean of raw depths    src/rabbitbin.cpp:4799-4806
//
// I/O (binary, native endianness): stdin  u64 n, u64 S, n*S f32 raw depths;
// stdout n*S f32 unit rows, then S f64 column norms.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <vector>

typedef float StoredDistance;  // the reference stores depths as float

static void average_ranks(std::vector<StoredDistance>& row) {
  const size_t S = row.size();
  std::vector<uint32_t> order(S);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return row[a] < row[b]; });
  std::vector<StoredDistance> out(S);
  for (uint32_t c = 0, reps; c < S; c += reps) {
    reps = 1;
    while (c + reps < S && row[order[c]] == row[order[c + reps]]) ++reps;
    for (uint32_t k = 0; k < reps; ++k) out[order[c + k]] = StoredDistance(2 * c + reps - 1) / 2 + 1;
  }
  row = out;
}

int main() {
  uint64_t hdr[2];
  if (std::fread(hdr, 8, 2, stdin) != 2) return 1;
  const size_t nobs = hdr[0], S = hdr[1];
  std::vector<StoredDistance> raw(nobs * S);
  if (std::fread(raw.data(), sizeof(float), raw.size(), stdin) != raw.size()) return 1;

  // column norms over the raw depths
  std::vector<double> colnorm(S, 1.0);
  for (size_t i = 0; i < S; ++i) {
    double s = 0.0;
    for (size_t r = 0; r < nobs; ++r) s += raw[r * S + i];
    double mean = s / std::max<size_t>(nobs, 1);
    colnorm[i] = (mean > 1e-12) ? 1.0 / mean : 0.0;
  }

  // rank transform per row, then centred + L2-normalised rows
  std::vector<StoredDistance> ranked(raw);
  std::vector<float> unit(nobs * S, 0.0f);
  if (S > 1) {
    for (size_t r = 0; r < nobs; ++r) {
      std::vector<StoredDistance> row(ranked.begin() + r * S, ranked.begin() + (r + 1) * S);
      average_ranks(row);
      std::copy(row.begin(), row.end(), ranked.begin() + r * S);
    }
    for (size_t r = 0; r < nobs; ++r) {
      const StoredDistance* x = ranked.data() + r * S;
      double mean = 0.0;
      for (size_t k = 0; k < S; ++k) mean += x[k];
      mean /= (double)S;
      double ss = 0.0;
      for (size_t k = 0; k < S; ++k) {
        double d = (double)x[k] - mean;
        ss += d * d;
      }
      if (ss > 0.0) {
        const double inv = 1.0 / std::sqrt(ss);
        float* u = unit.data() + r * S;
        for (size_t k = 0; k < S; ++k) u[k] = (float)(((double)x[k] - mean) * inv);
      }
    }
  }
  std::fwrite(unit.data(), sizeof(float), unit.size(), stdout);
  std::fwrite(colnorm.data(), sizeof(double), colnorm.size(), stdout);
  return 0;
}
