// Host-side pieces of librbg that stay on the CPU because the reference
// computes them with host libraries whose exact results are the contract:
// glibc rand()/log1p/expm1/exp, libstdc++ mt19937_64 and
// uniform_int_distribution.  Internal header (not part of the C ABI).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace rbg {

constexpr int kElems = 136;  // canonical 4-mer codes

// Canonical 4-mer codes in ascending order (the reference inserts them into
// the sketch in this order, src/rabbitbin.cpp:1500-1514).
const uint8_t* canonical_codes();   // [136] -> code v
const int16_t* element_of_code();   // [256] -> element index or -1

// Per-(element, register) proposal tables of the weighted ProbMinHash4
// stream (src/probmh.cpp:634-686).  For element x the stream's proposals do
// not depend on its weight w except through one multiply hv = (1/w) * T, so
// the sketch of a contig is  leaf[k] = min_x fl((1/w_x) * T[x][k])  with ties
// to the smaller x, and winner[k] = W[x*][k].
struct ProposalTables {
  uint32_t m = 0;
  uint64_t seed = 0;
  std::vector<double> T;     // [136][m]
  std::vector<uint64_t> W;   // [136][m] raw 64-bit winners
  std::vector<uint32_t> W32; // [136][m] folded 64 -> 32
  std::vector<uint16_t> W16; // [136][m] fold_win16
  std::vector<uint8_t> cls;  // [m][136] class byte: smallest x' with equal
                             // fold16 at this register; 0xFF if fold16 == 0
};
const ProposalTables& proposal_tables(uint32_t m, uint64_t seed);

// pmh_sv_from_count for cnt = 0..m (src/rabbitbin.cpp:1297-1305).
void sv_tables(uint32_t m, double b0, std::vector<double>& svd, std::vector<float>& svf,
               std::vector<uint16_t>& level);

// Depth features (src/rabbitbin.cpp:4784-4988).
void depth_features(const float* depth, size_t n, size_t S, double min_cv, float* ranked,
                    float* unit, float* raw, double* colnorm, uint8_t* anynz);

// glibc rand() stream with private state (srand(seed) then rand()).
class RandStream {
 public:
  explicit RandStream(unsigned seed);
  RandStream(const RandStream&) = delete;
  RandStream& operator=(const RandStream&) = delete;
  int next();
 private:
  struct random_data rd_;
  char state_[128];
};

// Calibration (src/impl/rb_graph.cpp:390-415, 502-525, 1243-1261, 1362-1384).
// rowmax: per-row maximum match count (small: over all partners; fused: over
// gated partners).  Returns simCutoff (per-mille).
double calibrate_small(const std::vector<uint32_t>& rowmax, const std::vector<double>& svd,
                       double coverage, unsigned rand_seed);
double calibrate_fused(const std::vector<uint32_t>& rowmax, const std::vector<float>& svf,
                       double coverage, unsigned rand_seed);

// The 30000 (i, j) index draws of estimate_pmh_baseline (libstdc++
// mt19937_64 + uniform_int_distribution, src/rabbitbin.cpp:1340-1352);
// pairs with i == j are dropped as in the reference.
void baseline_pairs(size_t n, std::vector<uint32_t>& pi, std::vector<uint32_t>& pj);
double baseline_median(const std::vector<uint32_t>& counts, uint32_t m);

}  // namespace rbg
