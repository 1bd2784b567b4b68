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
  // Per register k, the elements in ascending order of T[x][k] (ties: smaller
  // x first), laid out [rank][k] so consecutive registers are contiguous:
  // sT[r*m + k] = T[sX[r*m + k]][k].
  std::vector<double> sT;
  std::vector<uint8_t> sX;
};
const ProposalTables& proposal_tables(uint32_t m, uint64_t seed);

// pmh_sv_from_count for cnt = 0..m (src/rabbitbin.cpp:1297-1305).
void sv_tables(uint32_t m, double b0, std::vector<double>& svd, std::vector<float>& svf,
               std::vector<uint16_t>& level);

// Depth features (src/rabbitbin.cpp:4784-4988).  `parallel`: rows on the
// host's OpenMP threads (off when the caller already runs one call per thread).
// `stride` (0: S): unit and raw rows are written `stride` floats apart, the
// padding after each row's S values zeroed.
void depth_features(const float* depth, size_t n, size_t S, double min_cv, float* ranked,
                    float* unit, float* raw, double* colnorm, uint8_t* anynz, bool parallel = true,
                    size_t stride = 0);
// sum[c] += the n rows' depth[r][c], rows in order for every column.
void column_sums(const float* depth, size_t n, size_t S, double* sum);

// int8 image of the unit rows for the MFMA prefilter of the abundance gate.
// q = round(u * scale) with scale = 127 / max|u| (zero padded to kpad).  The
// prefilter keeps a pair iff its integer dot >= gate_qcut(thr); gate_qcut
// charges the quantisation error (Cauchy-Schwarz over the per-row error
// norms) and the forward error of the FP32 dot's rounding chain, so every
// pair whose exact FP32 gate value reaches thr is kept.
struct GateQuant {
  std::vector<int8_t> q;  // n x kpad
  uint32_t kpad = 0;
  double scale = 0;       // 0 when every unit entry is zero
  double max_qnorm = 0;   // max_i ||q_i / scale||
  double max_err = 0;     // max_i ||u_i - q_i / scale||
  double max_unorm = 0;   // max_i ||u_i||
  uint32_t S = 0;
};
// `parallel` (here and below): rows on the host's OpenMP threads.  Row r of
// the quantisation is unit row order[r] (r when order is null); rows are
// `stride` floats apart (0: S).
void gate_quantize(const float* unit, size_t n, uint32_t S, GateQuant& g, bool parallel = true,
                   const uint32_t* order = nullptr, size_t stride = 0);
// Rows appended to an existing quantisation: the same scale (values beyond
// +-127 clamp, and the clamp error counts in max_err), g.q = the new rows only,
// the norm maxima taken over the earlier and the new rows.  The cut stays a
// rigorous bound for every pair of rows quantised under g.
void gate_quantize_more(const float* unit, size_t n, uint32_t S, GateQuant& g, bool parallel = true,
                        const uint32_t* order = nullptr, size_t stride = 0);
// Row order of the fused pair pass: rows sorted by the indices of their three
// largest unit-vector entries (ties to the smaller sample, then by row), so
// contigs with the same abundance pattern are adjacent and the pairs that
// pass the abundance gate fall into few tiles.  perm[r] = original row.
void pair_pass_order(const float* unit, size_t n, uint32_t S, std::vector<uint32_t>& perm,
                     bool parallel = true, size_t stride = 0);
// Smallest integer dot a kept pair can have; INT32_MIN disables the prefilter.
int32_t gate_qcut(const GateQuant& g, float thr);
// Smallest integer dot that proves a pair kept: the same error terms taken the
// other way, so qdot >= gate_qcut_hi(thr) implies fl(u_i . u_j) >= thr (and,
// the cut being positive, both rows nonzero).  INT32_MAX: no pair is proven.
int32_t gate_qcut_hi(const GateQuant& g, float thr);

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

// The data-independent part of the calibration of n contigs under
// srand(seed): the contig draws (regime 1: n rand() calls whose order does
// not matter; regime 2: 10 rounds of a 2500-element Fisher-Yates prefix,
// src/impl/rb_graph.cpp:502-522) and the kCalibRand rand() values that
// follow them, which calib_converge consumes (at most ~135 per round).
constexpr uint32_t kCalibRand = 2048;
struct CalibPlan {
  uint32_t n = 0;
  unsigned seed = 0;
  int regime = 0;
  std::vector<int32_t> rnd;          // kCalibRand values
  std::vector<uint32_t> members;     // regime 2: contigs by lowest sampled round
  std::vector<uint32_t> round_off;   // regime 2: 11 offsets into members
};
void calibration_plan(uint32_t n, unsigned seed, int regime, CalibPlan& plan);
double baseline_median(const std::vector<uint32_t>& counts, uint32_t m);

}  // namespace rbg
