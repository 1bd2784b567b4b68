// CPU oracle for the RabbitBin graph-construction path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or, for bench.py, as the timed CPU restatement).  The product library
// (rabbitbin_amd/librbg.so) never links or calls it.
//
// Every function restates the reference algorithm it names (file:line cited
// at each definition).  Arithmetic follows the reference build: GCC 11,
// -std=c++17 -O3 -march=native.  GCC fuses a*b+c into one FMA under those
// flags (C++ keeps -ffp-contract=fast even in ISO mode), so each place where
// the reference expression is fused is written here as an explicit std::fma,
// and this file is compiled with -ffp-contract=off.
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

// ---- RNG primitives (src/probmh.cpp:60-87, 319-357; src/hash_int.h:79-87)
uint64_t orc_fmix64(uint64_t x, uint64_t seed);
void orc_wyrand_seq(uint64_t* state, int n, uint64_t* out);
void orc_wy_uniform_int_seq(uint64_t* state, uint32_t bound, int n, uint32_t* out);
void orc_zig_seq(uint64_t* state, int n, double* out);
void orc_ted_seq(double c1, double c2, double c3, uint64_t* state, int n, double* out);
// TED table: out[5*i + {0..4}] = boundary, gap, c1, c2, c3 for i < m-1.
void orc_ted_params(uint32_t m, double* out, double* first_boundary_inv);

// ---- streaming ProbMinHash4 (src/probmh.cpp:410-686)
void* orc_pmh_new(uint32_t m, uint64_t seed);
void orc_pmh_free(void* p);
void orc_pmh_reset(void* p);
void orc_pmh_add_hash(void* p, uint64_t h, double w);
void orc_pmh_get(void* p, uint64_t* winners, double* registers);

// ---- stage (a): build_pmh_winners, k = 4, GC-normalised weights
//      (src/rabbitbin.cpp:1436-1589, g_gc_norm = 1, cap 20)
void orc_kmer4_hist(const char* seq, size_t len, uint32_t cnt[256],
                    uint64_t base_cnt[4], uint64_t* total);
void orc_kmer4_weights(const uint32_t cnt[256], const uint64_t base_cnt[4],
                       uint64_t total, double cap, double w[256]);
void orc_build_pmh_winners(const char* seq, size_t len, uint32_t m,
                           uint64_t seed, uint32_t* out32);
void orc_sketch_batch(const char* seqs, const uint64_t* offsets, size_t n,
                      uint32_t m, uint64_t seed, uint32_t* out32, int threads);
void orc_fold16(const uint32_t* in, size_t n, uint16_t* out);

// ---- depth features (src/rabbitbin.cpp:4784-4988, src/ranker.h:55-133,
//      src/impl/rb_abundance.cpp:94-114)
// depth: N x S raw per-sample means (row-major, float).  Outputs (any may be
// null): ranked N x S, unit N x S, raw N x S, colnorm S, anynz N.
void orc_depth_features(const float* depth, size_t n, size_t S, double min_cv,
                        float* ranked, float* unit, float* raw,
                        double* colnorm, uint8_t* anynz);

// ---- stage (b) primitives
uint32_t orc_match_count16(const uint16_t* a, const uint16_t* b, uint32_t m);
double orc_pmh_baseline(const uint16_t* win16, size_t n, uint32_t m);
float orc_unit_dot_f(const float* a, const float* b, uint32_t S);
// depth_corr_fast (src/rabbitbin.cpp:705-746): mode 0 in the reference's
// AVX-512 order, mode 1 in FP64 throughout; the intrinsics form returns -1 on a
// host without AVX-512F.
double orc_depth_corr(const float* a, const float* b, uint32_t S, int mode);
int orc_depth_corr_avx512(const float* a, const float* b, uint32_t S, double* out);
double orc_sv_from_count(uint32_t cnt, uint32_t m, double b0);

// glibc rand() stream after srand(seed), through the reentrant interface.
void orc_rand_seq(unsigned seed, int n, int32_t* out);

// ---- stages (b)+(c): graph construction, both regimes
//      fused  (n > 25000): src/impl/rb_graph.cpp:501-1356
//      small  (n <= 25000): src/impl/rb_graph.cpp:1362-1384 + 7-72
struct OrcGraphIn {
  const uint16_t* win16;   // n x m
  const float* unit;       // n x S (centred, L2-normalised ranks) or null
  const uint8_t* anynz;    // n or null (null: is_nz always true)
  uint64_t n;
  uint32_t m;
  uint32_t S;
  uint32_t K;              // --max-edges
  double min_edge_weight;  // --min-edge-score / 100
  double coverage;         // --max-posterior / 100
  double w_comp;           // composition weight (auto: S<=2 -> 1, else 0)
  double b0;               // PMH baseline
  unsigned rand_seed;      // srand() seed consumed by calibration
  int threads;
  int force_regime;        // 0 auto, 1 small, 2 fused
};
struct OrcGraphOut {
  double sim_cutoff;       // simCutoff (x1000 scale, as returned by the ref)
  uint64_t n_edges;
  uint32_t* from;          // library-owned; free with orc_graph_free
  uint32_t* to;
  float* sv;
  // per-row diagnostics (library-owned): top-K lists sorted by
  // (sv desc, id asc); len <= K; rowmax = max count used for calibration.
  uint32_t* topk_id;       // n x K
  float* topk_sv;          // n x K
  uint32_t* topk_len;      // n
  uint32_t* rowmax_cnt;    // n
  int regime;              // 1 small, 2 fused
  uint64_t gated_pairs;    // pairs that reached the PMH compare
};
int orc_build_graph(const OrcGraphIn* in, OrcGraphOut* out);
void orc_graph_free(OrcGraphOut* out);

// ---- stage (d): edge weights (src/rabbitbin.cpp:5482-5509, 930-970,
//      685-746).  mode 0: reference arithmetic; mode 1: pure FP64.
//      Writes (float) weight per edge; cut edges get 0.
void orc_edge_weights(const uint32_t* from, const uint32_t* to,
                      const float* scomp, uint64_t ne, const float* unit,
                      const float* raw, const double* colnorm, uint32_t S,
                      double w_comp, double min_edge_weight,
                      double neg_depth_thr, int mode, float* out_w,
                      double* out_w64);
// Per edge 1 when its FP64 correlation lies within tol of neg_depth_thr or its
// FP64 weight before the cut within tol of min_edge_weight.
void orc_edge_near(const uint32_t* from, const uint32_t* to, const float* scomp, uint64_t ne,
                   const float* unit, const float* raw, const double* colnorm, uint32_t S,
                   double w_comp, double min_edge_weight, double neg_depth_thr, double tol,
                   uint8_t* near);

}  // extern "C"
