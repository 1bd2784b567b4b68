// Test accessor for the reference ProbMinHash4 (oracle/_ref only).
//
// This translation unit includes the reference's src/probmh.cpp as it lies
// under the reference tree (path supplied by the Makefile through -I), so the
// reference's own code is compiled unchanged with the reference's flags
// (-std=c++17 -O3 -march=native -funroll-loops, src/CMakeLists.txt:210-253).
// It only adds extern "C" entry points that call the reference's functions,
// giving the tests golden outputs for the RNG streams and the sketches.
// Nothing in the product links this file.
#include "probmh.cpp"

extern "C" {

void* ref_pmh_new(uint32_t m, int k, uint64_t seed) {
  return new Sketch::ProbMinHash4(m, k, seed);
}
void ref_pmh_free(void* p) { delete static_cast<Sketch::ProbMinHash4*>(p); }
void ref_pmh_reset(void* p) { static_cast<Sketch::ProbMinHash4*>(p)->reset(); }
void ref_pmh_add_hash(void* p, uint64_t h, double w) {
  static_cast<Sketch::ProbMinHash4*>(p)->addHash(h, w);
}
void ref_pmh_get(void* p, uint64_t* winners, double* registers) {
  auto* s = static_cast<Sketch::ProbMinHash4*>(p);
  const uint32_t m = s->getM();
  for (uint32_t r = 0; r < m; ++r) {
    if (winners) winners[r] = s->getWinners()[r];
    if (registers) registers[r] = s->getRegisters()[r];
  }
}
// n consecutive wyrand outputs starting from *state (state advanced).
void ref_wyrand_seq(uint64_t* state, int n, uint64_t* out) {
  for (int i = 0; i < n; ++i) out[i] = wyrand_next(*state);
}
void ref_wy_uniform_int_seq(uint64_t* state, uint32_t bound, int n, uint32_t* out) {
  for (int i = 0; i < n; ++i) out[i] = wy_uniform_int(bound, *state);
}
void ref_zig_seq(uint64_t* state, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = zig_exponential(*state);
}
void ref_ted_seq(double c1, double c2, double c3, uint64_t* state, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = ted_sample(c1, c2, c3, *state);
}
uint64_t ref_fmix64(uint64_t x, uint64_t seed) { return mc::murmur3_fmix(x, seed); }
void ref_zig_tables(double* x, double* y) {
  for (int i = 0; i < 257; ++i) { x[i] = zig_table_x[i]; y[i] = zig_table_y[i]; }
}

}  // extern "C"
