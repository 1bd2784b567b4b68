// Seeded synthetic assemblies for the graph-construction workloads.
//
// Genomes: order-3 Markov chains over ACGT, 64 context rows ~ Dirichlet(2),
// G/C probabilities scaled by g/2 and A/T by (1-g)/2 with g ~ U(0.3, 0.7),
// rows renormalised.  Contig lengths ~ LogNormal(mu, sigma) clamped to
// [2500, 200000].  Contig c belongs to genome gid[c] (an even split over the
// genomes, shuffled); its bases are a run of its genome's chain started from
// a uniform context -- each contig is an independent stretch of the genome's
// process, which gives it the composition of a cut from the genome.
// Depth (MetaBAT means): per genome and sample a ~ LogNormal(0, 1); per
// contig reads ~ Poisson(a * 30 * L / 150), mean = reads * 150 / L.
//
// Every draw comes from a splitmix64 stream keyed by (seed, purpose, index),
// so the output does not depend on the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * 0x1p-53; }           // [0,1)
  double open_unit() { return ((double)(next() >> 11) + 0.5) * 0x1p-53; }  // (0,1)
  double normal() {
    const double u1 = open_unit(), u2 = unit();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  double gamma2() { return -std::log(open_unit() * open_unit()); }  // Gamma(2, 1)
  // Poisson: multiplication method below 10, else PTRS (Hormann 1993).
  uint64_t poisson(double lam) {
    if (lam <= 0) return 0;
    if (lam < 10) {
      const double L = std::exp(-lam);
      uint64_t k = 0;
      double p = unit();
      while (p > L) {
        ++k;
        p *= unit();
      }
      return k;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
    for (;;) {
      const double U = unit() - 0.5, V = unit();
      const double us = 0.5 - std::fabs(U);
      const double kf = std::floor((2 * a / us + b) * U + lam + 0.43);
      if (us >= 0.07 && V <= vr) return (uint64_t)kf;
      if (kf < 0 || (us < 0.013 && V > us)) continue;
      if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
          -lam + kf * loglam - std::lgamma(kf + 1))
        return (uint64_t)kf;
    }
  }
};

inline uint64_t key(uint64_t seed, uint64_t purpose, uint64_t idx) {
  Rng r(seed ^ (purpose * 0xD1B54A32D192ED03ULL) ^ (idx * 0x8CB92BA72F3D8DD7ULL));
  r.next();
  return r.next();
}

const char kBase[4] = {'A', 'C', 'G', 'T'};

}  // namespace

extern "C" {

// Contig lengths and genome ids (n each).
void rbg_synth_layout(uint64_t seed, uint32_t n, uint32_t genomes, double mu, double sigma,
                      uint32_t* lengths, uint32_t* gid) {
  Rng r(key(seed, 1, 0));
  for (uint32_t c = 0; c < n; ++c) {
    double L = std::exp(mu + sigma * r.normal());
    if (L < 2500) L = 2500;
    if (L > 200000) L = 200000;
    lengths[c] = (uint32_t)L;
    gid[c] = c % genomes;
  }
  for (uint32_t c = n; c > 1; --c) {  // shuffle the genome labels
    const uint32_t j = (uint32_t)(r.next() % c);
    const uint32_t t = gid[c - 1];
    gid[c - 1] = gid[j];
    gid[j] = t;
  }
}

// Sequences (offsets[n+1] precomputed by the caller from lengths) and the
// n x S depth matrix of per-sample mean coverage.
void rbg_synth_fill(uint64_t seed, uint32_t n, uint32_t genomes, const uint32_t* lengths,
                    const uint32_t* gid, const uint64_t* offsets, char* seqs, uint32_t S,
                    float* depth, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  // genome models: 64 cumulative rows of 4
  std::vector<double> cum((size_t)genomes * 64 * 4);
  std::vector<double> abund((size_t)genomes * (S ? S : 1));
  for (uint32_t g = 0; g < genomes; ++g) {
    Rng r(key(seed, 2, g));
    const double gc = 0.3 + 0.4 * r.unit();
    for (int ctx = 0; ctx < 64; ++ctx) {
      double p[4], tot = 0;
      for (int b = 0; b < 4; ++b) p[b] = r.gamma2();
      for (int b = 0; b < 4; ++b) {
        p[b] *= (b == 1 || b == 2) ? gc / 2 : (1 - gc) / 2;
        tot += p[b];
      }
      double acc = 0;
      for (int b = 0; b < 4; ++b) {
        acc += p[b] / tot;
        cum[((size_t)g * 64 + ctx) * 4 + b] = acc;
      }
      cum[((size_t)g * 64 + ctx) * 4 + 3] = 1.0;
    }
    for (uint32_t s = 0; s < S; ++s) abund[(size_t)g * S + s] = std::exp(r.normal());
  }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
  for (uint32_t c = 0; c < n; ++c) {
    Rng r(key(seed, 3, c));
    const double* M = &cum[(size_t)gid[c] * 256];
    char* out = seqs + offsets[c];
    const uint32_t L = lengths[c];
    uint32_t ctx = (uint32_t)(r.next() & 63u);
    for (uint32_t t = 0; t < L; ++t) {
      const double u = r.unit();
      const double* row = M + ctx * 4;
      const uint32_t b = u < row[0] ? 0u : u < row[1] ? 1u : u < row[2] ? 2u : 3u;
      out[t] = kBase[b];
      ctx = ((ctx << 2) | b) & 63u;
    }
    if (S) {
      Rng d(key(seed, 4, c));
      for (uint32_t s = 0; s < S; ++s) {
        const double lam = abund[(size_t)gid[c] * S + s] * 30.0 * L / 150.0;
        const uint64_t reads = d.poisson(lam);
        depth[(size_t)c * S + s] = (float)((double)reads * 150.0 / (double)L);
      }
    }
  }
}

}  // extern "C"
