// Calibration helpers shared by the oracle's graph restatement and its
// performance port of the fused pass.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <utility>
#include <vector>

namespace orc_calib {

// glibc random_r stream after initstate_r(seed): the values rand() returns
// after srand(seed).
struct RandStream {
  struct random_data rd;
  char state[128];
  explicit RandStream(unsigned seed) {
    std::memset(&rd, 0, sizeof(rd));
    initstate_r(seed, state, sizeof(state), &rd);
  }
  int next() {
    int32_t r;
    random_r(&rd, &r);
    return r;
  }
};

// random_unique (src/rabbitbin.h:330-341).
template <class It>
inline void random_unique(It begin, It end, size_t num, RandStream& rs) {
  size_t left = (size_t)std::distance(begin, end);
  while (num--) {
    It r = begin;
    std::advance(r, (size_t)rs.next() % left);
    std::swap(*begin, *r);
    ++begin;
    --left;
  }
}

// calib_converge (src/impl/rb_graph.cpp:390-415), RB_SIM_FLOOR = 50.
inline size_t calib_converge(const std::vector<double>& maxsim, double coverage, RandStream& rs) {
  const size_t n = maxsim.size();
  size_t p = 999, pp = 1000;
  double cov = 0, pcov = 0;
  for (; p > 50;) {
    const double cutoff = (double)p / 1000.;
    size_t on = 0;
    for (size_t i = 0; i < n; ++i)
      if (maxsim[i] >= cutoff) on++;
    cov = (double)on / n;
    if (cov >= coverage) {
      if (cov - coverage > coverage - pcov) {
        p = pp;
        cov = pcov;
      }
      break;
    }
    pp = p;
    pcov = cov;
    if (p > 990)
      p -= rs.next() % 3 + 1;
    else if (p > 900)
      p -= rs.next() % 3 + 3;
    else
      p -= rs.next() % 3 + 9;
  }
  return p;
}

}  // namespace orc_calib
