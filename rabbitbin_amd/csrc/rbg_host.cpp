// Host-side pieces of librbg (see rbg_host.h).  Compiled with
// -ffp-contract=off: every multiply-add that the reference build fuses is an
// explicit std::fma here, every other operation is a separate IEEE op.
#include "rbg_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <random>

#include "zig_tables.h"

namespace rbg {
namespace {

// ---- wyrand stream (src/probmh.cpp:60-87) --------------------------------
inline uint64_t wy_next(uint64_t& s) {
  s += 0xa0761d6478bd642fULL;
  const __uint128_t r = (__uint128_t)s * (s ^ 0xe7037ed1a0b428dbULL);
  return (uint64_t)(r >> 64) ^ (uint64_t)r;
}
inline double to_unit(uint64_t bits) { return (double)(bits >> 11) * 0x1p-53; }

inline uint32_t wy_below(uint32_t bound, uint64_t& s) {
  uint64_t prod = (wy_next(s) >> 32) * (uint64_t)bound;
  if ((uint32_t)prod < bound) {
    const uint32_t reject_below = (uint32_t)(0u - bound) % bound;
    while ((uint32_t)prod < reject_below) prod = (wy_next(s) >> 32) * (uint64_t)bound;
  }
  return (uint32_t)(prod >> 32);
}

// ---- samplers -------------------------------------------------------------
struct TedRow {
  double boundary, gap, c1, c2, c3;
};

// Truncated exponential on [0,1) (src/probmh.cpp:344-357).
double trunc_exp(const TedRow& t, uint64_t& s) {
  double x = to_unit(wy_next(s)) * t.c1;
  if (x < 1.0) return x;
  for (;;) {
    x = to_unit(wy_next(s));
    if (x <= t.c2) return x;
    double y = to_unit(wy_next(s)) * 0.5;
    if (y > 1.0 - x) {
      x = 1.0 - x;
      y = 1.0 - y;
    }
    if (x <= t.c3 * (1.0 - y)) return x;
    if (t.c1 * y <= 1.0 - x) return x;
  }
}

// Exp(1) by the 256-layer ziggurat (src/probmh.cpp:319-338); the slow-path
// interpolations are single fused operations in the reference build.
double exp1(uint64_t& s) {
  double shift = 0.0;
  for (;;) {
    const uint64_t bits = wy_next(s);
    const int layer = (int)(bits & 0xFF);
    const double x = to_unit(bits) * kZigX[layer];
    if (x < kZigX[layer + 1]) return shift + x;
    if (layer == 0) {
      shift += kZigX[1];
      continue;
    }
    const double y01 = to_unit(wy_next(s));
    const double y = std::fma(y01, kZigY[layer + 1] - kZigY[layer], kZigY[layer]);
    const double above_upper =
        std::fma(kZigX[layer] - kZigX[layer + 1], y01, -(kZigX[layer] - x));
    const double above_lower =
        y - std::fma(kZigX[layer + 1] - x, kZigY[layer + 1], kZigY[layer + 1]);
    if (above_upper < 0 && (above_lower < 0 || y < std::exp(-x))) return x + shift;
  }
}

// Boundaries and sampler constants of the m-1 proposal intervals
// (src/probmh.cpp:506-572).
void ted_rows(uint32_t m, std::vector<TedRow>& rows, double& first_inv) {
  rows.assign(m - 1, TedRow{});
  const double first = std::log1p(1.0 / (double)(m - 1));
  first_inv = 1.0 / first;
  auto consts = [](double rate, TedRow& r) {
    r.c1 = (rate != 0.0) ? std::expm1(rate) / rate : 1.0;
    r.c2 = (rate != 0.0) ? -std::log1p(std::expm1(-rate) * 0.5) / rate : 0.5;
    r.c3 = (rate != 0.0) ? -std::expm1(-rate) / rate : 1.0;
  };
  rows[0].boundary = 1.0;
  rows[0].gap = 1.0;
  consts(first, rows[0]);
  double prev = first;
  for (uint32_t i = 1; i + 1 < m; ++i) {
    const double b = std::log1p((double)(i + 1) / (double)(m - i - 1));
    rows[i].boundary = b / first;
    rows[i].gap = rows[i].boundary - rows[i - 1].boundary;
    consts(b - prev, rows[i]);
    prev = b;
  }
}

// All m proposals of one element against an empty sketch: proposal p goes to
// the p-th register of a Fisher-Yates permutation drawn from the element's
// own stream; its winner identity is the stream state right after the draw.
void element_proposals(uint32_t m, uint64_t rng, const std::vector<TedRow>& ted,
                       double first_inv, double* T, uint64_t* W) {
  std::vector<uint32_t> perm(m);
  std::iota(perm.begin(), perm.end(), 0u);
  const double u0 = to_unit(rng) * ted[0].c1;
  double hv = (u0 < 1.0) ? u0 : trunc_exp(ted[0], rng);
  for (uint32_t p = 1; p <= m; ++p) {
    const uint32_t j = (p - 1) + wy_below(m - (p - 1), rng);
    const uint32_t reg = perm[j];
    perm[j] = perm[p - 1];
    T[reg] = hv;
    W[reg] = rng;
    if (p + 1 < m)
      hv = std::fma(ted[p].gap, trunc_exp(ted[p], rng), ted[p - 1].boundary);
    else if (p + 1 == m)
      hv = std::fma(first_inv, exp1(rng), ted[m - 2].boundary);
  }
}

inline uint64_t fmix64(uint64_t x, uint64_t seed) {
  x ^= seed;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

inline uint16_t fold16(uint32_t w) {
  if (w == 0) return 0;
  const uint16_t h = (uint16_t)(w ^ (w >> 16));
  return h ? h : (uint16_t)1;
}

struct CodeTables {
  uint8_t codes[kElems];
  int16_t elem[256];
  CodeTables() {
    int x = 0;
    for (int v = 0; v < 256; ++v) {
      int rc = 0;
      for (int d = 0; d < 4; ++d) rc |= (3 - ((v >> (2 * d)) & 3)) << (2 * (3 - d));
      if (v <= rc) {
        codes[x] = (uint8_t)v;
        elem[v] = (int16_t)x++;
      } else {
        elem[v] = -1;
      }
    }
  }
};
const CodeTables& code_tables() {
  static const CodeTables t;
  return t;
}

}  // namespace

const uint8_t* canonical_codes() { return code_tables().codes; }
const int16_t* element_of_code() { return code_tables().elem; }

const ProposalTables& proposal_tables(uint32_t m, uint64_t seed) {
  static std::mutex mu;
  static std::map<std::pair<uint32_t, uint64_t>, ProposalTables> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(m, seed);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  ProposalTables t;
  t.m = m;
  t.seed = seed;
  t.T.resize((size_t)kElems * m);
  t.W.resize((size_t)kElems * m);
  t.W32.resize((size_t)kElems * m);
  t.W16.resize((size_t)kElems * m);
  t.cls.resize((size_t)m * kElems);
  std::vector<TedRow> ted;
  double first_inv;
  ted_rows(m, ted, first_inv);
  const uint8_t* codes = canonical_codes();
  for (int x = 0; x < kElems; ++x)
    element_proposals(m, fmix64(codes[x], seed), ted, first_inv, &t.T[(size_t)x * m],
                      &t.W[(size_t)x * m]);
  for (size_t q = 0; q < t.W.size(); ++q) {
    t.W32[q] = (uint32_t)(t.W[q] ^ (t.W[q] >> 32));
    t.W16[q] = fold16(t.W32[q]);
  }
  for (uint32_t k = 0; k < m; ++k)
    for (int x = 0; x < kElems; ++x) {
      const uint16_t f = t.W16[(size_t)x * m + k];
      uint8_t c = 0xFF;
      if (f != 0)
        for (int y = 0; y <= x; ++y)
          if (t.W16[(size_t)y * m + k] == f) {
            c = (uint8_t)y;
            break;
          }
      t.cls[(size_t)k * kElems + x] = c;
    }
  t.sT.resize((size_t)kElems * m);
  t.sX.resize((size_t)kElems * m);
  std::vector<int> ord(kElems);
  for (uint32_t k = 0; k < m; ++k) {
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return t.T[(size_t)a * m + k] < t.T[(size_t)b * m + k];
    });
    for (int r = 0; r < kElems; ++r) {
      t.sT[(size_t)r * m + k] = t.T[(size_t)ord[r] * m + k];
      t.sX[(size_t)r * m + k] = (uint8_t)ord[r];
    }
  }
  return cache.emplace(key, std::move(t)).first->second;
}

void sv_tables(uint32_t m, double b0, std::vector<double>& svd, std::vector<float>& svf,
               std::vector<uint16_t>& level) {
  svd.resize(m + 1);
  svf.resize(m + 1);
  level.resize(m + 1);
  const double inv_m = 1.0 / (double)m;
  const bool rescale = b0 > 0.0 && b0 < 1.0;
  const double inv1mb0 = rescale ? 1.0 / (1.0 - b0) : 1.0;
  for (uint32_t c = 0; c <= m; ++c) {
    double j = (double)c * inv_m;
    if (rescale) {
      j = (j - b0) * inv1mb0;
      if (j < 0.0) j = 0.0;
    }
    if (j > 1.0 - 1e-6) j = 1.0 - 1e-6;
    svd[c] = j;
    svf[c] = (float)j;
  }
  // level: dense rank of the distinct float similarities (monotone in c)
  uint16_t lv = 0;
  for (uint32_t c = 0; c <= m; ++c) {
    if (c > 0 && svf[c] != svf[c - 1]) ++lv;
    level[c] = lv;
  }
}

void column_sums(const float* depth, size_t n, size_t S, double* sum) {
  // each column's sum runs over the rows in order (the reference's loop);
  // blocks of rows keep the rows in cache while every column passes them
  constexpr size_t kBlock = 4096;
  for (size_t r0 = 0; r0 < n; r0 += kBlock) {
    const size_t r1 = std::min(n, r0 + kBlock);
    for (size_t c = 0; c < S; ++c) {
      double v = sum[c];
      for (size_t r = r0; r < r1; ++r) v += depth[r * S + c];
      sum[c] = v;
    }
  }
}

void depth_features(const float* depth, size_t n, size_t S, double min_cv, float* ranked,
                    float* unit, float* raw, double* colnorm, uint8_t* anynz, bool parallel,
                    size_t stride) {
  if (stride == 0) stride = S;
  if (S > 1 && colnorm) {
    std::vector<double> sum(S, 0.0);
    column_sums(depth, n, S, sum.data());
    for (size_t c = 0; c < S; ++c) {
      const double mean = sum[c] / (double)std::max<size_t>(n, 1);
      colnorm[c] = (mean > 1e-12) ? 1.0 / mean : 0.0;
    }
  }
  // Row by row: the raw row, average ranks (ties share the mean position,
  // src/ranker.h:55-66), the centred unit row, is_nz.  Every output element
  // (padding included) is written by the thread that owns the row.
#pragma omp parallel if (parallel)
  {
    std::vector<uint32_t> order(S);
    std::vector<float> rk(S);
#pragma omp for schedule(static)
    for (size_t r = 0; r < n; ++r) {
      const float* row = depth + r * S;
      if (S > 1 && raw) {
        std::copy(row, row + S, raw + r * stride);
        std::fill(raw + r * stride + S, raw + (r + 1) * stride, 0.0f);
      }
      if (S > 1) {
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [row](uint32_t a, uint32_t b) { return row[a] < row[b]; });
        size_t c = 0;
        while (c < S) {
          size_t tie = 1;
          while (c + tie < S && row[order[c]] == row[order[c + tie]]) ++tie;
          const float v = (float)(2 * c + tie - 1) / 2 + 1;
          for (size_t q = 0; q < tie; ++q) rk[order[c + q]] = v;
          c += tie;
        }
      } else {
        std::copy(row, row + S, rk.begin());
      }
      if (ranked) std::copy(rk.begin(), rk.end(), ranked + r * S);
      if (unit) {
        float* u = unit + r * stride;
        std::fill(u, u + stride, 0.0f);
        if (S > 1) {
          double mean = 0.0;
          for (size_t k = 0; k < S; ++k) mean += rk[k];
          mean /= (double)S;
          double ss = 0.0;
          for (size_t k = 0; k < S; ++k) {
            const double d = (double)rk[k] - mean;
            // exact with or without fusion: ranks make d a multiple of 1/2
            ss = std::fma(d, d, ss);
          }
          if (ss > 0.0) {
            const double inv = 1.0 / std::sqrt(ss);
            for (size_t k = 0; k < S; ++k) u[k] = (float)(((double)rk[k] - mean) * inv);
          }
        }
      }
      if (anynz) {
        uint8_t any = 0;
        for (size_t c = 0; c < S && !any; ++c) any = (double)rk[c] > min_cv;
        anynz[r] = any;
      }
    }
  }
}

namespace {

// q rows of `unit` at g.scale (clamped to [-127, 127]; the clamp error counts
// in the error norm), folding the rows' norms into g's maxima.  Rows in
// parallel: each row's arithmetic is sequential, and maxima do not depend on
// the order they are taken in.
void quantize_rows(const float* unit, size_t n, uint32_t S, GateQuant& g, bool parallel,
                   const uint32_t* order, size_t stride) {
  g.q.assign(n * g.kpad, 0);
  double mq = g.max_qnorm, me = g.max_err, mu = g.max_unorm;
#pragma omp parallel for schedule(static) reduction(max : mq, me, mu) if (parallel)
  for (size_t r = 0; r < n; ++r) {
    double qq = 0, ee = 0, uu = 0;
    const float* row = unit + (order ? order[r] : r) * stride;
    for (uint32_t k = 0; k < S; ++k) {
      const double u = row[k];
      long v = g.scale > 0 ? std::lround(u * g.scale) : 0;
      v = std::max(-127L, std::min(127L, v));
      g.q[r * g.kpad + k] = (int8_t)v;
      const double x = g.scale > 0 ? (double)v / g.scale : 0.0;
      qq += x * x;
      ee += (u - x) * (u - x);
      uu += u * u;
    }
    mq = std::max(mq, std::sqrt(qq));
    me = std::max(me, std::sqrt(ee));
    mu = std::max(mu, std::sqrt(uu));
  }
  g.max_qnorm = mq;
  g.max_err = me;
  g.max_unorm = mu;
}

// max |u| over the n rows (S values each, stride apart)
double abs_max(const float* unit, size_t n, uint32_t S, size_t stride, bool parallel) {
  double amax = 0;
#pragma omp parallel for schedule(static) reduction(max : amax) if (parallel)
  for (size_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < S; ++k) amax = std::max(amax, (double)std::fabs(unit[r * stride + k]));
  return amax;
}

}  // namespace

void gate_quantize(const float* unit, size_t n, uint32_t S, GateQuant& g, bool parallel,
                   const uint32_t* order, size_t stride) {
  if (stride == 0) stride = S;
  g.S = S;
  g.kpad = 32 * ((S + 31) / 32);
  const double amax = abs_max(unit, n, S, stride, parallel);
  g.scale = amax > 0 ? 127.0 / amax : 0.0;
  g.max_qnorm = g.max_err = g.max_unorm = 0;
  quantize_rows(unit, n, S, g, parallel, order, stride);
}

void gate_quantize_more(const float* unit, size_t n, uint32_t S, GateQuant& g, bool parallel,
                        const uint32_t* order, size_t stride) {
  if (stride == 0) stride = S;
  g.S = S;
  g.kpad = 32 * ((S + 31) / 32);
  if (g.scale <= 0) {  // every earlier row is zero: any scale quantises them exactly
    const double amax = abs_max(unit, n, S, stride, parallel);
    g.scale = amax > 0 ? 127.0 / amax : 0.0;
  }
  quantize_rows(unit, n, S, g, parallel, order, stride);
}

void pair_pass_order(const float* unit, size_t n, uint32_t S, std::vector<uint32_t>& perm, bool parallel,
                     size_t stride) {
  if (stride == 0) stride = S;
  // (key << 32 | row): sorting these is the stable sort of the rows by key
  std::vector<uint64_t> kr(n), tmp(n);
  const uint32_t top = std::min<uint32_t>(3, S);
#pragma omp parallel for schedule(static) if (parallel)
  for (size_t r = 0; r < n; ++r) {
    const float* u = unit + r * stride;
    uint32_t k = 0;
    uint64_t used = 0;
    for (uint32_t t = 0; t < top; ++t) {
      uint32_t best = 0;
      bool have = false;
      for (uint32_t s = 0; s < S; ++s)
        if (!((used >> s) & 1u) && (!have || u[s] > u[best])) {
          best = s;
          have = true;
        }
      used |= 1ull << best;
      k = (k << 8) | best;
    }
    kr[r] = (uint64_t)k << 32 | r;
  }
  // least significant digit first over the key's 8 * top bits (the rows are
  // already in order): each pass is stable
  for (uint32_t pass = 0; pass < top; ++pass) {
    const uint32_t shift = 32 + 8 * pass;
    size_t cnt[257] = {0};
    for (size_t r = 0; r < n; ++r) ++cnt[((kr[r] >> shift) & 0xFF) + 1];
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (size_t r = 0; r < n; ++r) tmp[cnt[(kr[r] >> shift) & 0xFF]++] = kr[r];
    kr.swap(tmp);
  }
  perm.resize(n);
  for (size_t r = 0; r < n; ++r) perm[r] = (uint32_t)kr[r];
}

// Bound on |fl(u_i . u_j) - q_i . q_j / s^2| for every pair of rows of g.
static double gate_slack(const GateQuant& g) {
  // FP32 dot: each product passes through at most ceil(S/16) lane FMAs, the
  // acc0 + acc1 add and four reduction levels.
  const double u = std::ldexp(1.0, -24);
  const double depth = (double)((g.S + 15) / 16 + 6);
  const double gamma = depth * u / (1.0 - depth * u);
  // |dot(u_i,u_j) - dot(q_i,q_j)/s^2| <= 2 X E + E^2 (Cauchy-Schwarz), plus
  // a margin for the double arithmetic of this bound.
  return 2.0 * g.max_qnorm * g.max_err + g.max_err * g.max_err + gamma * g.max_unorm * g.max_unorm +
         1e-6;
}

int32_t gate_qcut_hi(const GateQuant& g, float thr) {
  if (g.scale <= 0) return INT32_MAX;
  const double cut = g.scale * g.scale * ((double)thr + gate_slack(g));
  if (cut <= 0.0 || cut >= (double)INT32_MAX) return INT32_MAX;
  return (int32_t)std::ceil(cut);
}

int32_t gate_qcut(const GateQuant& g, float thr) {
  if (g.scale <= 0) return INT32_MIN;
  const double slack = gate_slack(g);
  const double cut = g.scale * g.scale * ((double)thr - slack);
  if (cut <= (double)INT32_MIN + 1) return INT32_MIN;
  if (cut > (double)INT32_MAX) return INT32_MAX;
  return (int32_t)std::ceil(cut);
}

RandStream::RandStream(unsigned seed) {
  std::memset(&rd_, 0, sizeof(rd_));
  initstate_r(seed, state_, sizeof(state_), &rd_);
}
int RandStream::next() {
  int32_t r;
  random_r(&rd_, &r);
  return r;
}

namespace {
// calib_converge: walk p down from 999 until coverage is reached.
size_t converge(const std::vector<double>& maxsim, double coverage, RandStream& rs) {
  const size_t n = maxsim.size();
  size_t p = 999, prev_p = 1000;
  double cov = 0, prev_cov = 0;
  while (p > 50) {
    const double cut = (double)p / 1000.;
    const size_t on = (size_t)std::count_if(maxsim.begin(), maxsim.end(),
                                            [cut](double v) { return v >= cut; });
    cov = (double)on / n;
    if (cov >= coverage) {
      if (cov - coverage > coverage - prev_cov) p = prev_p;
      break;
    }
    prev_p = p;
    prev_cov = cov;
    const int r = rs.next() % 3;
    p -= (size_t)(p > 990 ? r + 1 : p > 900 ? r + 3 : r + 9);
  }
  return p;
}

void draw_prefix(std::vector<size_t>& v, size_t count, RandStream& rs) {
  size_t left = v.size();
  for (size_t s = 0; s < count; ++s, --left) {
    const size_t r = (size_t)rs.next() % left;
    std::swap(v[s], v[s + r]);
  }
}
}  // namespace

double calibrate_small(const std::vector<uint32_t>& rowmax, const std::vector<double>& svd,
                       double coverage, unsigned rand_seed) {
  const size_t n = rowmax.size();
  RandStream rs(rand_seed);
  std::vector<size_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  draw_prefix(idx, n, rs);  // consumes n draws; the order does not matter below
  std::vector<double> maxsim(n);
  for (size_t t = 0; t < n; ++t) maxsim[t] = n > 1 ? svd[rowmax[idx[t]]] : 0.0;
  return (double)converge(maxsim, coverage, rs);
}

double calibrate_fused(const std::vector<uint32_t>& rowmax, const std::vector<float>& svf,
                       double coverage, unsigned rand_seed) {
  const size_t n = rowmax.size();
  constexpr size_t R = 10, SAMP = 2500;
  RandStream rs(rand_seed);
  std::vector<int8_t> round_of(n, -1);
  std::vector<uint32_t> local_of(n, 0);
  for (size_t r = 0; r < R; ++r) {
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    draw_prefix(idx, SAMP, rs);
    for (size_t li = 0; li < SAMP; ++li)
      if (round_of[idx[li]] < 0) {  // a contig counts only in its lowest round
        round_of[idx[li]] = (int8_t)r;
        local_of[idx[li]] = (uint32_t)li;
      }
  }
  std::vector<std::vector<double>> maxsim(R, std::vector<double>(SAMP, 0.0));
  for (size_t c = 0; c < n; ++c)
    if (round_of[c] >= 0) maxsim[round_of[c]][local_of[c]] = (double)svf[rowmax[c]];
  double sum_p = 0;
  for (size_t r = 0; r < R; ++r) {
    double p = (double)converge(maxsim[r], coverage, rs);
    if (p < 51.0) p = 50.0;
    sum_p += p;
    if (r == 1 && sum_p / 2 < 51.0) {
      sum_p = 50.0 * R;
      break;
    }
  }
  return sum_p / (double)R;
}

void calibration_plan(uint32_t n, unsigned seed, int regime, CalibPlan& plan) {
  plan.n = n;
  plan.seed = seed;
  plan.regime = regime;
  plan.members.clear();
  plan.round_off.assign(1, 0);
  RandStream rs(seed);
  if (regime == 1) {
    for (uint32_t t = 0; t < n; ++t) rs.next();
  } else {
    constexpr size_t R = 10, SAMP = 2500;
    std::vector<int8_t> round_of(n, -1);
    std::vector<size_t> idx(n);
    for (size_t r = 0; r < R; ++r) {
      std::iota(idx.begin(), idx.end(), 0);
      draw_prefix(idx, SAMP, rs);
      for (size_t li = 0; li < SAMP; ++li)
        if (round_of[idx[li]] < 0) round_of[idx[li]] = (int8_t)r;
    }
    for (size_t r = 0; r < R; ++r) {
      for (uint32_t c = 0; c < n; ++c)
        if (round_of[c] == (int8_t)r) plan.members.push_back(c);
      plan.round_off.push_back((uint32_t)plan.members.size());
    }
  }
  plan.rnd.resize(kCalibRand);
  for (auto& v : plan.rnd) v = rs.next();
}

void baseline_pairs(size_t n, std::vector<uint32_t>& pi, std::vector<uint32_t>& pj) {
  pi.clear();
  pj.clear();
  if (n < 4) return;
  std::mt19937_64 rng(0x9E3779B97F4A7C15ULL);
  std::uniform_int_distribution<size_t> uni(0, n - 1);
  for (size_t s = 0; s < 30000; ++s) {
    const size_t i = uni(rng), j = uni(rng);
    if (i == j) continue;
    pi.push_back((uint32_t)i);
    pj.push_back((uint32_t)j);
  }
}

double baseline_median(const std::vector<uint32_t>& counts, uint32_t m) {
  if (counts.empty()) return 0.0;
  const double inv_m = 1.0 / (double)m;
  std::vector<float> vals(counts.size());
  for (size_t t = 0; t < counts.size(); ++t) vals[t] = (float)((double)counts[t] * inv_m);
  std::nth_element(vals.begin(), vals.begin() + vals.size() / 2, vals.end());
  return (double)vals[vals.size() / 2];
}

}  // namespace rbg
