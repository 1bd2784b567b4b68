// librbg: frozen-core recruitment (rbg_recruit_frozen_cores), after
// recruit_to_frozen_cores (src/impl/rb_split.cpp:729-919).
//
// The host builds the cores in the reference's arithmetic (member sums in
// each bin's order, the closed-form mean Spearman threshold) and the small
// contigs' unit rows (rank transform as for the large contigs); the device
// scores every candidate contig against every core (rbg_recruit.hip).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/rbg.h"
#include "rbg_build.h"
#include "rbg_host.h"
#include "rbg_internal.h"
#include "rbg_kernels.h"

using namespace rbg;

namespace {

// sum_k u_k^2 in sample order, as the reference's loop (contracted to one
// fused multiply-add per sample).
double norm2_of(const float* u, uint32_t S) {
  double s = 0.0;
  for (uint32_t k = 0; k < S; ++k) s = std::fma((double)u[k], (double)u[k], s);
  return s;
}

// sum_k v_k^2 over the core sum, as the reference build evaluates
// `for (double v : sum) sum_norm2 += v * v` (GCC 11, -O3 -march=native): from
// S = 4 on, the products of the first 4 * floor(S / 4) entries are formed four
// at a time and rounded, then added in order; the last S mod 4 entries are
// fused multiply-adds (all of them below S = 4).
double sum_sq(const double* v, uint32_t S) {
  const uint32_t q = S >= 4 ? (S & ~3u) : 0u;
  double s = 0.0;
  for (uint32_t k = 0; k < q; ++k) {
    const double p = v[k] * v[k];
    s = s + p;
  }
  for (uint32_t k = q; k < S; ++k) s = std::fma(v[k], v[k], s);
  return s;
}

}  // namespace

extern "C" rbg_status rbg_recruit_frozen_cores(rbg_ctx* ctx, uint32_t S, uint32_t n_large, const float* unit_large,
                                               const uint64_t* len_large, uint32_t n_small, const float* depth_small,
                                               const uint64_t* len_small, uint32_t n_bins, const int32_t* bin_id,
                                               const uint64_t* bin_off, const uint64_t* members, uint64_t core_floor,
                                               uint64_t min_cs, int32_t* assign_large, int32_t* assign_small,
                                               rbg_recruit_stats* stats) {
  if (!ctx || (n_large && (!assign_large || !len_large)) || (n_small && (!assign_small || !len_small)) ||
      (n_bins && (!bin_id || !bin_off || !members)))
    return RBG_ERR_ARG;
  rbg_recruit_stats st{};
  for (uint32_t c = 0; c < n_large; ++c) assign_large[c] = -1;
  for (uint32_t s = 0; s < n_small; ++s) assign_small[s] = -1;
  if (S > 64) return ctx->fail(RBG_ERR_UNSUPPORTED, "frozen-core recruitment supports S <= 64 samples");
  if (S >= 2 && ((n_large && !unit_large) || (n_small && !depth_small))) return RBG_ERR_ARG;
  const uint64_t ntot = (uint64_t)n_large + n_small;
  for (uint32_t b = 0; b < n_bins; ++b) {
    if (bin_off[b + 1] < bin_off[b]) return ctx->fail(RBG_ERR_ARG, "bin offsets must not decrease");
    for (uint64_t t = bin_off[b]; t < bin_off[b + 1]; ++t)
      if (members[t] >= ntot) return ctx->fail(RBG_ERR_ARG, "bin member out of range");
  }
  // no recruitment without two samples or without bins (the reference returns)
  if (S < 2 || n_bins == 0) {
    if (stats) *stats = st;
    return RBG_OK;
  }
  // cores, by ascending bin id
  std::vector<uint32_t> ids(n_bins);
  std::iota(ids.begin(), ids.end(), 0u);
  std::sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return bin_id[a] < bin_id[b]; });
  for (uint32_t t = 0; t < n_bins; ++t) {
    // (the reference's scan takes a negative id for "no match yet")
    if (bin_id[ids[t]] < 0) return ctx->fail(RBG_ERR_ARG, "bin ids must be non-negative");
    if (t && bin_id[ids[t]] == bin_id[ids[t - 1]]) return ctx->fail(RBG_ERR_ARG, "bin ids must be distinct");
  }
  const uint64_t need = std::max<uint64_t>(2, min_cs);
  std::vector<double> cent, thr;
  std::vector<int32_t> core_id;
  std::vector<double> sum(S);
  for (uint32_t b : ids) {
    uint64_t span = 0;
    for (uint64_t t = bin_off[b]; t < bin_off[b + 1]; ++t) {
      const uint64_t c = members[t];
      span += c < n_large ? len_large[c] : len_small[c - n_large];
    }
    if (span < core_floor) continue;
    std::fill(sum.begin(), sum.end(), 0.0);
    double self_sum = 0.0;
    uint64_t valid = 0;
    for (uint64_t t = bin_off[b]; t < bin_off[b + 1]; ++t) {
      const uint64_t c = members[t];
      if (c >= n_large) continue;
      const float* u = unit_large + c * S;
      const double nn = norm2_of(u, S);
      if (!(nn > 0.0)) continue;
      for (uint32_t k = 0; k < S; ++k) sum[k] += u[k];
      self_sum += nn;
      ++valid;
    }
    if (valid < need) continue;
    const double sum_norm2 = sum_sq(sum.data(), S);
    double tau = (sum_norm2 - self_sum) / ((double)valid * (double)(valid - 1));
    tau = std::max(-1.0, std::min(1.0, tau));
    for (uint32_t k = 0; k < S; ++k) cent.push_back(sum[k] / (double)valid);
    thr.push_back(tau);
    core_id.push_back(bin_id[b]);
  }
  st.cores = core_id.size();
  if (core_id.empty()) {
    if (stats) *stats = st;
    return RBG_OK;
  }
  // candidates: large contigs in no bin, then small contigs with a unit row
  std::vector<uint8_t> binned(n_large, 0);
  for (uint64_t t = 0; t < bin_off[n_bins] - bin_off[0]; ++t) {
    const uint64_t c = members[bin_off[0] + t];
    if (c < n_large) binned[c] = 1;
  }
  std::vector<uint32_t> lrow;
  for (uint32_t c = 0; c < n_large; ++c)
    if (!binned[c]) lrow.push_back(c);
  std::vector<float> su((size_t)n_small * S, 0.0f);
  std::vector<uint8_t> sok(n_small, 0);
  if (n_small) {
    // the small contigs' ranks and unit rows, as for the large contigs
    depth_features(depth_small, n_small, S, 0.0, nullptr, su.data(), nullptr, nullptr, nullptr);
    for (uint32_t s = 0; s < n_small; ++s) {
      // ss > 0 exactly when the ranks are not all equal, i.e. the unit row is not zero
      bool any = false;
      for (uint32_t k = 0; k < S && !any; ++k) any = su[(size_t)s * S + k] != 0.0f;
      sok[s] = any;
    }
  }
  std::vector<uint32_t> srow;
  for (uint32_t s = 0; s < n_small; ++s)
    if (sok[s]) srow.push_back(s);
  const uint64_t nc = lrow.size() + srow.size();
  std::vector<float> cu(nc * S);
  for (size_t t = 0; t < lrow.size(); ++t)
    std::copy(unit_large + (size_t)lrow[t] * S, unit_large + (size_t)(lrow[t] + 1) * S, cu.begin() + t * S);
  for (size_t t = 0; t < srow.size(); ++t)
    std::copy(su.begin() + (size_t)srow[t] * S, su.begin() + (size_t)(srow[t] + 1) * S,
              cu.begin() + (lrow.size() + t) * S);
  if (nc > 0xFFFFFFFFull) return ctx->fail(RBG_ERR_UNSUPPORTED, "too many candidate contigs");
  std::vector<int32_t> oid(nc, -1);
  std::vector<uint8_t> oamb(nc, 0);
  if (nc) {
    RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf du, dc, dt, di, dout, damb;
    RBG_TRY_HIP(ctx, du.ensure(cu.size() * 4));
    RBG_TRY_HIP(ctx, dc.ensure(cent.size() * 8));
    RBG_TRY_HIP(ctx, dt.ensure(thr.size() * 8));
    RBG_TRY_HIP(ctx, di.ensure(core_id.size() * 4));
    RBG_TRY_HIP(ctx, dout.ensure(nc * 4));
    RBG_TRY_HIP(ctx, damb.ensure(nc));
    hipStream_t sm = ctx->stream;
    RBG_TRY_HIP(ctx, hipMemcpyAsync(du.p, cu.data(), cu.size() * 4, hipMemcpyHostToDevice, sm));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(dc.p, cent.data(), cent.size() * 8, hipMemcpyHostToDevice, sm));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(dt.p, thr.data(), thr.size() * 8, hipMemcpyHostToDevice, sm));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(di.p, core_id.data(), core_id.size() * 4, hipMemcpyHostToDevice, sm));
    RBG_TRY_HIP(ctx, launch_recruit_match((const float*)du.p, (uint32_t)nc, S, (const double*)dc.p,
                                          (const double*)dt.p, (const int32_t*)di.p, (uint32_t)core_id.size(),
                                          (int32_t*)dout.p, (uint8_t*)damb.p, sm));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(oid.data(), dout.p, nc * 4, hipMemcpyDeviceToHost, sm));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(oamb.data(), damb.p, nc, hipMemcpyDeviceToHost, sm));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(sm));
  }
  for (size_t t = 0; t < lrow.size(); ++t) {
    assign_large[lrow[t]] = oid[t];
    if (oid[t] >= 0) ++st.long_assigned;
    else if (oamb[t]) ++st.long_ambiguous;
    else ++st.long_unmatched;
  }
  st.short_unmatched = n_small - srow.size();  // constant profiles (no unit row)
  for (size_t t = 0; t < srow.size(); ++t) {
    const size_t x = lrow.size() + t;
    assign_small[srow[t]] = oid[x];
    if (oid[x] >= 0) ++st.short_assigned;
    else if (oamb[x]) ++st.short_ambiguous;
    else ++st.short_unmatched;
  }
  if (stats) *stats = st;
  return RBG_OK;
}
