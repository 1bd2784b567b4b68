// librbg: context, device-resident assemblies and the graph pipeline behind
// the C ABI of include/rbg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <numeric>
#include <vector>

#include "../../include/rbg.h"
#include "rbg_host.h"
#include "rbg_kernels.h"
#include "rbg_internal.h"

using namespace rbg;

namespace rbg {

const char* kStatusText[] = {"ok", "HIP error", "invalid argument", "out of memory",
                             "unsupported configuration", "over HBM budget"};

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
hipError_t DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && p) return hipSuccess;
  release();
  if (nbytes == 0) nbytes = 1;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = nbytes;
  return hipSuccess;
}

}  // namespace rbg

struct rbg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  struct Tables {
    uint32_t m;
    uint64_t seed;
    DevBuf T, cls, codes, sT, sX;
  };
  std::vector<std::unique_ptr<Tables>> tables;
  hipEvent_t ev[12] = {};
  rbg_status fail(rbg_status st, const std::string& msg) {
    err = msg;
    return st;
  }
  rbg_status hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return RBG_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return RBG_ERR_HIP;
  }
};

#define RBG_TRY_HIP(ctx, expr)                                  \
  do {                                                          \
    const hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return (ctx)->hip(_e, #expr);         \
  } while (0)
#define RBG_TRY(expr)                      \
  do {                                     \
    const rbg_status _s = (expr);          \
    if (_s != RBG_OK) return _s;           \
  } while (0)

namespace {

uint32_t words_for(uint32_t m) {
  uint32_t w = (m + 31) / 32, p = 1;
  while (p < w) p <<= 1;
  return p;
}

rbg_status get_tables(rbg_ctx* ctx, uint32_t m, uint64_t seed, rbg_ctx::Tables** out) {
  for (auto& t : ctx->tables)
    if (t->m == m && t->seed == seed) {
      *out = t.get();
      return RBG_OK;
    }
  const ProposalTables& pt = proposal_tables(m, seed);
  auto t = std::make_unique<rbg_ctx::Tables>();
  t->m = m;
  t->seed = seed;
  RBG_TRY_HIP(ctx, t->T.ensure(pt.T.size() * sizeof(double)));
  RBG_TRY_HIP(ctx, t->cls.ensure(pt.cls.size()));
  RBG_TRY_HIP(ctx, t->codes.ensure(kElems));
  RBG_TRY_HIP(ctx, hipMemcpy(t->T.p, pt.T.data(), pt.T.size() * sizeof(double), hipMemcpyHostToDevice));
  RBG_TRY_HIP(ctx, hipMemcpy(t->cls.p, pt.cls.data(), pt.cls.size(), hipMemcpyHostToDevice));
  RBG_TRY_HIP(ctx, hipMemcpy(t->codes.p, canonical_codes(), kElems, hipMemcpyHostToDevice));
  RBG_TRY_HIP(ctx, t->sT.ensure(pt.sT.size() * sizeof(double)));
  RBG_TRY_HIP(ctx, t->sX.ensure(pt.sX.size()));
  RBG_TRY_HIP(ctx, hipMemcpy(t->sT.p, pt.sT.data(), pt.sT.size() * sizeof(double), hipMemcpyHostToDevice));
  RBG_TRY_HIP(ctx, hipMemcpy(t->sX.p, pt.sX.data(), pt.sX.size(), hipMemcpyHostToDevice));
  *out = t.get();
  ctx->tables.push_back(std::move(t));
  return RBG_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device-resident assembly: contigs, depth features and every intermediate
// of the pipeline.  Buffers persist across calls so a resident assembly can
// be rebuilt or queried without new allocations.
// ---------------------------------------------------------------------------
struct rbg_asm {
  uint32_t n = 0, S = 0;
  uint64_t seq_bytes = 0;
  DevBuf seqs, offs;
  // depth features (device) and their host copies
  DevBuf unit_pad, raw_pad, unit, raw, colnorm, anynz, q8, perm;
  uint32_t NC = 1;
  bool has_anynz = false;
  std::vector<uint8_t> anynz_h;
  GateQuant gq;  // bound statistics of q8 (the int8 rows live on the device)
  // pipeline intermediates
  DevBuf bscal, svd, cal_rnd, cal_members, cal_roff;
  uint32_t np = 0;            // b0 pair draws (uploaded once)
  bool pairs_ready = false;
  int cal_regime = 0;         // calibration plan cached for (n, seed, regime)
  unsigned cal_seed = 0;
  DevBuf wx, planes, pi, pj, pcnt, level, svf, mat, cand, nsurv, keep, fence, scnt, deg, rowmax, off, cursor,
      list, topk_id, topk_cnt, topk_len, row_edges, row_off, efrom, eto, ecnt, escomp, escore,
      scan_tmp, tiles;
  uint32_t ntiles = 0;
  uint64_t last_edges = 0;  // edge count of the last build (device buffers)
};

namespace {

rbg_status asm_upload(rbg_ctx* ctx, rbg_asm* a, const char* seqs, const uint64_t* offsets,
                      uint32_t n, const float* depth, uint32_t S, double min_cv) {
  a->n = n;
  a->S = S;
  a->seq_bytes = offsets[n] - offsets[0];
  std::vector<uint64_t> rel(n + 1);
  for (uint32_t r = 0; r <= n; ++r) rel[r] = offsets[r] - offsets[0];
  // The sketch kernel reads whole 16-byte chunks: keep 64 readable bytes
  // past the last contig (their contents are never counted).
  RBG_TRY_HIP(ctx, a->seqs.ensure(a->seq_bytes + 64));
  RBG_TRY_HIP(ctx, a->offs.ensure((n + 1) * sizeof(uint64_t)));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(a->seqs.p, seqs + offsets[0], a->seq_bytes, hipMemcpyHostToDevice,
                                  ctx->stream));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(a->offs.p, rel.data(), (n + 1) * sizeof(uint64_t),
                                  hipMemcpyHostToDevice, ctx->stream));
  a->has_anynz = false;
  if (S > 0) {
    std::vector<float> unit((size_t)n * S), raw((size_t)n * S);
    std::vector<double> colnorm(S, 0.0);
    a->anynz_h.assign(n, 0);
    depth_features(depth, n, S, min_cv, nullptr, unit.data(), raw.data(), colnorm.data(),
                   a->anynz_h.data());
    a->has_anynz = true;
    a->NC = (S + 15) / 16;
    std::vector<float> pad((size_t)n * 16 * a->NC, 0.0f), rpad(pad.size(), 0.0f);
    for (uint32_t r = 0; r < n; ++r) {
      std::memcpy(&pad[(size_t)r * 16 * a->NC], &unit[(size_t)r * S], S * sizeof(float));
      std::memcpy(&rpad[(size_t)r * 16 * a->NC], &raw[(size_t)r * S], S * sizeof(float));
    }
    RBG_TRY_HIP(ctx, a->unit_pad.ensure(pad.size() * sizeof(float)));
    RBG_TRY_HIP(ctx, a->raw_pad.ensure(rpad.size() * sizeof(float)));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->raw_pad.p, rpad.data(), rpad.size() * sizeof(float),
                                    hipMemcpyHostToDevice, ctx->stream));
    RBG_TRY_HIP(ctx, a->unit.ensure(unit.size() * sizeof(float)));
    RBG_TRY_HIP(ctx, a->raw.ensure(raw.size() * sizeof(float)));
    RBG_TRY_HIP(ctx, a->colnorm.ensure(S * sizeof(double)));
    RBG_TRY_HIP(ctx, a->anynz.ensure(n));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->unit_pad.p, pad.data(), pad.size() * sizeof(float),
                                    hipMemcpyHostToDevice, ctx->stream));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->unit.p, unit.data(), unit.size() * sizeof(float),
                                    hipMemcpyHostToDevice, ctx->stream));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->raw.p, raw.data(), raw.size() * sizeof(float),
                                    hipMemcpyHostToDevice, ctx->stream));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->colnorm.p, colnorm.data(), S * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(a->anynz.p, a->anynz_h.data(), n, hipMemcpyHostToDevice,
                                    ctx->stream));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (S > 1 && S <= 64) {
      // The pair pass visits rows in abundance order (pair_pass_order); the
      // int8 rows are stored in that order and candidates map back through perm.
      std::vector<uint32_t> perm;
      pair_pass_order(unit.data(), n, S, perm);
      std::vector<float> sorted((size_t)n * S);
      for (uint32_t r = 0; r < n; ++r)
        std::memcpy(&sorted[(size_t)r * S], &unit[(size_t)perm[r] * S], S * sizeof(float));
      gate_quantize(sorted.data(), n, S, a->gq);
      RBG_TRY_HIP(ctx, a->q8.ensure(a->gq.q.size()));
      RBG_TRY_HIP(ctx, hipMemcpy(a->q8.p, a->gq.q.data(), a->gq.q.size(), hipMemcpyHostToDevice));
      std::vector<int8_t>().swap(a->gq.q);
      RBG_TRY_HIP(ctx, a->perm.ensure((size_t)n * sizeof(uint32_t)));
      RBG_TRY_HIP(ctx, hipMemcpy(a->perm.p, perm.data(), (size_t)n * sizeof(uint32_t),
                                 hipMemcpyHostToDevice));
    }
  }
  if (!(S > 1 && S <= 64)) {
    // no abundance rows: an all-zero image (the gate is then off)
    a->NC = 1;
    a->gq = GateQuant{};
    a->gq.kpad = 32;
    RBG_TRY_HIP(ctx, a->q8.ensure((size_t)n * 32));
    RBG_TRY_HIP(ctx, hipMemsetAsync(a->q8.p, 0, (size_t)n * 32, ctx->stream));
    RBG_TRY_HIP(ctx, a->unit_pad.ensure((size_t)n * 16 * sizeof(float)));
    RBG_TRY_HIP(ctx, hipMemsetAsync(a->unit_pad.p, 0, (size_t)n * 16 * sizeof(float), ctx->stream));
    std::vector<uint32_t> ident(n);
    std::iota(ident.begin(), ident.end(), 0u);
    RBG_TRY_HIP(ctx, a->perm.ensure((size_t)n * sizeof(uint32_t)));
    RBG_TRY_HIP(ctx, hipMemcpy(a->perm.p, ident.data(), (size_t)n * sizeof(uint32_t),
                               hipMemcpyHostToDevice));
  }
  // work items of the fused pair pass (256-row stripes x up to 16 tiles)
  std::vector<uint32_t> tl;
  gate_work_items(n, tl);
  a->ntiles = (uint32_t)(tl.size() / 2);
  RBG_TRY_HIP(ctx, a->tiles.ensure(tl.size() * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(a->tiles.p, tl.data(), tl.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RBG_OK;
}

rbg_status run_sketch(rbg_ctx* ctx, rbg_asm* a, uint32_t m, uint64_t seed, bool want_wx) {
  rbg_ctx::Tables* t;
  RBG_TRY(get_tables(ctx, m, seed, &t));
  const uint32_t W = words_for(m);
  RBG_TRY_HIP(ctx, a->planes.ensure((size_t)a->n * 8 * W * sizeof(uint32_t)));
  const uint32_t mpad = 32 * W;
  if (want_wx) RBG_TRY_HIP(ctx, a->wx.ensure((size_t)a->n * mpad));
  RBG_TRY_HIP(ctx, launch_sketch((const uint8_t*)a->seqs.p, (const uint64_t*)a->offs.p, a->n, m, W,
                                 (const double*)t->sT.p, (const uint8_t*)t->sX.p, (const uint8_t*)t->cls.p,
                                 (const uint8_t*)t->codes.p, want_wx ? (uint8_t*)a->wx.p : nullptr,
                                 mpad, (uint32_t*)a->planes.p, ctx->stream));
  return RBG_OK;
}

rbg_status ensure_baseline_pairs(rbg_ctx* ctx, rbg_asm* a) {
  if (a->pairs_ready) return RBG_OK;
  std::vector<uint32_t> pi, pj;
  baseline_pairs(a->n, pi, pj);
  a->np = (uint32_t)pi.size();
  const size_t nb = std::max<size_t>(a->np, 1) * sizeof(uint32_t);
  RBG_TRY_HIP(ctx, a->pi.ensure(nb));
  RBG_TRY_HIP(ctx, a->pj.ensure(nb));
  RBG_TRY_HIP(ctx, a->pcnt.ensure(nb));
  if (a->np) {
    RBG_TRY_HIP(ctx, hipMemcpy(a->pi.p, pi.data(), a->np * 4, hipMemcpyHostToDevice));
    RBG_TRY_HIP(ctx, hipMemcpy(a->pj.p, pj.data(), a->np * 4, hipMemcpyHostToDevice));
  }
  a->pairs_ready = true;
  return RBG_OK;
}

rbg_status ensure_calib_plan(rbg_ctx* ctx, rbg_asm* a, int regime, unsigned seed) {
  if (a->cal_regime == regime && a->cal_seed == seed) return RBG_OK;
  CalibPlan plan;
  calibration_plan(a->n, seed, regime, plan);
  RBG_TRY_HIP(ctx, a->cal_rnd.ensure(kCalibRand * sizeof(int32_t)));
  RBG_TRY_HIP(ctx, a->cal_members.ensure(std::max<size_t>(plan.members.size(), 1) * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, a->cal_roff.ensure(11 * sizeof(uint32_t)));
  plan.round_off.resize(11, plan.round_off.back());
  RBG_TRY_HIP(ctx, hipMemcpy(a->cal_rnd.p, plan.rnd.data(), kCalibRand * 4, hipMemcpyHostToDevice));
  if (!plan.members.empty())
    RBG_TRY_HIP(ctx, hipMemcpy(a->cal_members.p, plan.members.data(), plan.members.size() * 4,
                               hipMemcpyHostToDevice));
  RBG_TRY_HIP(ctx, hipMemcpy(a->cal_roff.p, plan.round_off.data(), 11 * 4, hipMemcpyHostToDevice));
  a->cal_regime = regime;
  a->cal_seed = seed;
  return RBG_OK;
}

// Fused regime: int8 MFMA prefilter of the abundance gate (one host sync for
// the candidate count), exact gate + register compare in place, per-row
// candidate lists (off, list).
rbg_status fused_pair_pass(rbg_ctx* ctx, rbg_asm* a, const rbg_graph_params* P, uint32_t S,
                           double w_comp, uint32_t W, const uint8_t* anynz) {
  const uint32_t n = a->n;
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  // abundance-first gate (src/impl/rb_graph.cpp:693-706)
  const bool gate = S > 1 && w_comp < 1.0 && w_comp < P->min_edge_weight;
  if (gate && S > 64) return ctx->fail(RBG_ERR_UNSUPPORTED, "abundance gate needs S <= 64");
  const float thr = gate ? (float)((P->min_edge_weight - w_comp) / (1.0 - w_comp) - 1e-4)
                         : -__builtin_huge_valf();
  const uint32_t NC = a->NC;
  const uint32_t acc1_end = gate ? 2 * (S / 32) : 0;
  // prefilter cut: with the gate off every pair passes (cut = INT32_MIN, thr = -inf)
  const int32_t qcut = gate ? gate_qcut(a->gq, thr) : INT32_MIN;
  const uint64_t total_pairs = (uint64_t)n * (n - 1) / 2;
  uint64_t cap = std::min<uint64_t>(total_pairs, std::max<uint64_t>(1ull << 26, 2048ull * n));
  RBG_TRY_HIP(ctx, a->nsurv.ensure(2 * sizeof(unsigned long long)));
  unsigned long long* counters = (unsigned long long*)a->nsurv.p;  // [0] candidates
  uint64_t nc = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    RBG_TRY_HIP(ctx, a->cand.ensure(std::max<uint64_t>(cap, 1) * sizeof(uint64_t)));
    RBG_TRY_HIP(ctx, hipMemsetAsync(counters, 0, 2 * sizeof(unsigned long long), st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[2], st));
    RBG_TRY_HIP(ctx, launch_gate_mfma((const int8_t*)a->q8.p, a->gq.kpad, n, qcut,
                                      (const uint32_t*)a->perm.p, (const uint32_t*)a->tiles.p,
                                      a->ntiles, (uint64_t*)a->cand.p, counters, cap, st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[6], st));
    unsigned long long got = 0;
    RBG_TRY_HIP(ctx, hipMemcpyAsync(&got, counters, sizeof(got), hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
    nc = got;
    if (nc <= cap) break;
    cap = nc;  // rerun with room for every candidate
  }
  // exact FP32 gate and register compare of the candidates, in place
  RBG_TRY_HIP(ctx, a->deg.ensure((size_t)n * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, a->cursor.ensure((size_t)n * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, a->off.ensure((size_t)(n + 1) * sizeof(uint64_t)));
  RBG_TRY_HIP(ctx, hipMemsetAsync(a->deg.p, 0, (size_t)n * 4, st));
  RBG_TRY_HIP(ctx, hipMemsetAsync(a->cursor.p, 0, (size_t)n * 4, st));
  RBG_TRY_HIP(ctx, hipEventRecord(ev[7], st));
  RBG_TRY_HIP(ctx, launch_gate_compare((uint64_t*)a->cand.p, nc, (const float*)a->unit_pad.p, NC,
                                       anynz, thr, acc1_end, (const uint32_t*)a->planes.p, W,
                                       (uint32_t*)a->deg.p, st));
  RBG_TRY_HIP(ctx, hipEventRecord(ev[8], st));
  size_t tmp_bytes = 0;
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(nullptr, nullptr, n, nullptr, tmp_bytes, st));
  RBG_TRY_HIP(ctx, a->scan_tmp.ensure(tmp_bytes));
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)a->deg.p, (uint64_t*)a->off.p, n,
                                             a->scan_tmp.p, tmp_bytes, st));
  // off[n] = 2 x survivors <= 2 x candidates
  RBG_TRY_HIP(ctx, a->list.ensure(std::max<uint64_t>(2 * nc, 1) * sizeof(uint64_t)));
  RBG_TRY_HIP(ctx, launch_scatter((const uint64_t*)a->cand.p, nc, (const uint64_t*)a->off.p,
                                  (uint32_t*)a->cursor.p, (uint64_t*)a->list.p, st));
  return RBG_OK;
}

rbg_status build_graph(rbg_ctx* ctx, rbg_asm* a, const rbg_graph_params* P, rbg_graph_result* out,
                       rbg_graph_diag* diag) {
  std::memset(out, 0, sizeof(*out));
  const uint32_t n = a->n, S = a->S, m = P->m, K = P->max_edges;
  if (n == 0) return ctx->fail(RBG_ERR_ARG, "empty assembly");
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (K == 0 || K > 1024) return ctx->fail(RBG_ERR_UNSUPPORTED, "max-edges must be in [1, 1024]");
  if (n >= (1u << 24)) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^24 contigs");
  const int regime = P->regime ? P->regime : (n > 25000 ? 2 : 1);
  if (regime == 2 && n < 2500) return ctx->fail(RBG_ERR_ARG, "fused regime needs n >= 2500");
  const uint32_t W = words_for(m);
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  out->regime = regime;
  // composition weight: auto_w_comp(S) (src/rabbitbin.cpp:470-473)
  const double w_comp = S <= 2 ? 1.0 : 0.0;

  RBG_TRY_HIP(ctx, hipEventRecord(ev[0], st));
  RBG_TRY(run_sketch(ctx, a, m, P->pmh_seed, diag && diag->win16));
  RBG_TRY_HIP(ctx, hipEventRecord(ev[1], st));
  // b0 and the sv tables on the device (the pair draws depend on n only)
  RBG_TRY(ensure_baseline_pairs(ctx, a));
  RBG_TRY_HIP(ctx, a->bscal.ensure(sizeof(BuildScalars)));
  BuildScalars* bs = (BuildScalars*)a->bscal.p;
  RBG_TRY_HIP(ctx, a->svd.ensure((m + 1) * sizeof(double)));
  RBG_TRY_HIP(ctx, a->svf.ensure((m + 1) * sizeof(float)));
  RBG_TRY_HIP(ctx, a->level.ensure((m + 1) * sizeof(uint16_t)));
  if (a->np)
    RBG_TRY_HIP(ctx, launch_list_counts((const uint32_t*)a->planes.p, W, (const uint32_t*)a->pi.p,
                                        (const uint32_t*)a->pj.p, a->np, (uint32_t*)a->pcnt.p, st));
  RBG_TRY_HIP(ctx, launch_baseline_tables((const uint32_t*)a->pcnt.p, a->np, m, bs,
                                          (double*)a->svd.p, (float*)a->svf.p,
                                          (uint16_t*)a->level.p, st));
  // level bins above the number of distinct similarities stay empty
  const uint32_t nlevels = m + 1;
  // +16 ids: the mutual test reads whole 16-id segments of a row's list
  RBG_TRY_HIP(ctx, a->topk_id.ensure(((size_t)n * K + 16) * sizeof(uint32_t)));
  const uint32_t FW = fence_words(K);
  RBG_TRY_HIP(ctx, a->fence.ensure((size_t)n * FW * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, a->topk_cnt.ensure((size_t)n * K * sizeof(uint16_t)));
  RBG_TRY_HIP(ctx, a->topk_len.ensure((size_t)n * sizeof(uint32_t)));
  RBG_TRY_HIP(ctx, a->rowmax.ensure((size_t)n * sizeof(uint32_t)));
  const uint8_t* anynz = (S > 0 && a->has_anynz) ? (const uint8_t*)a->anynz.p : nullptr;

  RBG_TRY_HIP(ctx, hipEventRecord(ev[2], st));
  if (regime == 1) {
    RBG_TRY_HIP(ctx, a->mat.ensure((size_t)n * n * sizeof(uint16_t)));
    RBG_TRY_HIP(ctx, launch_dense_counts((const uint32_t*)a->planes.p, W, n, (uint16_t*)a->mat.p, st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[6], st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[7], st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[8], st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[3], st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[9], st));
    RBG_TRY_HIP(ctx, launch_select_dense((const uint16_t*)a->mat.p, anynz, n,
                                         (const uint16_t*)a->level.p, nlevels, K,
                                         (uint32_t*)a->topk_id.p, (uint16_t*)a->topk_cnt.p,
                                         (uint32_t*)a->topk_len.p, (uint32_t*)a->rowmax.p,
                                         (uint32_t*)a->fence.p, FW, st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[10], st));
  } else {
    RBG_TRY(fused_pair_pass(ctx, a, P, S, w_comp, W, anynz));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[3], st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[9], st));
    RBG_TRY_HIP(ctx, launch_select_list((const uint64_t*)a->list.p, (const uint64_t*)a->off.p, n,
                                        (const uint16_t*)a->level.p, nlevels, K,
                                        (uint32_t*)a->topk_id.p, (uint16_t*)a->topk_cnt.p,
                                        (uint32_t*)a->topk_len.p, (uint32_t*)a->rowmax.p,
                                        (uint32_t*)a->fence.p, FW, st));
    RBG_TRY_HIP(ctx, hipEventRecord(ev[10], st));
  }
  // calibration on the device: the glibc rand() draws are prepared per (n, seed)
  RBG_TRY(ensure_calib_plan(ctx, a, regime, P->rand_seed));
  RBG_TRY_HIP(ctx, launch_calibrate((const uint32_t*)a->rowmax.p, n, m, (const double*)a->svd.p,
                                    (const float*)a->svf.p, (const int32_t*)a->cal_rnd.p, kCalibRand,
                                    (const uint32_t*)a->cal_members.p,
                                    (const uint32_t*)a->cal_roff.p, regime, P->coverage, bs, st));
  RBG_TRY_HIP(ctx, a->row_edges.ensure((size_t)n * 4));
  RBG_TRY_HIP(ctx, a->row_off.ensure((size_t)(n + 1) * 8));
  const int mutual = regime == 2 ? 1 : 0;
  RBG_TRY_HIP(ctx, a->keep.ensure((size_t)n * ((K + 63) / 64) * sizeof(uint64_t)));
  RBG_TRY_HIP(ctx, launch_emit_mark((const uint32_t*)a->topk_id.p, (const uint16_t*)a->topk_cnt.p,
                                    (const uint32_t*)a->topk_len.p, (const uint32_t*)a->fence.p,
                                    FW, n, K, bs, mutual, (uint64_t*)a->keep.p,
                                    (uint32_t*)a->row_edges.p, st));
  size_t tmp_bytes = 0;
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(nullptr, nullptr, n, nullptr, tmp_bytes, st));
  RBG_TRY_HIP(ctx, a->scan_tmp.ensure(tmp_bytes));
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)a->row_edges.p, (uint64_t*)a->row_off.p,
                                             n, a->scan_tmp.p, tmp_bytes, st));
  // edge arrays at their capacity (n*K); the edge count stays on the device
  const uint64_t* ne_dev = (const uint64_t*)a->row_off.p + n;
  const uint64_t cap_e = (uint64_t)n * K;
  RBG_TRY_HIP(ctx, a->efrom.ensure(cap_e * 4));
  RBG_TRY_HIP(ctx, a->eto.ensure(cap_e * 4));
  RBG_TRY_HIP(ctx, a->ecnt.ensure(cap_e * 2));
  RBG_TRY_HIP(ctx, a->escomp.ensure(cap_e * 4));
  RBG_TRY_HIP(ctx, launch_emit_write((const uint32_t*)a->topk_id.p, (const uint16_t*)a->topk_cnt.p,
                                     (const uint32_t*)a->topk_len.p, n, K,
                                     (const uint64_t*)a->keep.p, (const uint64_t*)a->row_off.p,
                                     (uint32_t*)a->efrom.p, (uint32_t*)a->eto.p,
                                     (uint16_t*)a->ecnt.p, st));
  RBG_TRY_HIP(ctx, launch_cnt_to_sv((const uint16_t*)a->ecnt.p, cap_e, (const float*)a->svf.p,
                                    (float*)a->escomp.p, st, ne_dev));
  RBG_TRY_HIP(ctx, hipEventRecord(ev[4], st));
  if (S > 0) {
    RBG_TRY_HIP(ctx, a->escore.ensure(cap_e * 4));
    const bool padded = S > 1 && S <= 64;  // unit_pad / raw_pad hold the rows
    RBG_TRY_HIP(ctx, launch_edge_weights(
                         (const uint32_t*)a->efrom.p, (const uint32_t*)a->eto.p,
                         (const float*)a->escomp.p, cap_e,
                         (const float*)(padded ? a->unit_pad.p : a->unit.p),
                         (const float*)(padded ? a->raw_pad.p : a->raw.p),
                         S > 1 ? (const double*)a->colnorm.p : nullptr, S, w_comp,
                         P->min_edge_weight, P->neg_depth_thr, (float*)a->escore.p, nullptr, st,
                         padded, ne_dev));
  }
  RBG_TRY_HIP(ctx, hipEventRecord(ev[5], st));
  // the build's scalars back to the host
  BuildScalars hb{};
  uint64_t ne = 0, off_total = 0;
  RBG_TRY_HIP(ctx, hipMemcpyAsync(&hb, bs, sizeof(hb), hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(&ne, ne_dev, 8, hipMemcpyDeviceToHost, st));
  if (regime == 2)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(&off_total, (const uint64_t*)a->off.p + n, 8,
                                    hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  out->b0 = hb.b0;
  out->sim_cutoff = hb.sim_cutoff;
  out->gated_pairs = off_total / 2;
  out->n_edges = ne;
  a->last_edges = ne;
  if (!(P->flags & RBG_EDGES_ON_DEVICE)) {  // results to the host
    const size_t nb = std::max<uint64_t>(ne, 1);
    out->from = (uint32_t*)std::malloc(nb * 4);
    out->to = (uint32_t*)std::malloc(nb * 4);
    out->scomp = (float*)std::malloc(nb * 4);
    out->score = S > 0 ? (float*)std::malloc(nb * 4) : nullptr;
    if (!out->from || !out->to || !out->scomp || (S > 0 && !out->score))
      return ctx->fail(RBG_ERR_NOMEM, "host allocation of the edge list");
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->from, a->efrom.p, ne * 4, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->to, a->eto.p, ne * 4, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->scomp, a->escomp.p, ne * 4, hipMemcpyDeviceToHost, st));
    if (S > 0)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(out->score, a->escore.p, ne * 4, hipMemcpyDeviceToHost, st));
  }
  if (diag) {
    if (diag->topk_id)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_id, a->topk_id.p, (size_t)n * K * 4, hipMemcpyDeviceToHost, st));
    if (diag->topk_cnt)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_cnt, a->topk_cnt.p, (size_t)n * K * 2, hipMemcpyDeviceToHost, st));
    if (diag->topk_len)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_len, a->topk_len.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    if (diag->rowmax_cnt)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->rowmax_cnt, a->rowmax.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  }
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (diag && diag->win16) {
    const ProposalTables& pt = proposal_tables(m, P->pmh_seed);
    const uint32_t mpad = 32 * W;
    std::vector<uint8_t> wx((size_t)n * mpad);
    RBG_TRY_HIP(ctx, hipMemcpy(wx.data(), a->wx.p, wx.size(), hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t k = 0; k < m; ++k) {
        const uint8_t x = wx[(size_t)r * mpad + k];
        diag->win16[(size_t)r * m + k] = x == 0xFF ? 0 : pt.W16[(size_t)x * m + k];
      }
  }
  out->ms_sketch = elapsed_ms(ev[0], ev[1]);
  out->ms_pairs = elapsed_ms(ev[2], ev[3]);
  out->ms_select = elapsed_ms(ev[3], ev[4]);
  out->ms_edges = elapsed_ms(ev[4], ev[5]);
  out->ms_total = elapsed_ms(ev[0], ev[5]);
  out->ms_k_pair = elapsed_ms(ev[2], ev[6]);
  out->ms_k_compare = elapsed_ms(ev[7], ev[8]);
  out->ms_k_select = elapsed_ms(ev[9], ev[10]);
  return RBG_OK;
}

}  // namespace

extern "C" {

rbg_status rbg_open(int device, rbg_ctx** out) {
  if (!out) return RBG_ERR_ARG;
  *out = nullptr;
  auto ctx = std::make_unique<rbg_ctx>();
  ctx->device = device;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return RBG_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return RBG_ERR_HIP;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return RBG_ERR_HIP;
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return RBG_ERR_HIP;
  *out = ctx.release();
  return RBG_OK;
}

void rbg_close(rbg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  ctx->tables.clear();
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* rbg_last_error(const rbg_ctx* ctx) {
  static thread_local std::string copy;
  copy = ctx ? ctx->err : std::string("no context");
  return copy.c_str();
}

void rbg_default_params(rbg_graph_params* p) {
  if (!p) return;
  p->m = 500;
  p->max_edges = 200;
  p->min_edge_weight = 70 / 100.;
  p->coverage = 95 / 100.;
  p->min_cv = 1.0;
  p->neg_depth_thr = -0.3;
  p->pmh_seed = 42;
  p->rand_seed = 1;
  p->regime = 0;
  p->flags = 0;
}

rbg_status rbg_sketch_winners(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                              uint32_t m, uint64_t seed, uint32_t* out_win32, uint16_t* out_win16) {
  if (!ctx || !seqs || !offsets) return RBG_ERR_ARG;
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (n == 0) return RBG_OK;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  rbg_asm a;
  RBG_TRY(asm_upload(ctx, &a, seqs, offsets, n, nullptr, 0, 1.0));
  RBG_TRY(run_sketch(ctx, &a, m, seed, true));
  const uint32_t mpad = 32 * words_for(m);
  std::vector<uint8_t> wx((size_t)n * mpad);
  RBG_TRY_HIP(ctx, hipMemcpyAsync(wx.data(), a.wx.p, wx.size(), hipMemcpyDeviceToHost, ctx->stream));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const ProposalTables& pt = proposal_tables(m, seed);
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < m; ++k) {
      const uint8_t x = wx[(size_t)r * mpad + k];
      const size_t q = (size_t)x * m + k;
      if (out_win32) out_win32[(size_t)r * m + k] = x == 0xFF ? 0u : pt.W32[q];
      if (out_win16) out_win16[(size_t)r * m + k] = x == 0xFF ? 0 : pt.W16[q];
    }
  return RBG_OK;
}

rbg_status rbg_depth_features(const float* depth, uint32_t n, uint32_t S, double min_cv,
                              float* ranked, float* unit, float* raw, double* colnorm,
                              uint8_t* anynz) {
  if (!depth && n && S) return RBG_ERR_ARG;
  depth_features(depth, n, S, min_cv, ranked, unit, raw, colnorm, anynz);
  return RBG_OK;
}

rbg_status rbg_asm_create(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                          const float* depth, uint32_t S, double min_cv, rbg_asm** out) {
  if (!ctx || !out || !seqs || !offsets || (S > 0 && !depth)) return RBG_ERR_ARG;
  if (n == 0) return ctx->fail(RBG_ERR_ARG, "empty assembly");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  auto a = std::make_unique<rbg_asm>();
  RBG_TRY(asm_upload(ctx, a.get(), seqs, offsets, n, depth, S, min_cv));
  *out = a.release();
  return RBG_OK;
}

void rbg_asm_free(rbg_asm* a) { delete a; }

rbg_status rbg_asm_build_graph(rbg_ctx* ctx, rbg_asm* a, const rbg_graph_params* params,
                               rbg_graph_result* out, rbg_graph_diag* diag) {
  if (!ctx || !a || !params || !out) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const rbg_status s = build_graph(ctx, a, params, out, diag);
  if (s != RBG_OK) rbg_graph_result_free(out);
  return s;
}

rbg_status rbg_asm_copy_edges(rbg_ctx* ctx, rbg_asm* a, uint32_t* from, uint32_t* to,
                              float* scomp, float* score) {
  if (!ctx || !a) return RBG_ERR_ARG;
  const uint64_t ne = a->last_edges;
  if (ne == 0) return RBG_OK;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (from) RBG_TRY_HIP(ctx, hipMemcpyAsync(from, a->efrom.p, ne * 4, hipMemcpyDeviceToHost, st));
  if (to) RBG_TRY_HIP(ctx, hipMemcpyAsync(to, a->eto.p, ne * 4, hipMemcpyDeviceToHost, st));
  if (scomp) RBG_TRY_HIP(ctx, hipMemcpyAsync(scomp, a->escomp.p, ne * 4, hipMemcpyDeviceToHost, st));
  if (score && a->S > 0)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(score, a->escore.p, ne * 4, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  return RBG_OK;
}

rbg_status rbg_build_graph(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                           const float* depth, uint32_t S, const rbg_graph_params* params,
                           rbg_graph_result* out, rbg_graph_diag* diag) {
  if (!ctx || !params || !out) return RBG_ERR_ARG;
  rbg_asm* a = nullptr;
  RBG_TRY(rbg_asm_create(ctx, seqs, offsets, n, depth, S, params->min_cv, &a));
  const rbg_status s = rbg_asm_build_graph(ctx, a, params, out, diag);
  rbg_asm_free(a);
  return s;
}

void rbg_graph_result_free(rbg_graph_result* r) {
  if (!r) return;
  std::free(r->from);
  std::free(r->to);
  std::free(r->scomp);
  std::free(r->score);
  r->from = r->to = nullptr;
  r->scomp = r->score = nullptr;
  r->n_edges = 0;
}

rbg_status rbg_edge_weights(rbg_ctx* ctx, const uint32_t* from, const uint32_t* to,
                            const float* scomp, uint64_t ne, const float* unit, const float* raw,
                            const double* colnorm, uint32_t n, uint32_t S, double w_comp,
                            double min_edge_weight, double neg_depth_thr, float* score,
                            double* score64) {
  if (!ctx || (ne && (!from || !to || !scomp || !score))) return RBG_ERR_ARG;
  if (S > 1 && (!unit || !raw)) return RBG_ERR_ARG;
  if (ne == 0) return RBG_OK;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  for (uint64_t e = 0; e < ne; ++e)
    if (from[e] >= n || to[e] >= n) return ctx->fail(RBG_ERR_ARG, "edge endpoint out of range");
  hipStream_t st = ctx->stream;
  DevBuf df, dt, ds, du, dr, dc, dw, dw64;
  RBG_TRY_HIP(ctx, df.ensure(ne * 4));
  RBG_TRY_HIP(ctx, dt.ensure(ne * 4));
  RBG_TRY_HIP(ctx, ds.ensure(ne * 4));
  RBG_TRY_HIP(ctx, dw.ensure(ne * 4));
  if (score64) RBG_TRY_HIP(ctx, dw64.ensure(ne * 8));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(df.p, from, ne * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(dt.p, to, ne * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(ds.p, scomp, ne * 4, hipMemcpyHostToDevice, st));
  if (S > 1) {
    RBG_TRY_HIP(ctx, du.ensure((size_t)n * S * 4));
    RBG_TRY_HIP(ctx, dr.ensure((size_t)n * S * 4));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(du.p, unit, (size_t)n * S * 4, hipMemcpyHostToDevice, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(dr.p, raw, (size_t)n * S * 4, hipMemcpyHostToDevice, st));
    if (colnorm) {
      RBG_TRY_HIP(ctx, dc.ensure(S * 8));
      RBG_TRY_HIP(ctx, hipMemcpyAsync(dc.p, colnorm, S * 8, hipMemcpyHostToDevice, st));
    }
  }
  RBG_TRY_HIP(ctx, launch_edge_weights((const uint32_t*)df.p, (const uint32_t*)dt.p,
                                       (const float*)ds.p, ne, (const float*)du.p,
                                       (const float*)dr.p, colnorm ? (const double*)dc.p : nullptr,
                                       S, w_comp, min_edge_weight, neg_depth_thr, (float*)dw.p,
                                       score64 ? (double*)dw64.p : nullptr, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(score, dw.p, ne * 4, hipMemcpyDeviceToHost, st));
  if (score64) RBG_TRY_HIP(ctx, hipMemcpyAsync(score64, dw64.p, ne * 8, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  return RBG_OK;
}

rbg_status rbg_sketch_tables(uint32_t m, uint64_t seed, double* T, uint64_t* W, uint8_t* cls) {
  if (m < 2) return RBG_ERR_ARG;
  const ProposalTables& pt = proposal_tables(m, seed);
  if (T) std::memcpy(T, pt.T.data(), pt.T.size() * sizeof(double));
  if (W) std::memcpy(W, pt.W.data(), pt.W.size() * sizeof(uint64_t));
  if (cls) std::memcpy(cls, pt.cls.data(), pt.cls.size());
  return RBG_OK;
}

rbg_status rbg_calibrate(const uint32_t* rowmax_cnt, uint32_t n, uint32_t m, double b0,
                         double coverage, uint32_t rand_seed, int regime, double* sim_cutoff) {
  if ((!rowmax_cnt && n) || !sim_cutoff || m < 2 || (regime != 1 && regime != 2)) return RBG_ERR_ARG;
  if (regime == 2 && n < 2500) return RBG_ERR_ARG;
  std::vector<uint32_t> rm(rowmax_cnt, rowmax_cnt + n);
  for (uint32_t v : rm)
    if (v > m) return RBG_ERR_ARG;
  std::vector<double> svd;
  std::vector<float> svf;
  std::vector<uint16_t> level;
  sv_tables(m, b0, svd, svf, level);
  *sim_cutoff = regime == 1 ? calibrate_small(rm, svd, coverage, rand_seed)
                            : calibrate_fused(rm, svf, coverage, rand_seed);
  return RBG_OK;
}

rbg_status rbg_pmh_baseline(rbg_ctx* ctx, const uint16_t* win16, uint32_t n, uint32_t m,
                            double* b0) {
  if (!ctx || !b0 || (n && !win16) || m == 0) return RBG_ERR_ARG;
  // The pair draws are the reference's libstdc++ stream (host); the 16-bit
  // register compares of the drawn pairs run on the device.
  std::vector<uint32_t> pi, pj;
  baseline_pairs(n, pi, pj);
  if (pi.empty()) {
    *b0 = 0.0;
    return RBG_OK;
  }
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint32_t np = (uint32_t)pi.size();
  DevBuf dw, dpi, dpj, dc;
  RBG_TRY_HIP(ctx, dw.ensure((size_t)n * m * 2));
  RBG_TRY_HIP(ctx, dpi.ensure(np * 4));
  RBG_TRY_HIP(ctx, dpj.ensure(np * 4));
  RBG_TRY_HIP(ctx, dc.ensure(np * 4));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(dw.p, win16, (size_t)n * m * 2, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(dpi.p, pi.data(), np * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(dpj.p, pj.data(), np * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, launch_win16_counts((const uint16_t*)dw.p, m, (const uint32_t*)dpi.p,
                                       (const uint32_t*)dpj.p, np, (uint32_t*)dc.p, st));
  std::vector<uint32_t> cnt(np);
  RBG_TRY_HIP(ctx, hipMemcpyAsync(cnt.data(), dc.p, np * 4, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  *b0 = baseline_median(cnt, m);
  return RBG_OK;
}

}  // extern "C"
