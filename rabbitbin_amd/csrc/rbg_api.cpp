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


// Host side of one assembly of a batch.
struct AsmHost {
  uint32_t n = 0, S = 0, row0 = 0, colnorm0 = 0, pair0 = 0, np = 0;
  GateQuant gq;  // int8 row statistics (the rows live on the device)
};

// The build configuration a device plan is prepared for.
struct PlanKey {
  uint32_t m = 0, K = 0, rand_seed = 0, tiles = 0, rows = 0;
  int regime = -1;
  uint64_t pmh_seed = 0, cand_cap = 0;
  double min_edge = 0, coverage = 0, neg_thr = 0;
  bool operator==(const PlanKey& o) const {
    return m == o.m && K == o.K && rand_seed == o.rand_seed && tiles == o.tiles && rows == o.rows &&
           regime == o.regime && pmh_seed == o.pmh_seed && cand_cap == o.cand_cap &&
           min_edge == o.min_edge && coverage == o.coverage && neg_thr == o.neg_thr;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// A batch of device-resident assemblies (one assembly is a batch of one):
// contigs, depth features and every intermediate of the pipeline.  Rows of
// all assemblies are numbered consecutively (global rows).  Buffers persist
// across builds, so a resident batch is rebuilt without new allocations, and
// its build's launch sequence can be captured into a hipGraph and replayed.
// ---------------------------------------------------------------------------
struct rbg_asm {
  std::vector<AsmHost> asms;
  uint32_t nrows = 0;
  uint32_t NC = 1;     // unit / raw rows padded to 16*NC floats
  uint32_t kpad = 32;  // int8 row bytes
  uint32_t np_total = 0;
  DevBuf seqs, offs, asm_of, unit_pad, raw_pad, colnorm, anynz, q8, perm, pi, pj, pcnt;
  // plan of the current configuration
  bool planned = false;
  PlanKey key;
  uint64_t cand_cap_min = 0;  // grown after a candidate overflow
  std::vector<AsmDesc> desc_h;
  bool has_gate_items = false;
  uint32_t n_dense_items = 0, n_gate_items = 0;
  uint64_t mat_total = 0, cand_cap = 0, edge_cap = 0;
  rbg_schedule sched{16, 64};
  DevBuf desc, bscal, svd, svf, level, cal_rnd, cal_members, cal_roff, dense_items, gate_items;
  DevBuf wx, planes, mat, cand, counters, deg, cursor, off, list, topk_id, topk_cnt, topk_len,
      rowmax, fence, keep, row_edges, row_off, efrom, eto, ecnt, escomp, escore, scan_tmp, bounds;
  // captured build
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rbg_graph_params last{};  // parameters of the last build
  // host copies of the last build's per-assembly results
  std::vector<BuildScalars> bs_h;
  std::vector<uint64_t> bounds_h;
  uint64_t ncand_last = 0;
  bool built = false;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~rbg_asm() { drop_graph(); }
};

namespace {

rbg_status batch_upload(rbg_ctx* ctx, rbg_asm* b, const rbg_assembly_in* in, uint32_t count,
                        double min_cv) {
  hipStream_t st = ctx->stream;
  b->asms.assign(count, AsmHost{});
  uint64_t rows = 0, bytes = 0, pairs = 0, ncol = 0;
  uint32_t nc_max = 1, kc_max = 1;
  for (uint32_t a = 0; a < count; ++a) {
    const rbg_assembly_in& x = in[a];
    if (!x.seqs || !x.offsets || x.n == 0 || (x.S > 0 && !x.depth))
      return ctx->fail(RBG_ERR_ARG, "assembly " + std::to_string(a) + ": empty or missing input");
    if (x.S > 64)
      return ctx->fail(RBG_ERR_UNSUPPORTED, "graph builds support at most 64 depth samples");
    AsmHost& h = b->asms[a];
    h.n = x.n;
    h.S = x.S;
    h.row0 = (uint32_t)rows;
    h.colnorm0 = (uint32_t)ncol;
    rows += x.n;
    ncol += x.S;
    bytes += x.offsets[x.n] - x.offsets[0];
    if (x.S > 1) {
      nc_max = std::max(nc_max, (x.S + 15) / 16);
      kc_max = std::max(kc_max, (x.S + 31) / 32);
    }
  }
  if (rows >= (1u << 24)) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^24 contigs in a batch");
  b->nrows = (uint32_t)rows;
  b->NC = nc_max;
  b->kpad = 32 * kc_max;
  const uint32_t NC = nc_max, kpad = b->kpad;
  // The sketch kernel reads whole 16-byte chunks: keep 64 readable bytes past
  // the last contig (their contents are never counted).
  RBG_TRY_HIP(ctx, b->seqs.ensure(bytes + 64));
  std::vector<uint64_t> offs(rows + 1);
  std::vector<uint16_t> asm_of(rows);
  std::vector<uint8_t> anynz(rows, 1);  // no depth: is_nz always holds
  std::vector<float> upad(rows * 16 * NC, 0.0f), rpad(rows * 16 * NC, 0.0f);
  std::vector<double> colnorm(std::max<uint64_t>(ncol, 1), 0.0);
  std::vector<int8_t> q8(rows * kpad, 0);
  std::vector<uint32_t> perm(rows), pi, pj;
  uint64_t seq0 = 0;
  for (uint32_t a = 0; a < count; ++a) {
    const rbg_assembly_in& x = in[a];
    AsmHost& h = b->asms[a];
    const uint64_t nb = x.offsets[x.n] - x.offsets[0];
    if (nb)
      RBG_TRY_HIP(ctx, hipMemcpyAsync((char*)b->seqs.p + seq0, x.seqs + x.offsets[0], nb,
                                      hipMemcpyHostToDevice, st));
    for (uint32_t r = 0; r <= x.n; ++r) offs[h.row0 + r] = seq0 + (x.offsets[r] - x.offsets[0]);
    seq0 += nb;
    for (uint32_t r = 0; r < x.n; ++r) {
      asm_of[h.row0 + r] = (uint16_t)a;
      perm[h.row0 + r] = h.row0 + r;
    }
    if (x.S > 0) {
      std::vector<float> unit((size_t)x.n * x.S), raw((size_t)x.n * x.S);
      depth_features(x.depth, x.n, x.S, min_cv, nullptr, unit.data(), raw.data(),
                     colnorm.data() + h.colnorm0, anynz.data() + h.row0);
      for (uint32_t r = 0; r < x.n; ++r) {
        std::memcpy(&upad[(size_t)(h.row0 + r) * 16 * NC], &unit[(size_t)r * x.S], x.S * 4);
        std::memcpy(&rpad[(size_t)(h.row0 + r) * 16 * NC], &raw[(size_t)r * x.S], x.S * 4);
      }
      if (x.S > 1) {
        // The pair pass visits the assembly's rows in abundance order
        // (pair_pass_order); its int8 rows are stored in that order and
        // candidates map back to global rows through perm.
        std::vector<uint32_t> lp;
        pair_pass_order(unit.data(), x.n, x.S, lp);
        std::vector<float> sorted((size_t)x.n * x.S);
        for (uint32_t r = 0; r < x.n; ++r) {
          std::memcpy(&sorted[(size_t)r * x.S], &unit[(size_t)lp[r] * x.S], x.S * 4);
          perm[h.row0 + r] = h.row0 + lp[r];
        }
        gate_quantize(sorted.data(), x.n, x.S, h.gq);
        for (uint32_t r = 0; r < x.n; ++r)
          std::memcpy(&q8[(size_t)(h.row0 + r) * kpad], &h.gq.q[(size_t)r * h.gq.kpad], h.gq.kpad);
        std::vector<int8_t>().swap(h.gq.q);
      }
    }
    std::vector<uint32_t> ai, aj;
    baseline_pairs(x.n, ai, aj);
    h.pair0 = (uint32_t)pi.size();
    h.np = (uint32_t)ai.size();
    for (size_t t = 0; t < ai.size(); ++t) {
      pi.push_back(h.row0 + ai[t]);
      pj.push_back(h.row0 + aj[t]);
    }
    pairs += ai.size();
  }
  b->np_total = (uint32_t)pairs;
  RBG_TRY_HIP(ctx, b->offs.ensure(offs.size() * 8));
  RBG_TRY_HIP(ctx, b->asm_of.ensure(asm_of.size() * 2));
  RBG_TRY_HIP(ctx, b->anynz.ensure(anynz.size()));
  RBG_TRY_HIP(ctx, b->unit_pad.ensure(upad.size() * 4));
  RBG_TRY_HIP(ctx, b->raw_pad.ensure(rpad.size() * 4));
  RBG_TRY_HIP(ctx, b->colnorm.ensure(colnorm.size() * 8));
  RBG_TRY_HIP(ctx, b->q8.ensure(q8.size()));
  RBG_TRY_HIP(ctx, b->perm.ensure(perm.size() * 4));
  RBG_TRY_HIP(ctx, b->pi.ensure(std::max<size_t>(pi.size(), 1) * 4));
  RBG_TRY_HIP(ctx, b->pj.ensure(std::max<size_t>(pj.size(), 1) * 4));
  RBG_TRY_HIP(ctx, b->pcnt.ensure(std::max<size_t>(pi.size(), 1) * 4));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->offs.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->asm_of.p, asm_of.data(), asm_of.size() * 2, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->anynz.p, anynz.data(), anynz.size(), hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->unit_pad.p, upad.data(), upad.size() * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->raw_pad.p, rpad.data(), rpad.size() * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->colnorm.p, colnorm.data(), colnorm.size() * 8, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->q8.p, q8.data(), q8.size(), hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->perm.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, st));
  if (!pi.empty()) {
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->pi.p, pi.data(), pi.size() * 4, hipMemcpyHostToDevice, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->pj.p, pj.data(), pj.size() * 4, hipMemcpyHostToDevice, st));
  }
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));  // the host staging vectors go out of scope
  return RBG_OK;
}

// Descriptors, work items, capacities and calibration plans of one build
// configuration, and every buffer at its size.  Reused while the
// configuration stays the same.
rbg_status batch_plan(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P) {
  const uint32_t m = P->m, K = P->max_edges;
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (K == 0 || K > 1024) return ctx->fail(RBG_ERR_UNSUPPORTED, "max-edges must be in [1, 1024]");
  PlanKey key;
  key.m = m;
  key.K = K;
  key.rand_seed = P->rand_seed;
  key.tiles = b->sched.gate_tiles_per_item;
  key.rows = b->sched.dense_rows_per_item;
  key.regime = P->regime;
  key.pmh_seed = P->pmh_seed;
  key.cand_cap = b->cand_cap_min;
  key.min_edge = P->min_edge_weight;
  key.coverage = P->coverage;
  key.neg_thr = P->neg_depth_thr;
  if (b->planned && b->key == key) return RBG_OK;
  b->drop_graph();
  b->planned = false;
  rbg_ctx::Tables* t;
  RBG_TRY(get_tables(ctx, m, P->pmh_seed, &t));  // uploaded before any capture
  const uint32_t A = (uint32_t)b->asms.size(), N = b->nrows, W = words_for(m);
  std::vector<AsmDesc>& D = b->desc_h;
  D.assign(A, AsmDesc{});
  uint64_t mat_total = 0, cap = 0;
  std::vector<uint32_t> gfull, gpart;
  std::vector<int32_t> rnd((size_t)A * kCalibRandDev);
  std::vector<uint32_t> members, roff((size_t)A * 11, 0);
  for (uint32_t a = 0; a < A; ++a) {
    const AsmHost& h = b->asms[a];
    AsmDesc& d = D[a];
    const int regime = P->regime ? P->regime : (h.n > 25000 ? 2 : 1);
    if (regime != 1 && regime != 2) return ctx->fail(RBG_ERR_ARG, "regime must be 0, 1 or 2");
    if (regime == 2 && h.n < 2500)
      return ctx->fail(RBG_ERR_ARG, "fused regime needs n >= 2500 (assembly " + std::to_string(a) + ")");
    d.row0 = h.row0;
    d.n = h.n;
    d.S = h.S;
    d.regime = (uint32_t)regime;
    // composition weight: auto_w_comp(S) (src/rabbitbin.cpp:470-473)
    d.w_comp = h.S <= 2 ? 1.0 : 0.0;
    // abundance-first gate (src/impl/rb_graph.cpp:693-706)
    const bool gate = regime == 2 && h.S > 1 && d.w_comp < 1.0 && d.w_comp < P->min_edge_weight;
    d.thr = gate ? (float)((P->min_edge_weight - d.w_comp) / (1.0 - d.w_comp) - 1e-4)
                 : -__builtin_huge_valf();
    // prefilter cut: with the gate off every pair is a candidate
    d.qcut = gate ? gate_qcut(h.gq, d.thr) : INT32_MIN;
    d.acc1_end = gate ? 2 * (h.S / 32) : 0;
    d.pair0 = h.pair0;
    d.np = h.np;
    d.colnorm0 = h.colnorm0;
    if (regime == 1) {
      d.mat0 = mat_total;
      mat_total += (uint64_t)h.n * h.n;
    } else {
      gate_work_items(a, h.n, b->sched.gate_tiles_per_item, gfull, gpart);
      const uint64_t pairs = (uint64_t)h.n * (h.n - 1) / 2;
      cap += std::min<uint64_t>(pairs, std::max<uint64_t>(1ull << 26, 2048ull * h.n));
    }
    CalibPlan plan;
    calibration_plan(h.n, P->rand_seed, regime, plan);
    std::memcpy(&rnd[(size_t)a * kCalibRandDev], plan.rnd.data(), kCalibRandDev * 4);
    const uint32_t base = (uint32_t)members.size();
    for (uint32_t c : plan.members) members.push_back(h.row0 + c);
    for (uint32_t r = 0; r < 11; ++r)
      roff[(size_t)a * 11 + r] =
          base + (r < plan.round_off.size() ? plan.round_off[r] : plan.round_off.back());
  }
  cap = std::max(cap, b->cand_cap_min);
  std::vector<uint32_t> ditems;
  dense_work_items(D, b->sched.dense_rows_per_item, ditems);
  gfull.insert(gfull.end(), gpart.begin(), gpart.end());
  b->n_dense_items = (uint32_t)(ditems.size() / 4);
  b->n_gate_items = (uint32_t)(gfull.size() / 4);
  b->has_gate_items = b->n_gate_items > 0;
  b->mat_total = mat_total;
  b->cand_cap = cap;
  b->edge_cap = (uint64_t)N * K;
  hipStream_t st = ctx->stream;
  RBG_TRY_HIP(ctx, b->desc.ensure(A * sizeof(AsmDesc)));
  RBG_TRY_HIP(ctx, b->dense_items.ensure(std::max<size_t>(ditems.size(), 4) * 4));
  RBG_TRY_HIP(ctx, b->gate_items.ensure(std::max<size_t>(gfull.size(), 4) * 4));
  RBG_TRY_HIP(ctx, b->cal_rnd.ensure(rnd.size() * 4));
  RBG_TRY_HIP(ctx, b->cal_members.ensure(std::max<size_t>(members.size(), 1) * 4));
  RBG_TRY_HIP(ctx, b->cal_roff.ensure(roff.size() * 4));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->desc.p, D.data(), A * sizeof(AsmDesc), hipMemcpyHostToDevice, st));
  if (!ditems.empty())
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->dense_items.p, ditems.data(), ditems.size() * 4,
                                    hipMemcpyHostToDevice, st));
  if (!gfull.empty())
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->gate_items.p, gfull.data(), gfull.size() * 4,
                                    hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->cal_rnd.p, rnd.data(), rnd.size() * 4, hipMemcpyHostToDevice, st));
  if (!members.empty())
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->cal_members.p, members.data(), members.size() * 4,
                                    hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->cal_roff.p, roff.data(), roff.size() * 4, hipMemcpyHostToDevice, st));
  const uint32_t FW = fence_words(K), KW = (K + 63) / 64;
  RBG_TRY_HIP(ctx, b->planes.ensure((size_t)N * 8 * W * 4));
  RBG_TRY_HIP(ctx, b->bscal.ensure(A * sizeof(BuildScalars)));
  RBG_TRY_HIP(ctx, b->svd.ensure((size_t)A * (m + 1) * 8));
  RBG_TRY_HIP(ctx, b->svf.ensure((size_t)A * (m + 1) * 4));
  RBG_TRY_HIP(ctx, b->level.ensure((size_t)A * (m + 1) * 2));
  RBG_TRY_HIP(ctx, b->mat.ensure(std::max<uint64_t>(mat_total, 1) * 2));
  RBG_TRY_HIP(ctx, b->cand.ensure(std::max<uint64_t>(cap, 1) * 8));
  RBG_TRY_HIP(ctx, b->counters.ensure(16));
  RBG_TRY_HIP(ctx, b->deg.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->cursor.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->off.ensure((size_t)(N + 1) * 8));
  RBG_TRY_HIP(ctx, b->list.ensure(std::max<uint64_t>(2 * cap, 1) * 8));
  // +16 ids: the mutual test reads whole 16-id segments of a row's list
  RBG_TRY_HIP(ctx, b->topk_id.ensure(((size_t)N * K + 16) * 4));
  RBG_TRY_HIP(ctx, b->topk_cnt.ensure((size_t)N * K * 2));
  RBG_TRY_HIP(ctx, b->topk_len.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->rowmax.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->fence.ensure((size_t)N * FW * 4));
  RBG_TRY_HIP(ctx, b->keep.ensure((size_t)N * KW * 8));
  RBG_TRY_HIP(ctx, b->row_edges.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->row_off.ensure((size_t)(N + 1) * 8));
  const uint64_t ce = std::max<uint64_t>(b->edge_cap, 1);
  RBG_TRY_HIP(ctx, b->efrom.ensure(ce * 4));
  RBG_TRY_HIP(ctx, b->eto.ensure(ce * 4));
  RBG_TRY_HIP(ctx, b->ecnt.ensure(ce * 2));
  RBG_TRY_HIP(ctx, b->escomp.ensure(ce * 4));
  RBG_TRY_HIP(ctx, b->escore.ensure(ce * 4));
  RBG_TRY_HIP(ctx, b->bounds.ensure((size_t)A * 4 * 8));
  size_t tmp = 0;
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(nullptr, nullptr, N, nullptr, tmp, st));
  RBG_TRY_HIP(ctx, b->scan_tmp.ensure(tmp));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  b->key = key;
  b->planned = true;
  return RBG_OK;
}

// Every launch of one build, on ctx->stream, without a host round trip (so it
// can be captured).  `timed` records the stage events of the eager path.
rbg_status enqueue_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P, bool want_wx,
                         bool timed) {
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  const uint32_t m = P->m, K = P->max_edges, W = words_for(m), N = b->nrows;
  const uint32_t A = (uint32_t)b->asms.size(), FW = fence_words(K);
  rbg_ctx::Tables* t;
  RBG_TRY(get_tables(ctx, m, P->pmh_seed, &t));
  const AsmDesc* desc = (const AsmDesc*)b->desc.p;
  const uint16_t* asm_of = (const uint16_t*)b->asm_of.p;
  BuildScalars* bs = (BuildScalars*)b->bscal.p;
  auto mark = [&](int k) -> rbg_status {
    if (timed) RBG_TRY_HIP(ctx, hipEventRecord(ev[k], st));
    return RBG_OK;
  };
  RBG_TRY(mark(0));
  const uint32_t mpad = 32 * W;
  if (want_wx) RBG_TRY_HIP(ctx, b->wx.ensure((size_t)N * mpad));
  RBG_TRY_HIP(ctx, launch_sketch((const uint8_t*)b->seqs.p, (const uint64_t*)b->offs.p, N, m, W,
                                 (const double*)t->sT.p, (const uint8_t*)t->sX.p,
                                 (const uint8_t*)t->cls.p, (const uint8_t*)t->codes.p,
                                 want_wx ? (uint8_t*)b->wx.p : nullptr, mpad,
                                 (uint32_t*)b->planes.p, st));
  RBG_TRY(mark(1));
  // b0 and the sv tables (the pair draws depend on n only)
  if (b->np_total)
    RBG_TRY_HIP(ctx, launch_list_counts((const uint32_t*)b->planes.p, W, (const uint32_t*)b->pi.p,
                                        (const uint32_t*)b->pj.p, b->np_total,
                                        (uint32_t*)b->pcnt.p, st));
  RBG_TRY_HIP(ctx, launch_baseline_tables((const uint32_t*)b->pcnt.p, desc, A, m, bs,
                                          (double*)b->svd.p, (float*)b->svf.p,
                                          (uint16_t*)b->level.p, st));
  RBG_TRY(mark(2));
  if (b->n_dense_items)
    RBG_TRY_HIP(ctx, launch_dense_counts((const uint32_t*)b->planes.p, W, desc,
                                         (const uint32_t*)b->dense_items.p, b->n_dense_items,
                                         (uint16_t*)b->mat.p, st));
  unsigned long long* ncand = (unsigned long long*)b->counters.p;
  if (b->has_gate_items) {
    RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->counters.p, 4, 0u, st));
    RBG_TRY_HIP(ctx, launch_gate_mfma((const int8_t*)b->q8.p, b->kpad, desc,
                                      (const uint32_t*)b->perm.p, (const uint32_t*)b->gate_items.p,
                                      b->n_gate_items, (uint64_t*)b->cand.p, ncand, b->cand_cap, st));
  }
  RBG_TRY(mark(6));
  RBG_TRY(mark(7));
  if (b->has_gate_items) {
    // exact FP32 gate and register compare of the candidates, in place
    RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->deg.p, N, 0u, st));
    RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->cursor.p, N, 0u, st));
    RBG_TRY_HIP(ctx, launch_gate_compare((uint64_t*)b->cand.p, b->cand_cap, ncand,
                                         (const float*)b->unit_pad.p, b->NC,
                                         (const uint8_t*)b->anynz.p, desc, asm_of,
                                         (const uint32_t*)b->planes.p, W, (uint32_t*)b->deg.p, st));
  }
  RBG_TRY(mark(8));
  if (b->has_gate_items) {
    size_t tmp = b->scan_tmp.bytes;
    RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)b->deg.p, (uint64_t*)b->off.p, N,
                                               b->scan_tmp.p, tmp, st));
    RBG_TRY_HIP(ctx, launch_scatter((const uint64_t*)b->cand.p, b->cand_cap, ncand,
                                    (const uint64_t*)b->off.p, (uint32_t*)b->cursor.p,
                                    (uint64_t*)b->list.p, st));
  }
  RBG_TRY(mark(3));
  RBG_TRY(mark(9));
  RBG_TRY_HIP(ctx, launch_select((const uint64_t*)b->list.p, (const uint64_t*)b->off.p,
                                 (const uint16_t*)b->mat.p, (const uint8_t*)b->anynz.p, desc, asm_of,
                                 N, (const uint16_t*)b->level.p, m, K, (uint32_t*)b->topk_id.p,
                                 (uint16_t*)b->topk_cnt.p, (uint32_t*)b->topk_len.p,
                                 (uint32_t*)b->rowmax.p, (uint32_t*)b->fence.p, FW, st));
  RBG_TRY(mark(10));
  // calibration (the glibc rand() draws are prepared per (n, seed))
  RBG_TRY_HIP(ctx, launch_calibrate((const uint32_t*)b->rowmax.p, desc, A, m,
                                    (const double*)b->svd.p, (const float*)b->svf.p,
                                    (const int32_t*)b->cal_rnd.p, (const uint32_t*)b->cal_members.p,
                                    (const uint32_t*)b->cal_roff.p, P->coverage, bs, st));
  RBG_TRY_HIP(ctx, launch_emit_mark((const uint32_t*)b->topk_id.p, (const uint16_t*)b->topk_cnt.p,
                                    (const uint32_t*)b->topk_len.p, (const uint32_t*)b->fence.p, FW,
                                    N, K, bs, desc, asm_of, (uint64_t*)b->keep.p,
                                    (uint32_t*)b->row_edges.p, st));
  {
    size_t tmp = b->scan_tmp.bytes;
    RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)b->row_edges.p,
                                               (uint64_t*)b->row_off.p, N, b->scan_tmp.p, tmp, st));
  }
  const uint64_t* ne = (const uint64_t*)b->row_off.p + N;
  RBG_TRY_HIP(ctx, launch_emit_write((const uint32_t*)b->topk_id.p, (const uint16_t*)b->topk_cnt.p,
                                     (const uint32_t*)b->topk_len.p, N, K, (const uint64_t*)b->keep.p,
                                     (const uint64_t*)b->row_off.p, (uint32_t*)b->efrom.p,
                                     (uint32_t*)b->eto.p, (uint16_t*)b->ecnt.p, st));
  RBG_TRY_HIP(ctx, launch_cnt_to_sv((const uint16_t*)b->ecnt.p, b->edge_cap, ne,
                                    (const uint32_t*)b->efrom.p, asm_of, m, (const float*)b->svf.p,
                                    (float*)b->escomp.p, st));
  RBG_TRY(mark(4));
  RBG_TRY_HIP(ctx, launch_edge_weights_batch((const uint32_t*)b->efrom.p, (const uint32_t*)b->eto.p,
                                             (const float*)b->escomp.p, b->edge_cap, ne,
                                             (const float*)b->unit_pad.p, (const float*)b->raw_pad.p,
                                             b->NC, (const double*)b->colnorm.p, desc, asm_of,
                                             P->min_edge_weight, P->neg_depth_thr,
                                             (float*)b->escore.p, st));
  RBG_TRY_HIP(ctx, launch_gather_bounds((const uint64_t*)b->row_off.p,
                                        b->has_gate_items ? (const uint64_t*)b->off.p : nullptr, desc,
                                        A, (uint64_t*)b->bounds.p, st));
  RBG_TRY(mark(5));
  return RBG_OK;
}

// One build: plan, launches (eager with stage events, or a captured graph),
// then the per-assembly scalars and bounds back to the host.  A candidate
// overflow grows the capacity and builds again.
rbg_status run_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P, bool want_wx,
                     bool capture, bool* replayed) {
  hipStream_t st = ctx->stream;
  const uint32_t A = (uint32_t)b->asms.size();
  if (replayed) *replayed = false;
  for (int attempt = 0; attempt < 2; ++attempt) {
    RBG_TRY(batch_plan(ctx, b, P));
    if (capture && !want_wx) {
      const bool had = b->exec != nullptr;
      if (!had) {
        RBG_TRY_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const rbg_status s = enqueue_build(ctx, b, P, false, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (s != RBG_OK) {
          if (g) (void)hipGraphDestroy(g);
          return s;
        }
        RBG_TRY_HIP(ctx, e);
        b->graph = g;
        RBG_TRY_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      }
      if (replayed) *replayed = had;
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[0], st));
      RBG_TRY_HIP(ctx, hipGraphLaunch(b->exec, st));
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[5], st));
    } else {
      RBG_TRY(enqueue_build(ctx, b, P, want_wx, true));
    }
    b->bs_h.resize(A);
    b->bounds_h.resize((size_t)A * 4);
    unsigned long long nc = 0;
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->bs_h.data(), b->bscal.p, A * sizeof(BuildScalars),
                                    hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->bounds_h.data(), b->bounds.p, (size_t)A * 32,
                                    hipMemcpyDeviceToHost, st));
    if (b->has_gate_items)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(&nc, b->counters.p, 8, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
    b->ncand_last = nc;
    if (nc <= b->cand_cap) {
      b->built = true;
      return RBG_OK;
    }
    b->cand_cap_min = nc;  // rebuild with room for every candidate
    b->built = false;
  }
  return ctx->fail(RBG_ERR_NOMEM, "candidate list overflow after a resize");
}

// Assembly `a` of the last build into `out` (edges with contig ids local to
// the assembly, copied to the host unless on_device); diag as rbg_graph_diag.
rbg_status assembly_result(rbg_ctx* ctx, rbg_asm* b, uint32_t a, const rbg_graph_params* P,
                           bool on_device, rbg_graph_result* out, rbg_graph_diag* diag) {
  std::memset(out, 0, sizeof(*out));
  if (!b->built) return ctx->fail(RBG_ERR_ARG, "no graph has been built");
  if (a >= b->asms.size()) return ctx->fail(RBG_ERR_ARG, "assembly index out of range");
  hipStream_t st = ctx->stream;
  const AsmHost& h = b->asms[a];
  const BuildScalars& s = b->bs_h[a];
  const uint64_t e0 = b->bounds_h[4 * a], e1 = b->bounds_h[4 * a + 1];
  const uint64_t ne = e1 - e0;
  out->n_edges = ne;
  out->b0 = s.b0;
  out->sim_cutoff = s.sim_cutoff;
  out->regime = (int)b->desc_h[a].regime;
  out->gated_pairs = (b->bounds_h[4 * a + 3] - b->bounds_h[4 * a + 2]) / 2;
  const uint32_t K = P->max_edges, m = P->m, n = h.n, r0 = h.row0;
  if (!on_device) {
    const size_t nb = std::max<uint64_t>(ne, 1);
    out->from = (uint32_t*)std::malloc(nb * 4);
    out->to = (uint32_t*)std::malloc(nb * 4);
    out->scomp = (float*)std::malloc(nb * 4);
    out->score = h.S > 0 ? (float*)std::malloc(nb * 4) : nullptr;
    if (!out->from || !out->to || !out->scomp || (h.S > 0 && !out->score))
      return ctx->fail(RBG_ERR_NOMEM, "host allocation of the edge list");
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->from, (uint32_t*)b->efrom.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->to, (uint32_t*)b->eto.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(out->scomp, (float*)b->escomp.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
    if (h.S > 0)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(out->score, (float*)b->escore.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
  }
  if (diag) {
    if (diag->topk_id)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_id, (uint32_t*)b->topk_id.p + (size_t)r0 * K,
                                      (size_t)n * K * 4, hipMemcpyDeviceToHost, st));
    if (diag->topk_cnt)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_cnt, (uint16_t*)b->topk_cnt.p + (size_t)r0 * K,
                                      (size_t)n * K * 2, hipMemcpyDeviceToHost, st));
    if (diag->topk_len)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->topk_len, (uint32_t*)b->topk_len.p + r0, (size_t)n * 4,
                                      hipMemcpyDeviceToHost, st));
    if (diag->rowmax_cnt)
      RBG_TRY_HIP(ctx, hipMemcpyAsync(diag->rowmax_cnt, (uint32_t*)b->rowmax.p + r0, (size_t)n * 4,
                                      hipMemcpyDeviceToHost, st));
  }
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (r0) {  // global rows -> ids local to the assembly
    if (!on_device)
      for (uint64_t e = 0; e < ne; ++e) {
        out->from[e] -= r0;
        out->to[e] -= r0;
      }
    if (diag && diag->topk_id && diag->topk_len)
      for (uint32_t r = 0; r < n; ++r)
        for (uint32_t t2 = 0; t2 < diag->topk_len[r]; ++t2) diag->topk_id[(size_t)r * K + t2] -= r0;
  }
  if (diag && diag->win16) {
    if (!b->wx.p) return ctx->fail(RBG_ERR_ARG, "winners were not kept by this build");
    const ProposalTables& pt = proposal_tables(m, P->pmh_seed);
    const uint32_t mpad = 32 * words_for(m);
    std::vector<uint8_t> wx((size_t)n * mpad);
    RBG_TRY_HIP(ctx, hipMemcpy(wx.data(), (uint8_t*)b->wx.p + (size_t)r0 * mpad, wx.size(),
                               hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t k = 0; k < m; ++k) {
        const uint8_t x = wx[(size_t)r * mpad + k];
        diag->win16[(size_t)r * m + k] = x == 0xFF ? 0 : pt.W16[(size_t)x * m + k];
      }
  }
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
  const rbg_assembly_in in{seqs, offsets, n, nullptr, 0};
  RBG_TRY(batch_upload(ctx, &a, &in, 1, 1.0));
  rbg_ctx::Tables* t;
  RBG_TRY(get_tables(ctx, m, seed, &t));
  const uint32_t W = words_for(m), mpad = 32 * W;
  RBG_TRY_HIP(ctx, a.planes.ensure((size_t)n * 8 * W * 4));
  RBG_TRY_HIP(ctx, a.wx.ensure((size_t)n * mpad));
  RBG_TRY_HIP(ctx, launch_sketch((const uint8_t*)a.seqs.p, (const uint64_t*)a.offs.p, n, m, W,
                                 (const double*)t->sT.p, (const uint8_t*)t->sX.p,
                                 (const uint8_t*)t->cls.p, (const uint8_t*)t->codes.p,
                                 (uint8_t*)a.wx.p, mpad, (uint32_t*)a.planes.p, ctx->stream));
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

rbg_status rbg_asm_create(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                          const float* depth, uint32_t S, double min_cv, rbg_asm** out) {
  if (!ctx || !out || !seqs || !offsets || (S > 0 && !depth)) return RBG_ERR_ARG;
  const rbg_assembly_in in{seqs, offsets, n, depth, S};
  return rbg_batch_create(ctx, &in, 1, min_cv, out);
}

void rbg_asm_free(rbg_asm* a) { delete a; }

rbg_status rbg_asm_build_graph(rbg_ctx* ctx, rbg_asm* a, const rbg_graph_params* params,
                               rbg_graph_result* out, rbg_graph_diag* diag) {
  if (!ctx || !a || !params || !out) return RBG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  if (a->asms.size() != 1) return ctx->fail(RBG_ERR_ARG, "rbg_asm_build_graph needs a single assembly");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const bool capture = (params->flags & RBG_BUILD_CAPTURE) != 0;
  const bool want_wx = diag && diag->win16;
  bool replayed = false;
  RBG_TRY(run_build(ctx, a, params, want_wx, capture, &replayed));
  a->last = *params;
  const rbg_status s =
      assembly_result(ctx, a, 0, params, (params->flags & RBG_EDGES_ON_DEVICE) != 0, out, diag);
  if (s != RBG_OK) {
    rbg_graph_result_free(out);
    return s;
  }
  hipEvent_t* ev = ctx->ev;
  out->ms_total = elapsed_ms(ev[0], ev[5]);
  if (!(capture && !want_wx)) {
    out->ms_sketch = elapsed_ms(ev[0], ev[1]);
    out->ms_pairs = elapsed_ms(ev[2], ev[3]);
    out->ms_select = elapsed_ms(ev[3], ev[4]);
    out->ms_edges = elapsed_ms(ev[4], ev[5]);
    out->ms_k_pair = elapsed_ms(ev[2], ev[6]);
    out->ms_k_compare = elapsed_ms(ev[7], ev[8]);
    out->ms_k_select = elapsed_ms(ev[9], ev[10]);
  }
  return RBG_OK;
}

rbg_status rbg_asm_copy_edges(rbg_ctx* ctx, rbg_asm* a, uint32_t* from, uint32_t* to,
                              float* scomp, float* score) {
  if (!ctx || !a) return RBG_ERR_ARG;
  if (!a->built) return ctx->fail(RBG_ERR_ARG, "no graph has been built");
  const uint64_t e0 = a->bounds_h[0], ne = a->bounds_h[1] - e0;
  if (ne == 0) return RBG_OK;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (from) RBG_TRY_HIP(ctx, hipMemcpyAsync(from, (uint32_t*)a->efrom.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
  if (to) RBG_TRY_HIP(ctx, hipMemcpyAsync(to, (uint32_t*)a->eto.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
  if (scomp) RBG_TRY_HIP(ctx, hipMemcpyAsync(scomp, (float*)a->escomp.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
  if (score && a->asms[0].S > 0)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(score, (float*)a->escore.p + e0, ne * 4, hipMemcpyDeviceToHost, st));
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

rbg_status rbg_batch_create(rbg_ctx* ctx, const rbg_assembly_in* asms, uint32_t count,
                            double min_cv, rbg_asm** out) {
  if (!ctx || !out || (count && !asms)) return RBG_ERR_ARG;
  *out = nullptr;
  if (count == 0 || count > 65535) return ctx->fail(RBG_ERR_ARG, "a batch holds 1..65535 assemblies");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  auto b = std::make_unique<rbg_asm>();
  RBG_TRY(batch_upload(ctx, b.get(), asms, count, min_cv));
  *out = b.release();
  return RBG_OK;
}

rbg_status rbg_batch_set_schedule(rbg_asm* b, const rbg_schedule* s) {
  if (!b || !s) return RBG_ERR_ARG;
  const uint32_t t = s->gate_tiles_per_item, r = s->dense_rows_per_item;
  if (t < 1 || t > 64 || r < 1 || r > 4096) return RBG_ERR_ARG;
  b->sched = *s;
  return RBG_OK;
}

rbg_status rbg_batch_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* params,
                           rbg_batch_stats* stats) {
  if (!ctx || !b || !params) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  bool replayed = false;
  const bool capture = (params->flags & RBG_BUILD_CAPTURE) != 0;
  RBG_TRY(run_build(ctx, b, params, false, capture, &replayed));
  b->last = *params;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->candidates = b->ncand_last;
    for (size_t a = 0; a < b->asms.size(); ++a)
      stats->edges += b->bounds_h[4 * a + 1] - b->bounds_h[4 * a];
    stats->ms_total = elapsed_ms(ctx->ev[0], ctx->ev[5]);
    stats->replayed = replayed ? 1 : 0;
  }
  return RBG_OK;
}

rbg_status rbg_batch_result(rbg_ctx* ctx, rbg_asm* b, uint32_t index, rbg_graph_result* out) {
  if (!ctx || !b || !out) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  const rbg_status s = assembly_result(ctx, b, index, &b->last, false, out, nullptr);
  if (s != RBG_OK) rbg_graph_result_free(out);
  return s;
}

rbg_status rbg_depth_features(const float* depth, uint32_t n, uint32_t S, double min_cv,
                              float* ranked, float* unit, float* raw, double* colnorm,
                              uint8_t* anynz) {
  if (!depth && n && S) return RBG_ERR_ARG;
  depth_features(depth, n, S, min_cv, ranked, unit, raw, colnorm, anynz);
  return RBG_OK;
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
