// librbg: the C ABI of include/rbg.h over the graph-build pipeline
// (rbg_build.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rbg_build.h"

using namespace rbg;

extern "C" {

rbg_status rbg_open(int device, rbg_ctx** out) {
  if (!out) return RBG_ERR_ARG;
  *out = nullptr;
  auto ctx = std::make_unique<rbg_ctx>();
  ctx->device = device;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return RBG_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return RBG_ERR_HIP;
  if (noalloc_setup(device) != hipSuccess) return RBG_ERR_HIP;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return RBG_ERR_HIP;
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) return RBG_ERR_HIP;
  if (hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess) return RBG_ERR_HIP;
  if (hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) return RBG_ERR_HIP;
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return RBG_ERR_HIP;
  for (auto& e : ctx->kev)
    if (hipEventCreate(&e) != hipSuccess) return RBG_ERR_HIP;
  *out = ctx.release();
  return RBG_OK;
}

void rbg_close(rbg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  ctx->tables.clear();
  release_staging(ctx);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->kev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* rbg_last_error(const rbg_ctx* ctx) {
  static thread_local std::string copy;
  copy = ctx ? ctx->err : std::string("no context");
  return copy.c_str();
}

}  // extern "C"

namespace {

// The reference's two ways of reading a variable: rb_getenv treats an empty
// value as unset (src/rabbitbin.cpp:428-431), getenv does not.
const char* env_nonempty(const char* key) {
  const char* e = std::getenv(key);
  return (e && e[0]) ? e : nullptr;
}

struct Knob {
  const char* name;
  uint32_t bit;
  bool (*changed)(const char* key);  // value differs from the default in effect
};

// Each predicate restates how the reference parses the variable and when the
// parsed value differs from the default (file:line of the parse).
const Knob kKnobs[] = {
    {"RABBIT_PMH", RBG_SEM_NO_PMH,  // rabbitbin.cpp:433-436
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '0'; }},
    {"RABBIT_PMHK", RBG_SEM_PMH_K,  // rabbitbin.cpp:4668
     [](const char* k) { const char* e = env_nonempty(k); return e && std::atoi(e) != 4; }},
    {"RABBIT_PMH_BASE", RBG_SEM_NO_PMH_BASE,  // rabbitbin.cpp:4669
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '0'; }},
    {"RABBIT_GC_NORM", RBG_SEM_GC_NORM,  // rabbitbin.cpp:441-444
     [](const char* k) { const char* e = env_nonempty(k); return e && std::atoi(e) != 1; }},
    {"RABBIT_GC_NORM_CAP", RBG_SEM_GC_NORM_CAP,  // rabbitbin.cpp:3987-3988
     [](const char* k) {
       const char* e = env_nonempty(k);
       if (!e) return false;
       const double v = std::atof(e);
       return (v < 1.0 ? 20.0 : v) != 20.0;
     }},
    {"RABBIT_IDF_NORM", RBG_SEM_IDF_NORM,  // rabbitbin.cpp:3989
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '1'; }},
    {"RABBIT_EXACT_COS", RBG_SEM_EXACT_COS,  // rabbitbin.cpp:3990-3991
     [](const char* k) { const char* e = env_nonempty(k); return e && std::atoi(e) >= 1; }},
    {"RABBIT_WIN_BITS", RBG_SEM_WIN_BITS,  // rabbitbin.cpp:4935-4936
     [](const char* k) { const char* e = std::getenv(k); return e && std::atoi(e) != 16; }},
    {"RABBIT_LSH", RBG_SEM_LSH,  // rabbitbin.cpp:4937-4938
     [](const char* k) { const char* e = std::getenv(k); return e && e[0] == '1'; }},
    {"RABBIT_PMH_INDEX", RBG_SEM_PMH_INDEX,  // rabbitbin.cpp:4711
     [](const char* k) { return std::getenv(k) != nullptr; }},
    {"RABBIT_GRAPH_INDEX", RBG_SEM_GRAPH_INDEX,  // rabbitbin.cpp:4657-4660
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '1'; }},
    {"RABBIT_NO_ABDFIRST", RBG_SEM_NO_ABDFIRST,  // impl/rb_graph.cpp:693
     [](const char* k) { return std::getenv(k) != nullptr; }},
    {"RABBIT_MUTUAL_KNN", RBG_SEM_NO_MUTUAL,  // rabbitbin.cpp:437-440
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '0'; }},
    {"RABBIT_W_COMP", RBG_SEM_W_COMP,  // rabbitbin.cpp:451-456, 4671-4672
     [](const char* k) { return env_nonempty(k) != nullptr; }},
    {"RABBIT_EDGE_CUT", RBG_SEM_EDGE_CUT,  // rabbitbin.cpp:505-509
     [](const char* k) {
       const char* e = env_nonempty(k);
       return e && (!std::strcmp(e, "otsu") || !std::strcmp(e, "gmm"));
     }},
    {"RABBIT_EDGE_POWER", RBG_SEM_EDGE_POWER,  // rabbitbin.cpp:4701
     [](const char* k) { const char* e = env_nonempty(k); return e && std::atof(e) != 1.0; }},
    {"RABBIT_SNN", RBG_SEM_SNN,  // rabbitbin.cpp:484-485
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] != '0'; }},
    {"RABBIT_DEPTH_SIM", RBG_SEM_DEPTH_SIM,  // rabbitbin.cpp:654-660
     [](const char* k) {
       const char* e = env_nonempty(k);
       if (!e) return false;
       const char c = e[0];
       return c == 'c' || c == 'C' || c == '0' || c == 'w' || c == 'W' || c == '1';
     }},
    {"RABBIT_DEPTH_FUSE", RBG_SEM_DEPTH_FUSE,  // rabbitbin.cpp:661-667
     [](const char* k) {
       const char* e = env_nonempty(k);
       if (!e) return false;
       const char c = e[0];
       return c == 'g' || c == 'G' || c == '0' || c == 'p' || c == 'P' || c == '2';
     }},
    {"RABBIT_DEPTH_WJAC_NORM", RBG_SEM_WJAC_NORM,  // rabbitbin.cpp:668-671
     [](const char* k) { const char* e = env_nonempty(k); return e && e[0] == '0'; }},
    {"RABBIT_DEPTH_WJAC_CUT", RBG_SEM_WJAC_CUT,  // rabbitbin.cpp:672-675, 942
     [](const char* k) { const char* e = env_nonempty(k); return e && std::atof(e) > 0.0; }},
    {"RABBIT_DUAL_CONJ", RBG_SEM_DUAL_CONJ,  // rabbitbin.cpp:3686-3688, 951
     [](const char* k) { const char* e = std::getenv(k); return e && std::atoi(e) != 0; }},
    {"RABBIT_SNV_EDGE", RBG_SEM_SNV_EDGE,  // rabbitbin.cpp:3488-3490
     [](const char* k) { const char* e = std::getenv(k); return e && std::atoi(e) != 0; }},
};

}  // namespace

extern "C" {

uint32_t rbg_env_semantics(char* names, size_t names_cap) {
  uint32_t mask = 0;
  std::string list;
  for (const Knob& k : kKnobs)
    if (k.changed(k.name)) {
      mask |= k.bit;
      if (!list.empty()) list += ",";
      list += k.name;
    }
  if (names && names_cap) {
    const size_t n = std::min(list.size(), names_cap - 1);
    std::memcpy(names, list.data(), n);
    names[n] = 0;
  }
  return mask;
}

void rbg_default_params(rbg_graph_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->m = 500;
  p->max_edges = 200;
  p->min_edge_weight = 70 / 100.;
  p->coverage = 95 / 100.;
  p->min_cv = 1.0;
  // RABBIT_NEG_DEPTH (src/rabbitbin.cpp:445-449): an explicit parameter here
  const char* e = env_nonempty("RABBIT_NEG_DEPTH");
  p->neg_depth_thr = e ? std::atof(e) : -0.3;
  p->pmh_seed = 42;
  p->rand_seed = 1;
  p->regime = 0;
  p->flags = 0;
  p->near_tol = 1e-6;
  p->semantics = rbg_env_semantics(nullptr, 0);
}

rbg_status rbg_sketch_winners(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                              uint32_t m, uint64_t seed, uint32_t* out_win32, uint16_t* out_win16) {
  if (!ctx || !seqs || !offsets) return RBG_ERR_ARG;
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (n == 0) return RBG_OK;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  rbg_asm a;
  AsmInput in;
  in.in = rbg_assembly_in{seqs, offsets, n, nullptr, 0};
  RBG_TRY(batch_upload(ctx, &a, &in, 1, 1.0));
  rbg_ctx::Tables* t;
  RBG_TRY(get_tables(ctx, m, seed, &t));
  const uint32_t W = words_for(m), mpad = 32 * W;
  RBG_TRY_HIP(ctx, a.planes.ensure((size_t)n * 8 * W * 4));
  RBG_TRY_HIP(ctx, a.wx.ensure((size_t)n * mpad));
  RBG_TRY_HIP(ctx, launch_sketch((const uint8_t*)a.seqs.p, (const uint64_t*)a.offs.p, n, m, W,
                                 (const double*)t->sT.p, (const uint8_t*)t->sX.p,
                                 (const uint8_t*)t->cls.p, (const uint8_t*)t->codes.p,
                                 (uint8_t*)a.wx.p, mpad, (uint32_t*)a.planes.p, nullptr, ctx->stream));
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
    out->ms_k_pair = 0.0;  // the prefilter runs on the side stream: its own events
    out->ms_k_compare = elapsed_ms(ev[7], ev[8]);
    out->ms_k_select = elapsed_ms(ev[9], ev[10]);
    out->n_kernels = ctx->nk;
    for (uint32_t k = 0; k < ctx->nk; ++k) {
      out->kernels[k] = rbg_kernel_time{ctx->kname[k], elapsed_ms(ctx->kev[2 * k], ctx->kev[2 * k + 1])};
      if (std::strcmp(ctx->kname[k], "gate_mfma_kernel") == 0) out->ms_k_pair = out->kernels[k].ms;
    }
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

rbg_status rbg_build_graph_from_sketches(rbg_ctx* ctx, const uint16_t* win16, uint32_t n, uint32_t m,
                                         const double* b0, const float* unit, const float* raw,
                                         const double* colnorm, const uint8_t* anynz, uint32_t S,
                                         const rbg_graph_params* params, rbg_graph_result* out,
                                         rbg_graph_diag* diag) {
  if (!ctx || !params || !out || (n && !win16)) return RBG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  if (n == 0) return ctx->fail(RBG_ERR_ARG, "no contigs");
  if (m != params->m) return ctx->fail(RBG_ERR_ARG, "the sketches' m differs from params->m");
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (S > 1 && (!unit || !raw)) return ctx->fail(RBG_ERR_ARG, "S > 1 needs unit and raw rows");
  if (b0 && !(*b0 >= 0.0 && *b0 < 1.0)) return ctx->fail(RBG_ERR_ARG, "b0 must be in [0, 1)");
  if (diag && diag->win16) return ctx->fail(RBG_ERR_ARG, "the winners are the caller's input");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  // Per register, the distinct non-zero 16-bit winners are numbered in
  // ascending order; two registers match exactly when their numbers do.
  const uint32_t mpad = 32 * words_for(m);
  std::vector<uint8_t> classes((size_t)n * mpad, 0xFF);
  std::vector<uint16_t> present((size_t)m * 4096, 0);  // m x 65536 bits
  for (size_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < m; ++k) {
      const uint16_t v = win16[r * m + k];
      present[(size_t)k * 4096 + (v >> 4)] |= (uint16_t)(1u << (v & 15));
    }
  std::vector<uint8_t> rank((size_t)m * 65536, 0xFF);
  for (uint32_t k = 0; k < m; ++k) {
    uint32_t c = 0;
    for (uint32_t v = 1; v < 65536; ++v)
      if (present[(size_t)k * 4096 + (v >> 4)] >> (v & 15) & 1u) {
        if (c == 255)
          return ctx->fail(RBG_ERR_UNSUPPORTED, "register " + std::to_string(k) +
                                                    " holds more than 255 distinct winners");
        rank[(size_t)k * 65536 + v] = (uint8_t)c++;
      }
  }
  for (size_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < m; ++k) classes[r * mpad + k] = rank[(size_t)k * 65536 + win16[r * m + k]];
  SketchInput sk;
  sk.classes = classes.data();
  sk.mpad = mpad;
  sk.unit = unit;
  sk.raw = raw;
  sk.colnorm = colnorm;
  sk.anynz = anynz;
  if (b0) sk.b0 = *b0;
  AsmInput in;
  in.in = rbg_assembly_in{nullptr, nullptr, n, nullptr, S};
  in.sk = &sk;
  rbg_asm a;
  RBG_TRY(batch_upload(ctx, &a, &in, 1, params->min_cv));
  rbg_graph_params P = *params;
  P.flags &= ~(uint32_t)RBG_BUILD_CAPTURE;
  RBG_TRY(run_build(ctx, &a, &P, false, false, nullptr));
  const rbg_status st = assembly_result(ctx, &a, 0, &P, (P.flags & RBG_EDGES_ON_DEVICE) != 0, out, diag);
  if (st != RBG_OK) {
    rbg_graph_result_free(out);
    return st;
  }
  out->ms_total = elapsed_ms(ctx->ev[0], ctx->ev[5]);
  return RBG_OK;
}

rbg_status rbg_batch_create(rbg_ctx* ctx, const rbg_assembly_in* asms, uint32_t count,
                            double min_cv, rbg_asm** out) {
  if (!ctx || !out || (count && !asms)) return RBG_ERR_ARG;
  *out = nullptr;
  if (count == 0 || count > 65535) return ctx->fail(RBG_ERR_ARG, "a batch holds 1..65535 assemblies");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  for (uint32_t a = 0; a < count; ++a)
    if (asms[a].n == 0) return ctx->fail(RBG_ERR_ARG, "assembly " + std::to_string(a) + ": no contigs");
  std::vector<AsmInput> in(count);
  for (uint32_t a = 0; a < count; ++a) in[a].in = asms[a];
  auto b = std::make_unique<rbg_asm>();
  RBG_TRY(batch_upload(ctx, b.get(), in.data(), count, min_cv));
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

rbg_status rbg_batch_autotune(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* params, uint32_t reps,
                              rbg_schedule* best, double* best_ms) {
  if (!ctx || !b || !params || reps == 0) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  rbg_graph_params P = *params;
  P.flags = 0;
  bool fused = false, small = false;
  for (const AsmHost& h : b->asms) (regime_for(P.regime, h.n) == 2 ? fused : small) = true;
  // time of one eager build under schedule s (best of reps)
  auto timed = [&](const rbg_schedule& s, double* ms) -> rbg_status {
    b->sched = s;
    *ms = 1e300;
    for (uint32_t r = 0; r < reps; ++r) {
      RBG_TRY(run_build(ctx, b, &P, false, false, nullptr));
      *ms = std::min(*ms, elapsed_ms(ctx->ev[0], ctx->ev[5]));
    }
    return RBG_OK;
  };
  rbg_schedule cur = b->sched;
  double cur_ms;
  RBG_TRY(timed(cur, &cur_ms));
  static const uint32_t kTiles[] = {2, 4, 8, 16, 32}, kRows[] = {16, 32, 64, 128, 256};
  if (fused)
    for (uint32_t t : kTiles) {
      if (t == cur.gate_tiles_per_item) continue;
      rbg_schedule s = cur;
      s.gate_tiles_per_item = t;
      double ms;
      RBG_TRY(timed(s, &ms));
      if (ms < cur_ms) {
        cur = s;
        cur_ms = ms;
      }
    }
  if (small)
    for (uint32_t r : kRows) {
      if (r == cur.dense_rows_per_item) continue;
      rbg_schedule s = cur;
      s.dense_rows_per_item = r;
      double ms;
      RBG_TRY(timed(s, &ms));
      if (ms < cur_ms) {
        cur = s;
        cur_ms = ms;
      }
    }
  b->sched = cur;
  RBG_TRY(run_build(ctx, b, &P, false, false, nullptr));  // leave the batch built under it
  b->last = P;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
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
  std::free(r->near_edge);
  std::free(r->near_score64);
  r->from = r->to = nullptr;
  r->scomp = r->score = nullptr;
  r->near_edge = nullptr;
  r->near_score64 = nullptr;
  r->n_edges = 0;
  r->n_near = 0;
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
