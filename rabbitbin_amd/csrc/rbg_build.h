// The graph-build pipeline of librbg, shared by the C ABI (rbg_api.cpp) and
// the request server (rbg_server.cpp).  Internal header.
//
// A build runs over a batch of assemblies whose rows (contigs) are numbered
// consecutively (global rows).  An assembly's rows come from host input, or,
// for a served session, partly from the device state its earlier builds left
// (Resident): then only the new contigs are uploaded and sketched, and in the
// fused regime only the pairs with a new contig are evaluated; the session's
// earlier candidate lists are merged in on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/rbg.h"
#include "rbg_host.h"
#include "rbg_internal.h"
#include "rbg_kernels.h"

struct rbg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // A second stream (default priority) for work of a build that does not
  // depend on the stages before it: the int8 prefilter runs there beside the
  // sketch, between fork_ev (recorded on `stream`) and join_ev.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::string err;
  struct Tables {
    uint32_t m;
    uint64_t seed;
    rbg::DevBuf T, cls, codes, sT, sX;
  };
  std::vector<std::unique_ptr<Tables>> tables;
  hipEvent_t ev[12] = {};
  // per-kernel timing of an eager build: events 2k, 2k+1 around launch k
  hipEvent_t kev[2 * RBG_MAX_KERNEL_TIMES] = {};
  const char* kname[RBG_MAX_KERNEL_TIMES] = {};
  uint32_t nk = 0;
  // Pinned host staging for large transfers: two halves, so the pageable side
  // of one chunk is copied while the other chunk's DMA runs.
  struct Staging {
    char* p = nullptr;
    size_t half = 0;
    hipEvent_t done[2] = {nullptr, nullptr};
  } pin;
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

#define RBG_TRY_HIP(ctx, expr)                          \
  do {                                                  \
    const hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return (ctx)->hip(_e, #expr); \
  } while (0)
#define RBG_TRY(expr)             \
  do {                            \
    const rbg_status _s = (expr); \
    if (_s != RBG_OK) return _s;  \
  } while (0)

namespace rbg {

uint32_t words_for(uint32_t m);
// Host <-> device copies of large pageable buffers through ctx->pin (the
// pageable side copied by several threads).  h2d returns once the source may
// be reused; d2h returns when dst holds the data.  d2h_u32 subtracts `sub`
// from every 32-bit word on the way.
rbg_status h2d_staged(rbg_ctx* ctx, void* dst, const void* src, size_t bytes);
rbg_status d2h_staged(rbg_ctx* ctx, void* dst, const void* src, size_t bytes);
rbg_status d2h_u32_staged(rbg_ctx* ctx, uint32_t* dst, const uint32_t* src, size_t count, uint32_t sub);
void release_staging(rbg_ctx* ctx);
// Host memory for result arrays (released with free()).
void* alloc_host_result(size_t bytes);
rbg_status get_tables(rbg_ctx* ctx, uint32_t m, uint64_t seed, rbg_ctx::Tables** out);
double elapsed_ms(hipEvent_t a, hipEvent_t b);

// Unit and raw depth rows are padded to 16 * NC floats, int8 rows to kpad bytes.
inline uint32_t nc_for(uint32_t S) { return S > 1 ? (S + 15) / 16 : 1; }
inline uint32_t kpad_for(uint32_t S) { return S > 1 ? 32 * ((S + 31) / 32) : 32; }

// What a served session keeps between its builds.  Device rows use the
// session's own padding (nc_for(S), kpad_for(S)); ids in the candidate lists
// are local to the session.
struct Resident {
  uint32_t n = 0, S = 0;
  int regime = 0;             // regime of the last build (0: never built)
  DevBuf planes;              // n x 8W u32 bit-sliced register classes
  DevBuf unit_pad, raw_pad;   // n x 16 nc_for(S) f32 (S > 1)
  DevBuf q8;                  // n x kpad_for(S) i8 in pair-pass order (S > 1)
  DevBuf loff, list;          // regime 2: gated pairs of every row, CSR
  uint64_t nlist = 0;
  DevBuf mat;                 // dense counts: n x n match counts (the diagonal unused)
  bool dense = false;         // the last build counted pairs into mat (no candidate lists)
  std::vector<uint32_t> perm; // pair-pass position -> local row
  std::vector<uint8_t> anynz;
  std::vector<double> colsum; // column sums of the raw depths, accumulated in row order
  GateQuant gq;               // scale and error norms of q8 (the rows are on the device)
  size_t device_bytes() const {
    return planes.bytes + unit_pad.bytes + raw_pad.bytes + q8.bytes + loff.bytes + list.bytes +
           mat.bytes;
  }
};

// Sketches and depth rows a caller supplies instead of contigs and depths
// (rbg_build_graph_from_sketches): per-register class bytes n x mpad (0xFF
// unset), unit/raw rows n x S, column norms, anynz, and b0 (NaN: estimate).
struct SketchInput {
  const uint8_t* classes = nullptr;
  uint32_t mpad = 0;
  const float* unit = nullptr;
  const float* raw = nullptr;
  const double* colnorm = nullptr;
  const uint8_t* anynz = nullptr;
  double b0 = __builtin_nan("");
};

// One assembly of a build.
struct AsmInput {
  rbg_assembly_in in{};          // host contigs (and depth): the new rows
  const Resident* src = nullptr; // rows [0, src->n) come from here when set
  const SketchInput* sk = nullptr;  // set: in.seqs/in.depth are unused
  bool seqs_on_device = false;   // in.seqs is device memory (ingest); offsets and depth stay host
};

// Host side of one assembly of a batch.
struct AsmHost {
  uint32_t n = 0, S = 0, row0 = 0, colnorm0 = 0, pair0 = 0, np = 0;
  uint32_t n_old = 0;             // rows taken from src
  const Resident* src = nullptr;
  double b0_in = __builtin_nan("");  // b0 given by the caller (NaN: estimated on the device)
  uint32_t n_zero = 0;            // rows with anynz = 0 (they pass is_nz only with a nonzero partner)
  bool incr = false;              // fused pass over pairs with a new row, merged with src's lists
  bool dense_incr = false;        // small regime: counts of pairs with a new row, src's matrix kept
  GateQuant gq;                   // int8 row statistics (the rows live on the device)
  // the assembly's host state after this build (moved into its Resident by the server)
  std::vector<uint32_t> perm_local;
  std::vector<uint8_t> anynz_local;
  std::vector<double> colsum;
};

// The build configuration a device plan is prepared for.
struct PlanKey {
  uint32_t m = 0, K = 0, rand_seed = 0, tiles = 0, rows = 0;
  int regime = -1;
  uint64_t pmh_seed = 0, cand_cap = 0;
  double min_edge = 0, coverage = 0, neg_thr = 0, near_tol = 0;
  bool operator==(const PlanKey& o) const {
    return m == o.m && K == o.K && rand_seed == o.rand_seed && tiles == o.tiles && rows == o.rows &&
           regime == o.regime && pmh_seed == o.pmh_seed && cand_cap == o.cand_cap &&
           min_edge == o.min_edge && coverage == o.coverage && neg_thr == o.neg_thr &&
           near_tol == o.near_tol;
  }
};

// Regime of an assembly of n contigs under `regime` (0 = auto: n > 25000 fused,
// src/rabbitbin.cpp:5035-5036).
inline int regime_for(int regime, uint32_t n) { return regime ? regime : (n > 25000 ? 2 : 1); }
// Composition weight auto_w_comp(S) (src/rabbitbin.cpp:470-473).
inline double w_comp_for(uint32_t S) { return S <= 2 ? 1.0 : 0.0; }
// Whether the fused pass applies the abundance gate (src/impl/rb_graph.cpp:693-695).
inline bool gate_on(int regime, uint32_t S, double min_edge) {
  const double wc = w_comp_for(S);
  return regime == 2 && S > 1 && wc < 1.0 && wc < min_edge;
}
// Whether an assembly's pair counts go through the n x n matrix: the small
// regime, and the fused regime without the gate (every pair is a candidate,
// so the matrix, 2 bytes per pair, is cheaper than candidate lists).
inline bool dense_counts(int regime, uint32_t S, double min_edge) {
  return regime == 1 || !gate_on(regime, S, min_edge);
}
// Largest n x n count matrix of one assembly (n <= 185,363).
constexpr uint64_t kDenseMaxBytes = 64ull << 30;

}  // namespace rbg

// ---------------------------------------------------------------------------
// A batch of device-resident assemblies (one assembly is a batch of one):
// contigs, depth features and every intermediate of the pipeline.  Buffers
// persist across builds, so a resident batch is rebuilt without new
// allocations, and its build's launch sequence can be captured into a
// hipGraph and replayed.
// ---------------------------------------------------------------------------
struct rbg_asm {
  std::vector<rbg::AsmHost> asms;
  uint32_t nrows = 0;
  uint32_t NC = 1;     // unit / raw rows padded to 16*NC floats
  uint32_t kpad = 32;  // int8 row bytes
  uint32_t np_total = 0;
  uint32_t n_sketch = 0;      // rows the sketch kernel computes (the others come from sessions)
  bool has_src = false;       // some rows come from resident sessions
  uint32_t n_incr = 0;        // incremental assemblies (their list merge descriptors)
  uint32_t merge_max_old = 0; // most earlier rows of an incremental assembly
  uint64_t list_extra = 0;    // candidate-list entries merged from sessions
  uint32_t cls_rows = 0, cls_mpad = 0;  // rows given as class bytes (graph from sketches)
  rbg::DevBuf seqs, offs, sk_rows, asm_of, unit_pad, raw_pad, colnorm, anynz, q8, perm, pi, pj,
      pcnt, merge, cls_in;
  // plan of the current configuration
  bool planned = false;
  rbg::PlanKey key;
  uint64_t cand_cap_min = 0;  // grown after a candidate overflow
  std::vector<rbg::AsmDesc> desc_h;
  bool has_gate_items = false;
  uint32_t n_dense_items = 0, n_gate_items = 0;
  uint64_t mat_total = 0, cand_cap = 0, edge_cap = 0;
  uint32_t n_sel_static = 0;  // rows of dense assemblies at the head of selq
  uint32_t cal_max_items = 0;  // most contigs one assembly samples for calibration
  rbg_schedule sched{16, 64};
  rbg::DevBuf desc, bscal, svd, svf, level, cal_rnd, cal_members, cal_roff, cal_hist, dense_items, gate_items;
  rbg::DevBuf wx, planes, mat, cand, counters, deg, slot, off, list, topk_id, topk_cnt, topk_len, selq,
      rowmax, fence, keep, row_edges, row_off, efrom, eto, ecnt, escomp, escore, scan_tmp, bounds,
      near_e, near_w;
  uint64_t near_cap = 0;  // near-cut list entries (edge index, FP64 weight)
  // captured build
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rbg_graph_params last{};  // parameters of the last build
  // host copies of the last build's per-assembly results
  std::vector<rbg::BuildScalars> bs_h;
  std::vector<uint64_t> bounds_h;
  std::vector<uint64_t> near_e_h;  // the last build's near-cut edges (global edge index)
  std::vector<double> near_w_h;
  uint64_t ncand_last = 0;
  bool built = false;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  size_t device_bytes() const;
  void release_buffers();
  ~rbg_asm() { drop_graph(); }
};

namespace rbg {

// Host preparation and upload of a build's inputs (contigs, depth features,
// int8 rows, pair-pass order, b0 pair draws); rows from resident sessions are
// copied on the device.
rbg_status batch_upload(rbg_ctx* ctx, rbg_asm* b, const AsmInput* in, uint32_t count, double min_cv);
// Host -> device copies through the context's pinned staging.
struct H2DPart {
  void* dst;
  const void* src;
  size_t bytes;
};
rbg_status h2d_parts(rbg_ctx* ctx, const H2DPart* parts, size_t np);
// Descriptors, work items, capacities and calibration plans of one build
// configuration, and every buffer at its size.
rbg_status batch_plan(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P);
// One build: plan, launches (eager, or a captured graph), then the
// per-assembly scalars and bounds back to the host.
rbg_status run_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P, bool want_wx,
                     bool capture, bool* replayed);
// Assembly `a` of the last build (edges with ids local to the assembly).
rbg_status assembly_result(rbg_ctx* ctx, rbg_asm* b, uint32_t a, const rbg_graph_params* P,
                           bool on_device, rbg_graph_result* out, rbg_graph_diag* diag);
// Copies assembly `a` of the last build into `r` (device rows, lists, host
// state), reusing r's buffers when they are large enough.
rbg_status save_resident(rbg_ctx* ctx, rbg_asm* b, uint32_t a, const rbg_graph_params* P,
                         Resident* r);
// Device bytes a build of these inputs needs for its batch buffers (an
// upper estimate used for admission).
size_t estimate_build_bytes(const AsmInput* in, uint32_t count, const rbg_graph_params* P);

}  // namespace rbg
