/*
 * rbg.h -- C ABI of librbg, the MI355X (gfx950) implementation of RabbitBin's
 * graph-construction path: canonical 4-mer weighted ProbMinHash sketching,
 * all-pairs register-match similarity, per-row top-N with calibration and the
 * mutual filter, and Spearman / weighted-Jaccard edge weighting.
 *
 * Every entry point returns an rbg_status (0 = success) and never throws.
 * Host buffers are caller-owned unless a function says otherwise; buffers the
 * library allocates are released with the matching *_free call.
 * One context owns one HIP device and one stream; calls on a context are not
 * re-entrant (drive each context from one host thread).
 *
 * Reference seams replaced (paths relative to the RabbitBin source tree):
 *   build_pmh_winners           src/rabbitbin.cpp:1436-1441   -> rbg_sketch_winners
 *   estimate_pmh_baseline       src/rabbitbin.cpp:1340-1369   -> rbg_pmh_baseline
 *   gen_fused_calib_graph       src/impl/rb_graph.cpp:501     -> rbg_build_graph (fused)
 *   calibrate_sim_cutoff +      src/impl/rb_graph.cpp:1362,
 *   build_similarity_graph      src/impl/rb_graph.cpp:199     -> rbg_build_graph (small)
 *   graph dispatch              src/rabbitbin.cpp:5035-5054   -> rbg_build_graph (auto)
 *   the same graph seams over the caller's g_win16 / g_depth_unit / g_anynz
 *                                                             -> rbg_build_graph_from_sketches
 *   edge-weight loop            src/rabbitbin.cpp:5482-5509   -> rbg_edge_weights
 *   depth features              src/rabbitbin.cpp:4784-4988   -> rbg_depth_features
 */
#ifndef RBG_H
#define RBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int rbg_status;
#define RBG_OK 0
#define RBG_ERR_HIP (-1)         /* a HIP runtime call failed */
#define RBG_ERR_ARG (-2)         /* invalid argument */
#define RBG_ERR_NOMEM (-3)       /* device or host allocation failed */
#define RBG_ERR_UNSUPPORTED (-4) /* configuration outside the implemented path */
#define RBG_ERR_BUDGET (-5)      /* request exceeds the HBM budget */
#define RBG_ERR_TIMEOUT (-6)     /* a wait ended before the request was done */

typedef struct rbg_ctx rbg_ctx;

/* Open device `device` (HIP ordinal) and create the context's stream. */
rbg_status rbg_open(int device, rbg_ctx** out);
void rbg_close(rbg_ctx* ctx);
/* Message of the last failure on this context (static storage, never NULL). */
const char* rbg_last_error(const rbg_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Stage (a): sketches.  Replaces build_pmh_winners (src/rabbitbin.cpp:1436)
 * for k = 4 with the default GC-normalised weights (cap 20) and PMH seed
 * `seed` (the reference uses 42).  seqs holds the n contigs back to back;
 * contig r is seqs[offsets[r] .. offsets[r+1]).  Outputs (either may be NULL):
 *   out_win32  n x m  folded 32-bit winners, identical to the reference's
 *                     g_win_flat rows (src/rabbitbin.cpp:1582-1588)
 *   out_win16  n x m  16-bit folds (fold_win16, src/rabbitbin.cpp:249-253)
 * ------------------------------------------------------------------------- */
rbg_status rbg_sketch_winners(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets,
                              uint32_t n, uint32_t m, uint64_t seed,
                              uint32_t* out_win32, uint16_t* out_win16);

/* ---------------------------------------------------------------------------
 * Depth features (src/rabbitbin.cpp:4784-4988, src/ranker.h:55-133,
 * src/impl/rb_abundance.cpp:94-114).  depth: n x S raw per-sample means.
 * Any output may be NULL.
 * ------------------------------------------------------------------------- */
rbg_status rbg_depth_features(const float* depth, uint32_t n, uint32_t S, double min_cv,
                              float* ranked, float* unit, float* raw, double* colnorm,
                              uint8_t* anynz);

/* ---------------------------------------------------------------------------
 * Graph construction: the dispatch of src/rabbitbin.cpp:5035-5054 with the
 * default options.  Sketches, b0, the pair pass, top-N, calibration, the
 * mutual filter and (when S > 0) edge weights, in one call.
 * ------------------------------------------------------------------------- */
typedef struct rbg_graph_params {
  uint32_t m;              /* --sketch-m (default 500) */
  uint32_t max_edges;      /* --max-edges (default 200) */
  double min_edge_weight;  /* --min-edge-score / 100 (default 0.70) */
  double coverage;         /* --max-posterior / 100 (default 0.95) */
  double min_cv;           /* --min-coverage (default 1.0) */
  double neg_depth_thr;    /* RABBIT_NEG_DEPTH default -0.3 */
  uint64_t pmh_seed;       /* g_pmh_seed (42) */
  uint32_t rand_seed;      /* srand() seed (--seed); calibration consumes rand() */
  int regime;              /* 0 auto (n > 25000 -> fused), 1 small, 2 fused */
  uint32_t flags;          /* RBG_EDGES_ON_DEVICE: leave the edge list in HBM */
  double near_tol;         /* edges whose FP64 weight (or correlation) lies within
                              near_tol of the min-edge (or negative) cut are
                              listed in the result (default 1e-6) */
  uint32_t semantics;      /* RBG_SEM_* bits of non-default reference semantics
                              (rbg_env_semantics); any bit set makes every
                              build return RBG_ERR_UNSUPPORTED */
} rbg_graph_params;

/* ---------------------------------------------------------------------------
 * Reference semantics selected by RABBIT_* environment variables.  The
 * reference reads them at run time (src/rabbitbin.cpp:428-680, 3975-3992,
 * 4655-4715, 4930-4940, 3480-3495, 3683-3688; src/impl/rb_graph.cpp:693).
 * This library implements the defaults only.  rbg_env_semantics reads the
 * process environment the way the reference does (an empty value counts as
 * unset where the reference uses rb_getenv, and as set where it calls getenv)
 * and returns the bits of the variables whose value changes this path's
 * results; `names` (may be NULL) receives them comma-separated.  Variables that
 * change speed only (RABBIT_TILE, RABBIT_ABD_Q8, RABBIT_NO_COMP_EE,
 * RABBIT_NO_ABD_PRELOAD, ingest and NUMA knobs) and those acting before or
 * after this path (depth parsing and prefilter, clustering, splitting,
 * recruitment) are not reported.  RABBIT_NEG_DEPTH is a parameter here:
 * rbg_default_params reads it into neg_depth_thr.
 * ------------------------------------------------------------------------- */
#define RBG_SEM_NO_PMH          (1u << 0)   /* RABBIT_PMH=0: OPH sketches */
#define RBG_SEM_PMH_K           (1u << 1)   /* RABBIT_PMHK != 4 */
#define RBG_SEM_NO_PMH_BASE     (1u << 2)   /* RABBIT_PMH_BASE=0: no b0 rescale */
#define RBG_SEM_GC_NORM         (1u << 3)   /* RABBIT_GC_NORM != 1 */
#define RBG_SEM_GC_NORM_CAP     (1u << 4)   /* RABBIT_GC_NORM_CAP != 20 */
#define RBG_SEM_IDF_NORM        (1u << 5)   /* RABBIT_IDF_NORM=1 */
#define RBG_SEM_EXACT_COS       (1u << 6)   /* RABBIT_EXACT_COS >= 1 */
#define RBG_SEM_WIN_BITS        (1u << 7)   /* RABBIT_WIN_BITS != 16 */
#define RBG_SEM_LSH             (1u << 8)   /* RABBIT_LSH=1 */
#define RBG_SEM_PMH_INDEX       (1u << 9)   /* RABBIT_PMH_INDEX set */
#define RBG_SEM_GRAPH_INDEX     (1u << 10)  /* RABBIT_GRAPH_INDEX=1 */
#define RBG_SEM_NO_ABDFIRST     (1u << 11)  /* RABBIT_NO_ABDFIRST set: no abundance gate */
#define RBG_SEM_NO_MUTUAL       (1u << 12)  /* RABBIT_MUTUAL_KNN=0 */
#define RBG_SEM_W_COMP          (1u << 13)  /* RABBIT_W_COMP set */
#define RBG_SEM_EDGE_CUT        (1u << 14)  /* RABBIT_EDGE_CUT=otsu|gmm */
#define RBG_SEM_EDGE_POWER      (1u << 15)  /* RABBIT_EDGE_POWER != 1 */
#define RBG_SEM_SNN             (1u << 16)  /* RABBIT_SNN on */
#define RBG_SEM_DEPTH_SIM       (1u << 17)  /* RABBIT_DEPTH_SIM=corr|wjac */
#define RBG_SEM_DEPTH_FUSE      (1u << 18)  /* RABBIT_DEPTH_FUSE=geo|product */
#define RBG_SEM_WJAC_NORM       (1u << 19)  /* RABBIT_DEPTH_WJAC_NORM=0 */
#define RBG_SEM_WJAC_CUT        (1u << 20)  /* RABBIT_DEPTH_WJAC_CUT > 0 */
#define RBG_SEM_DUAL_CONJ       (1u << 21)  /* RABBIT_DUAL_CONJ != 0: conjunctive dual-depth corr */
#define RBG_SEM_SNV_EDGE        (1u << 22)  /* RABBIT_SNV_EDGE != 0 */
uint32_t rbg_env_semantics(char* names, size_t names_cap);

/* With this flag rbg_asm_build_graph returns the counts, cutoffs and timings
 * only (the result's array pointers are NULL); the edge list stays in the
 * assembly's device buffers until the next build and is copied out with
 * rbg_asm_copy_edges. */
#define RBG_EDGES_ON_DEVICE 1u
/* With this flag the build's launch sequence is captured into a hipGraph the
 * first time a resident assembly (or batch) is built with a configuration,
 * and the graph is replayed by later builds with the same configuration (no
 * host round trip inside a build).  Per-stage timings are then not reported. */
#define RBG_BUILD_CAPTURE 2u

/* Fills `p` with the reference defaults listed above (rand_seed = 1,
 * near_tol = 1e-6), neg_depth_thr from RABBIT_NEG_DEPTH when that is set (as
 * src/rabbitbin.cpp:446-449 reads it) and semantics = rbg_env_semantics(). */
void rbg_default_params(rbg_graph_params* p);

/* Device time of one kernel launch of a build (name: static storage). */
typedef struct rbg_kernel_time {
  const char* name;
  double ms;
} rbg_kernel_time;
#define RBG_MAX_KERNEL_TIMES 16

typedef struct rbg_graph_result {
  uint64_t n_edges;
  uint32_t* from;          /* edges sorted by (from, to), from < to */
  uint32_t* to;
  float* scomp;            /* composition similarity sComp */
  float* score;            /* edgeScore (0 for cut edges); NULL when S == 0 */
  double sim_cutoff;       /* simCutoff (per-mille scale, as the reference) */
  double b0;               /* PMH baseline */
  int regime;              /* 1 small, 2 fused */
  uint64_t gated_pairs;    /* pairs that reached the register compare */
  /* stage timings in milliseconds, measured with HIP events on the stream */
  double ms_sketch, ms_pairs, ms_select, ms_edges, ms_total;
  /* single-kernel durations: the pair kernel (the int8 prefilter of the
   * abundance gate in the fused regime; 0 in the small regime), the exact
   * gate + register compare of its candidates, and the top-N selection */
  double ms_k_pair, ms_k_compare, ms_k_select;
  /* Edges whose weight the reference's mixed FP32/FP64 arithmetic and FP64
   * could decide differently: the FP64 correlation within near_tol of the
   * negative cut, or the FP64 weight within near_tol of the min-edge cut.
   * near_edge: their indices in the edge list (ascending); near_score64:
   * their weights computed in FP64 throughout (0 when FP64 cuts the edge).
   * score itself is the reference's arithmetic.  NULL when n_near == 0. */
  uint64_t n_near;
  uint64_t* near_edge;
  double* near_score64;
  /* Per-kernel device times of an eager rbg_asm_build_graph / rbg_build_graph
   * (HIP events around each launch on the context's stream), in launch
   * order; n_kernels = 0 for a replayed graph and for the other entry points. */
  uint32_t n_kernels;
  rbg_kernel_time kernels[RBG_MAX_KERNEL_TIMES];
  uint64_t candidates;     /* int8 prefilter candidates of the build (fused regime) */
} rbg_graph_result;

/* Optional per-row diagnostics for parity tests (any pointer may be NULL):
 * top-N lists sorted by neighbour id, their lengths, and the per-row maximum
 * match count used by calibration. */
typedef struct rbg_graph_diag {
  uint32_t* topk_id;       /* n x max_edges */
  uint16_t* topk_cnt;      /* n x max_edges */
  uint32_t* topk_len;      /* n */
  uint32_t* rowmax_cnt;    /* n */
  uint16_t* win16;         /* n x m folded winners */
} rbg_graph_diag;

/* seqs/offsets as rbg_sketch_winners; depth n x S raw means (S may be 0). */
rbg_status rbg_build_graph(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets,
                           uint32_t n, const float* depth, uint32_t S,
                           const rbg_graph_params* params, rbg_graph_result* out,
                           rbg_graph_diag* diag);
void rbg_graph_result_free(rbg_graph_result* r);

/* ---------------------------------------------------------------------------
 * Graph from the caller's sketches and depth rows: the seams of
 * gen_fused_calib_graph (src/impl/rb_graph.cpp:501) and calibrate_sim_cutoff +
 * build_similarity_graph (src/impl/rb_graph.cpp:1362, 199), dispatched as
 * src/rabbitbin.cpp:5035-5054, with the globals they read passed explicitly:
 *   win16    n x m  g_win16: 16-bit folded winners, 0 = unset register
 *   b0       g_pmh_baseline; NULL estimates it (estimate_pmh_baseline)
 *   unit     n x S  g_depth_unit (S > 1)     raw  n x S  g_depth_raw (S > 1)
 *   colnorm  S      g_depth_colnorm (NULL: 1) anynz n    g_anynz (NULL: all 1)
 * m must equal params->m.  Per register the distinct non-zero winners are
 * renumbered to byte classes, at most 255 per register (PMH sketches of
 * 4-mers have at most 136); the match counts are those of
 * pmh_match_count16 (src/rabbitbin.cpp:1261-1293).  Edge weights as
 * rbg_build_graph.  diag->win16 must be NULL (the winners are the input).
 * ------------------------------------------------------------------------- */
rbg_status rbg_build_graph_from_sketches(rbg_ctx* ctx, const uint16_t* win16, uint32_t n, uint32_t m,
                                         const double* b0, const float* unit, const float* raw,
                                         const double* colnorm, const uint8_t* anynz, uint32_t S,
                                         const rbg_graph_params* params, rbg_graph_result* out,
                                         rbg_graph_diag* diag);

/* ---------------------------------------------------------------------------
 * Device-resident assembly: the contigs and depth features are uploaded once
 * (depth features are derived at upload), and rbg_asm_build_graph runs the
 * whole graph construction on the resident copy.  rbg_build_graph is
 * rbg_asm_create + rbg_asm_build_graph + rbg_asm_free.
 * ------------------------------------------------------------------------- */
typedef struct rbg_asm rbg_asm;
rbg_status rbg_asm_create(rbg_ctx* ctx, const char* seqs, const uint64_t* offsets, uint32_t n,
                          const float* depth, uint32_t S, double min_cv, rbg_asm** out);
void rbg_asm_free(rbg_asm* a);
rbg_status rbg_asm_build_graph(rbg_ctx* ctx, rbg_asm* a, const rbg_graph_params* params,
                               rbg_graph_result* out, rbg_graph_diag* diag);
/* Copies the last build's n_edges edges into caller buffers (any may be NULL;
 * score is only written when the assembly has depth samples). */
rbg_status rbg_asm_copy_edges(rbg_ctx* ctx, rbg_asm* a, uint32_t* from, uint32_t* to,
                              float* scomp, float* score);

/* ---------------------------------------------------------------------------
 * Batched form: many assemblies resident together.  One build runs every
 * stage as one launch over all assemblies' work items (rows of all
 * assemblies are numbered consecutively; each assembly keeps its own regime,
 * b0, calibration and parameters), so each assembly's result is identical to
 * building it alone.  The per-assembly results have contig ids local to the
 * assembly.  A single-assembly rbg_asm is a batch of one.
 * No counterpart in the reference (it builds one assembly per process).
 * ------------------------------------------------------------------------- */
typedef struct rbg_assembly_in {
  const char* seqs;          /* contigs back to back, as rbg_sketch_winners */
  const uint64_t* offsets;   /* n + 1 */
  uint32_t n;
  const float* depth;        /* n x S raw per-sample means, or NULL when S == 0 */
  uint32_t S;                /* at most 64 */
} rbg_assembly_in;

/* Work schedule of the pair pass: 256-column tiles per work item of the
 * abundance gate, and rows per work item of the all-pairs counts. */
typedef struct rbg_schedule {
  uint32_t gate_tiles_per_item;   /* default 16, 1..64 */
  uint32_t dense_rows_per_item;   /* default 64, 1..4096 */
} rbg_schedule;

typedef struct rbg_batch_stats {
  uint64_t candidates;       /* int8 prefilter candidates (all fused assemblies) */
  uint64_t edges;            /* edges of all assemblies */
  double ms_total;           /* device time of the build */
  int replayed;              /* 1 when a captured graph was replayed */
} rbg_batch_stats;

rbg_status rbg_batch_create(rbg_ctx* ctx, const rbg_assembly_in* asms, uint32_t count,
                            double min_cv, rbg_asm** out);
rbg_status rbg_batch_set_schedule(rbg_asm* b, const rbg_schedule* s);
/* flags of params: RBG_BUILD_CAPTURE (RBG_EDGES_ON_DEVICE is implied: the
 * edges stay in HBM until rbg_batch_result). */
rbg_status rbg_batch_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* params,
                           rbg_batch_stats* stats);
/* Autotuner of the batch's work schedule: builds the batch (eagerly) under
 * candidate schedules, one dimension at a time (gate tiles per work item in
 * {2, 4, 8, 16, 32} when the batch has fused assemblies, then dense rows per
 * work item in {16, 32, 64, 128, 256} when it has small ones), times each
 * build with HIP events on the context's stream (best of `reps`), and keeps
 * the fastest schedule in the batch (and in *best, with its time in
 * *best_ms).  The batch is left built under that schedule.  Results do not
 * depend on the schedule; only the time does. */
rbg_status rbg_batch_autotune(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* params, uint32_t reps,
                              rbg_schedule* best, double* best_ms);
/* Assembly `index` of the last build: scalars and the edge list on the host. */
rbg_status rbg_batch_result(rbg_ctx* ctx, rbg_asm* b, uint32_t index, rbg_graph_result* out);

/* ---------------------------------------------------------------------------
 * Served form: a request loop on one device.  rbg_server_submit starts a
 * session with an assembly; rbg_server_extend adds a batch of contigs to a
 * session's assembly (appended after its earlier contigs, with the session's
 * sample count).  A worker thread serves pending requests in steps: each
 * step is one batched build over the requests admitted to it (at most one per
 * session, in submission order), so requests that arrive while others are on
 * the device share the next step's launches.  Admission keeps the device
 * memory of the resident sessions plus the step's working set within
 * hbm_budget; a request that cannot fit even alone fails with RBG_ERR_BUDGET.
 * A session keeps its sketches, depth rows and (fused regime) gated
 * candidate lists on the device: an extension uploads and sketches only its
 * new contigs and evaluates only the pairs that involve them.  Every result
 * equals rbg_build_graph on the session's whole assembly as of that request
 * (ids local to the session: earlier contigs first, in order).
 * Input buffers are caller-owned and must stay valid until the request's
 * rbg_server_wait returns.  All calls are thread-safe.
 * No counterpart in the reference (it builds one assembly per process).
 * ------------------------------------------------------------------------- */
typedef struct rbg_server rbg_server;

typedef struct rbg_server_config {
  uint64_t hbm_budget;          /* device bytes for sessions + step working set (0: 90% of free) */
  uint32_t max_step_requests;   /* requests per step (default 64) */
  uint32_t max_step_rows;       /* contigs per step, < 2^24 (default 2^22) */
  rbg_graph_params params;      /* build parameters of every request (flags ignored) */
  rbg_schedule schedule;        /* pair-pass work schedule (see rbg_batch_autotune) */
} rbg_server_config;

typedef struct rbg_request_info {
  uint64_t session;
  uint32_t n;              /* contigs of the session's assembly after this request */
  uint32_t n_new;          /* contigs this request added */
  int incremental;         /* 1: only pairs with a new contig were evaluated */
  int status;              /* the request's rbg_status */
  uint64_t step;           /* the step that served it (1-based) */
  uint32_t step_requests;  /* requests in that step */
  uint64_t pass_pairs;     /* pairs the step evaluated for it */
  double t_submit, t_admit, t_done;  /* seconds on the server's clock */
  double latency_ms;       /* t_done - t_submit */
  char error[160];         /* message when status != 0 */
} rbg_request_info;

typedef struct rbg_server_stats {
  uint64_t steps, requests, sessions, pending, running;
  uint64_t graph_pairs;    /* sum over served requests of n(n-1)/2 of their graphs */
  uint64_t pass_pairs;     /* pairs evaluated */
  uint64_t resident_bytes, workspace_bytes, budget;
  double busy_ms;          /* wall time inside steps, split into: */
  double upload_ms;        /*   host preparation and upload of the steps' inputs */
  double build_ms;         /*   the batched builds on the device */
  double result_ms;        /*   edge lists to the host */
  double save_ms;          /*   sessions' device state kept for later extensions */
} rbg_server_stats;

void rbg_server_default_config(rbg_server_config* c);
rbg_status rbg_server_start(int device, const rbg_server_config* cfg, rbg_server** out);
/* Fails the requests still pending, waits for the running step, frees all. */
void rbg_server_stop(rbg_server* s);
const char* rbg_server_last_error(const rbg_server* s);
rbg_status rbg_server_submit(rbg_server* s, const rbg_assembly_in* a, uint64_t* session,
                             uint64_t* request);
rbg_status rbg_server_extend(rbg_server* s, uint64_t session, const rbg_assembly_in* more,
                             uint64_t* request);
/* Blocks until the request is done (timeout_s < 0: no limit) and hands over
 * its result (free with rbg_graph_result_free) and info; returns the
 * request's status.  Each request is waited for once. */
rbg_status rbg_server_wait(rbg_server* s, uint64_t request, double timeout_s,
                           rbg_graph_result* out, rbg_request_info* info);
/* Frees the session's device state once its queued requests are done. */
rbg_status rbg_server_release(rbg_server* s, uint64_t session);
rbg_status rbg_server_get_stats(rbg_server* s, rbg_server_stats* out);

/* ---------------------------------------------------------------------------
 * Stage (d) alone (src/rabbitbin.cpp:5482-5509): edge weights for a given
 * edge list.  unit/raw are n x S, colnorm S.  score is the reference's
 * arithmetic (the correlation summed as depth_corr_fast sums it on an AVX-512
 * build, src/rabbitbin.cpp:705-746; any S); score64, if non-NULL, the same
 * weight with the correlation summed in FP64 throughout.
 * ------------------------------------------------------------------------- */
rbg_status rbg_edge_weights(rbg_ctx* ctx, const uint32_t* from, const uint32_t* to,
                            const float* scomp, uint64_t ne, const float* unit,
                            const float* raw, const double* colnorm, uint32_t n,
                            uint32_t S, double w_comp, double min_edge_weight,
                            double neg_depth_thr, float* score, double* score64);

/* rbg_edge_weights for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and kernel as
 * rbg_edge_weights; the edge list, depth rows and results are staged in a
 * fixed 64 MiB device array that each GPU holds until the process ends, so
 * the call allocates nothing.  Inputs needing more (24 bytes per edge plus
 * 8 bytes per depth-row entry, each part 256-byte aligned) return
 * RBG_ERR_UNSUPPORTED.  All calls on one GPU share the array: calls that
 * overlap on one GPU race and must be serialised by the caller. */
rbg_status rbg_edge_weights_noalloc(rbg_ctx* ctx, const uint32_t* from, const uint32_t* to,
                                    const float* scomp, uint64_t ne, const float* unit,
                                    const float* raw, const double* colnorm, uint32_t n,
                                    uint32_t S, double w_comp, double min_edge_weight,
                                    double neg_depth_thr, float* score, double* score64);

/* ---------------------------------------------------------------------------
 * Host-side pieces, callable without a GPU.
 * rbg_sketch_tables: the per-(element, register) tables the sketch kernel
 *   consumes.  Element x is the x-th canonical 4-mer in ascending code order.
 *   T[x*m+k]: value of x's proposal to register k at weight 1 (the weighted
 *   value is (1/w)*T, src/probmh.cpp:651-677); W[x*m+k]: x's winner identity
 *   at k (src/probmh.cpp:667); cls[k*136+x]: class byte (smallest element with
 *   the same 16-bit fold at k, 0xFF when that fold is 0).  Any may be NULL.
 * rbg_calibrate: simCutoff from per-row maximum match counts, as
 *   calibrate_sim_cutoff(coverage, true) (regime 1, src/impl/rb_graph.cpp:1362)
 *   or the 10 x 2500 sampled calibration of gen_fused_calib_graph (regime 2,
 *   src/impl/rb_graph.cpp:502-525, 1243-1254), consuming the glibc rand()
 *   stream that srand(rand_seed) starts.
 * ------------------------------------------------------------------------- */
rbg_status rbg_sketch_tables(uint32_t m, uint64_t seed, double* T, uint64_t* W, uint8_t* cls);
rbg_status rbg_calibrate(const uint32_t* rowmax_cnt, uint32_t n, uint32_t m, double b0,
                         double coverage, uint32_t rand_seed, int regime, double* sim_cutoff);

/* ---------------------------------------------------------------------------
 * RabbitBin's graph cache file, RBCACHE1 version 3 (src/impl/rb_cache.cpp:
 * 94-175 writes it, :178-322 reads it; `rabbitbin bin --save-cache` /
 * `--load-cache`).  rbg_cache_write writes a graph built here (with the
 * caller's contig names, lengths and depth state) in the reference's layout;
 * rbg_cache_read reads versions 1-3 (size_t endpoints of versions 1-2 are
 * range-checked into 32-bit ids, as the reference does) into arrays it
 * allocates (release with rbg_cache_free).  Host only, no GPU.  Names are
 * stored back to back: name i is names[name_off[i] .. name_off[i+1]).  An
 * array may be NULL (written as empty).  Returns RBG_ERR_UNSUPPORTED for a
 * file that is not an RBCACHE1 cache of a supported version and ABI, and
 * RBG_ERR_ARG for a truncated or inconsistent one.
 * ------------------------------------------------------------------------- */
typedef struct rbg_cache {
  uint32_t version;                 /* read: the file's version; written: 3 */
  uint64_t nobs, nobs1, samples;    /* large and small contigs, depth samples */
  uint64_t seed, total_size, total_size1;
  double sim_cutoff, b0;            /* simCutoff (per-mille), g_pmh_baseline */
  uint64_t min_contig, min_small_contig;
  int has_depth;
  char* names;        uint64_t* name_off;        /* contig_names (nobs) */
  char* small_names;  uint64_t* small_name_off;  /* small_contig_names (nobs1) */
  uint64_t* seq_lens;                            /* nobs */
  uint64_t* small_seq_lens;                      /* nobs1 */
  uint64_t depth_rows, depth_cols;   float* depth;              /* depth_matrix (ranked) */
  uint64_t small_depth_rows, small_depth_cols; float* small_depth;
  uint64_t n_large_means;  float* large_means;   /* g_large_means */
  uint64_t n_depth_raw;    float* depth_raw;     /* g_depth_raw */
  uint64_t n_colnorm;      double* colnorm;      /* g_depth_colnorm */
  uint64_t n_edges;  uint32_t* from;  uint32_t* to;  float* scomp;  /* the graph */
  int has_snv;  uint64_t n_snv;  float* snv_pi;  uint32_t* snv_nsnv;  uint32_t* snv_cov;
} rbg_cache;
rbg_status rbg_cache_write(const char* path, const rbg_cache* c);
rbg_status rbg_cache_read(const char* path, rbg_cache* out);
void rbg_cache_free(rbg_cache* c);

/* ---------------------------------------------------------------------------
 * Fisher label propagation (host): cluster_by_propagation with its default
 * options (src/impl/rb_cluster.cpp:376-717) on the graph the reference
 * prepares for it (src/rabbitbin.cpp:5597-5651): the edges with score > 0,
 * scores clamped to 1 - 1e-6, each node's incident edges in ascending
 * neighbour id, the node order iota(n) shuffled by std::shuffle with
 * std::default_random_engine(seed) (seed = --seed).  membership (n) receives
 * each node's label (a node id); rounds_out (may be NULL) the number of
 * sweeps.  With no positive edge every node keeps its own label.  Scores must
 * be non-negative; endpoints distinct and below n.  Sequential by definition
 * (each update sees the labels changed before it in the order), so it runs
 * on the host.  rbg_lpa_node_order returns that order alone.
 * ------------------------------------------------------------------------- */
rbg_status rbg_label_propagation(const uint32_t* from, const uint32_t* to, const float* score, uint64_t ne,
                                 uint32_t n, uint64_t seed, uint32_t* membership, uint32_t* rounds_out);
rbg_status rbg_lpa_node_order(uint32_t n, uint64_t seed, uint32_t* order);

/* ---------------------------------------------------------------------------
 * Frozen-core recruitment: recruit_to_frozen_cores (src/impl/rb_split.cpp:
 * 729-919), the bins' recruitment of unbinned contigs by mean Spearman score.
 * Bins come in any order as (bin_id[b], members bin_off[b] .. bin_off[b+1]),
 * members in the bin's own order: member c < n_large is large contig c,
 * n_large + s is small contig s.  The cores are the bins, by ascending id,
 * whose length (the sum of all members' lengths) reaches core_floor and
 * which hold at least max(2, min_cs) large members with a nonzero unit row;
 * a core's centroid is the mean of those unit rows and its threshold their
 * mean pairwise Spearman correlation in the reference's closed form
 * (||sum u||^2 - sum ||u||^2) / (n(n-1)), clamped to [-1, 1].  Every large
 * contig in no bin, then every small contig (its unit row from its raw
 * depths, ranked as the reference ranks them), is scored against every core
 * (u . centroid in FP64, summed in sample order) and goes to the one core
 * whose threshold it reaches; with none or several it stays unassigned.
 * assign_large (n_large) and assign_small (n_small) receive the core's bin
 * id, or -1 (binned large contigs: -1).  Appending the assigned contigs to
 * their bins in ascending order, large before small, gives the reference's
 * bins.  unit_large: the large contigs' unit rows (n_large x S, as
 * rbg_depth_features returns them); depth_small: the small contigs' raw
 * depths (n_small x S).  The contig x core scores run on the device.
 * Bin ids must be distinct and non-negative.  S < 2 or no core: nothing is
 * assigned (as in the reference).  S > 64: RBG_ERR_UNSUPPORTED.  stats may be
 * NULL.
 * ------------------------------------------------------------------------- */
typedef struct rbg_recruit_stats {
  uint64_t cores;
  uint64_t long_assigned, long_ambiguous, long_unmatched;
  uint64_t short_assigned, short_ambiguous, short_unmatched;
} rbg_recruit_stats;
rbg_status rbg_recruit_frozen_cores(rbg_ctx* ctx, uint32_t S, uint32_t n_large, const float* unit_large,
                                    const uint64_t* len_large, uint32_t n_small, const float* depth_small,
                                    const uint64_t* len_small, uint32_t n_bins, const int32_t* bin_id,
                                    const uint64_t* bin_off, const uint64_t* members, uint64_t core_floor,
                                    uint64_t min_cs, int32_t* assign_large, int32_t* assign_small,
                                    rbg_recruit_stats* stats);

/* Chance baseline b0 over 16-bit winners (src/rabbitbin.cpp:1340-1369). */
rbg_status rbg_pmh_baseline(rbg_ctx* ctx, const uint16_t* win16, uint32_t n, uint32_t m,
                            double* b0);

/* ---------------------------------------------------------------------------
 * Ingest on the device side of the boundary: the assembly FASTA and the depth
 * file are parsed in HBM, and the large contigs stay there for the graph.
 *
 * rbg_ingest_fasta replaces parse_fasta_mmap_parallel (src/rabbitbin.cpp:
 * 1890-2248) with its default options: the file's bytes (e.g. an mmap) are
 * copied to the device, the records found and classified there (large:
 * >= min_contig bases, small: >= min_small_contig, the rest tiny and
 * dropped), and the large contigs' bases packed on the device in file order
 * (multi-line records compacted, '\r' before a line end dropped).  Names
 * (the header up to its first blank) and lengths come back to the host.
 *
 * rbg_ingest_depth replaces parse_depth_buf (src/rabbitbin.cpp:2832-2942)
 * and the merge by name (:4495-4545): the depth file (header line, then per
 * contig name, contigLen, totalAvgDepth and S mean / variance pairs;
 * cv_ext: name, contigLen and S means) is parsed on the device with the
 * C library's strtod / strtof rounding (exact), rows whose contigLen is below
 * min_depth_len skipped (the reference's default prefilter, min_depth_len =
 * min_small_contig; 0 keeps every row), the first row of a name kept, and
 * each large and small contig given its row.  S = 0 derives the sample count
 * from the first data line as the reference does (ncols - 3, halved; ncols - 1
 * with cv_ext).  Returns RBG_ERR_ARG when a large contig has no row (the
 * reference leaves that row undefined), and RBG_ERR_UNSUPPORTED for
 * duplicate large-contig names (the reference drops the later ones), for
 * hexadecimal floats, NaN payloads, more than 19 significant digits or
 * magnitudes beyond the exact conversion's range (about 1e-30 .. 1e38).
 *
 * rbg_ingest_assembly makes a resident assembly (rbg_asm) of the large
 * contigs and their depth rows without taking the bases back to the host;
 * it then builds graphs like any rbg_asm_create assembly.
 * ------------------------------------------------------------------------- */
typedef struct rbg_ingest rbg_ingest;
typedef struct rbg_ingest_info {
  uint64_t records;            /* FASTA records with a header line (the reference's num_seqs) */
  uint32_t n_large, n_small;   /* large and small contigs, file order */
  uint64_t large_bases;        /* total length of the large contigs */
  const char* names;           /* large contigs' names back to back: name i = */
  const uint64_t* name_off;    /*   names[name_off[i] .. name_off[i+1]) */
  const uint64_t* lens;        /* large contigs' lengths */
  const char* small_names;
  const uint64_t* small_name_off;
  const uint64_t* small_lens;
  uint32_t S;                  /* depth samples (0 before rbg_ingest_depth) */
  uint64_t depth_rows;         /* rows of the depth file's name map */
  double ms_fasta, ms_depth;   /* device time of the parses (upload included) */
} rbg_ingest_info;
rbg_status rbg_ingest_fasta(rbg_ctx* ctx, const char* data, uint64_t len, uint64_t min_contig,
                            uint64_t min_small_contig, rbg_ingest** out);
rbg_status rbg_ingest_depth(rbg_ctx* ctx, rbg_ingest* ig, const char* data, uint64_t len, uint32_t S,
                            int cv_ext, uint64_t min_depth_len);
/* The pointers stay valid until rbg_ingest_free. */
rbg_status rbg_ingest_get(const rbg_ingest* ig, rbg_ingest_info* info);
/* Copies to the host (any may be NULL): the large contigs' bases (large_bases
 * bytes), their offsets (n_large + 1), and the depth rows of the large and
 * small contigs (n x S raw means). */
rbg_status rbg_ingest_copy(rbg_ctx* ctx, const rbg_ingest* ig, char* seqs, uint64_t* offsets, float* depth,
                           float* small_depth);
rbg_status rbg_ingest_assembly(rbg_ctx* ctx, const rbg_ingest* ig, double min_cv, rbg_asm** out);
void rbg_ingest_free(rbg_ingest* ig);

#ifdef __cplusplus
}
#endif
#endif /* RBG_H */
