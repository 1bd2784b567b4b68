// Kernel launchers (internal to librbg).  All launch on `st` and return the
// launch status; shapes are checked by the callers in rbg_api.cpp.
//
// A graph build runs over a batch of assemblies (one assembly is a batch of
// one).  Rows (contigs) of all assemblies are numbered consecutively; every
// per-row array is indexed by that global row, asm_of[row] gives the row's
// assembly, and AsmDesc holds the assembly's parameters.  Pairs, candidate
// lists, top-K ids and edges carry global rows.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rbg {

// Scalars of one assembly's graph build, kept in device memory and written
// by the kernels that compute them (b0 and the sv tables, then calibration).
struct BuildScalars {
  double b0;
  double sim_cutoff;   // per-mille, as the reference's simCutoff
  uint32_t cmin;       // first count whose float similarity exceeds simCutoff / 1000
  uint32_t nlevels;    // distinct float similarities
  uint32_t rand_used;  // rand() values consumed by calib_converge
  uint32_t pad;
};

// Per-assembly parameters of a batch (device-resident, one per assembly).
struct AsmDesc {
  uint32_t row0;      // first global row of the assembly
  uint32_t n;         // contigs
  uint32_t S;         // depth samples
  uint32_t regime;    // 1: all pairs (n x n counts); 2: abundance-gated candidates
  float thr;          // exact abundance-gate threshold (-inf: gate off)
  int32_t qcut;       // int8 prefilter cut
  int32_t qcut_hi;    // int8 dots from here on pass the exact gate (INT32_MAX: none)
  uint32_t acc1_end;  // odd chunks below this accumulate into the second register
  uint32_t pair0;     // b0 pair draws: [pair0, pair0 + np) of the batch list
  uint32_t np;
  uint32_t colnorm0;  // first of the assembly's S colnorm entries
  uint32_t jmin;      // regime 1: pairs with j < jmin (local) were counted by an earlier build
  uint64_t mat0;      // regime 1: first entry of the assembly's n x n count matrix
  double w_comp;      // composition weight
  double b0_in;       // b0 given by the caller; NaN: the median of the pair draws
  uint32_t dense;     // 1: every pair counted into the n x n matrix (regime 1, and regime 2
                      // without the abundance gate); 0: gated candidate lists
  uint32_t pad0;
};

// Per-assembly tables: svd/svf/level hold m+1 entries per assembly; the
// calibration plan holds kCalibRandDev rand() values and 11 round offsets
// per assembly.
constexpr uint32_t kCalibRandDev = 2048;
// words of calibration histogram scratch per assembly (10 rounds x 520 counts)
constexpr uint32_t kCalibHistWords = 10 * 520;

// ---- frozen-core recruitment: out_id[r] = the id of the one core whose
// threshold unit row r's FP64 score reaches (sample order, fused multiply-
// adds), -1 with none, several or an all-zero row; out_amb[r] = 1 with
// several.  S in [1, 64].
hipError_t launch_recruit_match(const float* unit, uint32_t n, uint32_t S, const double* cent, const double* thr,
                                const int32_t* core_id, uint32_t ncores, int32_t* out_id, uint8_t* out_amb,
                                hipStream_t st);

// ---- stage (a)
// Sketches of contigs [0, n) of (seqs, offs); contig c goes to row out_row[c]
// of planes / wx (row c when out_row is null).
hipError_t launch_sketch(const uint8_t* seqs, const uint64_t* offs, uint32_t n, uint32_t m,
                         uint32_t W, const double* sT, const uint8_t* sX, const uint8_t* cls,
                         const uint8_t* codes, uint8_t* wx, uint32_t mpad, uint32_t* planes,
                         const uint32_t* out_row, hipStream_t st);
// Bit planes of rows [row0, row0 + n) from per-register class bytes (n x mpad,
// registers >= m and unset ones 0xFF).
hipError_t launch_planes_from_classes(const uint8_t* cls, uint32_t n, uint32_t m, uint32_t mpad,
                                      uint32_t W, uint32_t row0, uint32_t* planes, hipStream_t st);
hipError_t launch_list_counts(const uint32_t* planes, uint32_t W, const uint32_t* pi,
                              const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st);
hipError_t launch_win16_counts(const uint16_t* win16, uint32_t m, const uint32_t* pi,
                               const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st);

// ---- b0, sv tables, calibration (one block per assembly)
hipError_t launch_baseline_tables(const uint32_t* pcnt, const AsmDesc* desc, uint32_t nasm,
                                  uint32_t m, BuildScalars* bs, double* svd, float* svf,
                                  uint16_t* level, hipStream_t st);
// simCutoff and cmin from the per-row maxima (regime 1: all rows of the
// assembly; regime 2: its round members), consuming its rand() values in order.
// max_items: the most contigs one assembly samples (its rows, or its members);
// hist: nasm x 10 x 520 words of scratch for the rounds' count histograms.
hipError_t launch_calibrate(const uint32_t* rowmax, const AsmDesc* desc, uint32_t nasm, uint32_t m,
                            const double* svd, const float* svf, const int32_t* rnd,
                            const uint32_t* members, const uint32_t* round_off, uint32_t max_items,
                            uint32_t* hist, double coverage, BuildScalars* bs, hipStream_t st);

// ---- stage (b), regime 1: n x n counts per assembly (each unordered pair
// computed once and written at both places; the diagonal is not written)
// Work items (assembly, column block of 256, row block, rows per block), only
// blocks holding some pair r < j with j >= the assembly's jmin.
void dense_work_items(const std::vector<AsmDesc>& desc, uint32_t rows_per_item,
                      std::vector<uint32_t>& items);
hipError_t launch_dense_counts(const uint32_t* planes, uint32_t W, const AsmDesc* desc,
                               const uint32_t* items, uint32_t nitems, uint16_t* mat,
                               hipStream_t st);

// ---- stage (b), regime 2: abundance gate + register compare
// Work items (assembly, bi << 16 | bj0, tile count, jmin): a 256-row stripe bi
// of the assembly against up to tiles_per_item 256-column tiles from bj0, over
// the pairs i < j with jmin <= j (pair-pass positions; jmin > 0 when the
// pairs among the first jmin rows were evaluated by an earlier build).
void gate_work_items(uint32_t a, uint32_t n, uint32_t jmin, uint32_t tiles_per_item,
                     std::vector<uint32_t>& full, std::vector<uint32_t>& part);
// q holds each assembly's int8 rows in its pair-pass order (kpad bytes per
// row); perm maps that order back to global rows.  Candidates are written as
// (min row, max row).
hipError_t launch_gate_mfma(const int8_t* q, uint32_t kpad, const AsmDesc* desc,
                            const uint32_t* perm, const uint32_t* items, uint32_t nitems,
                            uint64_t* cand, unsigned long long* ncand, uint64_t cap,
                            hipStream_t st);
// Exact gate + register compare of the min(*ncand, cap) candidates, in place
// (rejected -> all ones, survivors -> count in the low 16 bits).  Each
// survivor t = (i, j) takes the next entry of row i's and of row j's list:
// deg counts the entries per row and slot[2t], slot[2t+1] receive the
// survivor's positions in row i's and row j's list (cap < 2^31).  unit rows
// are zero-padded to 16*NC floats.  n_asm: assemblies in the build (with one,
// its gate parameters are read once instead of per candidate).
hipError_t launch_gate_compare(uint64_t* cand, uint64_t cap, const unsigned long long* ncand,
                               const float* unit, uint32_t NC, const uint8_t* anynz,
                               const AsmDesc* desc, const uint16_t* asm_of, uint32_t n_asm,
                               const uint32_t* planes, uint32_t W, uint32_t* deg, uint32_t* slot,
                               hipStream_t st);
// Incremental assemblies: a served session's candidate lists from its earlier
// build (CSR over its first n_old rows, local ids), merged into this build's
// lists: merge_deg sets deg to their lengths (before the exact gate claims
// the new pairs' slots after them), merge_lists copies them to the head of
// each row's new list (after the offsets are scanned).
struct MergeDesc {
  const uint64_t* loff;  // n_old + 1
  const uint64_t* list;  // entries (count << 32 | local id)
  uint32_t row0, n_old;
};
hipError_t launch_merge_deg(const MergeDesc* md, uint32_t nmd, uint32_t max_old, uint32_t* deg,
                            hipStream_t st);
hipError_t launch_merge_lists(const MergeDesc* md, uint32_t nmd, uint32_t max_old, const uint64_t* off,
                              uint64_t* list, hipStream_t st);
// Rows [row0, row0 + n) of a build's lists as a local CSR (loff: n + 1 offsets
// from off[row0]; dst: entries with ids made local).
hipError_t launch_extract_lists(const uint64_t* off, const uint64_t* list, uint32_t row0, uint32_t n,
                                uint64_t* loff, uint64_t* dst, hipStream_t st);
// Survivors into both rows' lists at off[row] + their slots (no atomics).
hipError_t launch_scatter(const uint64_t* cand, uint64_t cap, const unsigned long long* ncand,
                          const uint32_t* slot, const uint64_t* off, uint64_t* list, hipStream_t st);

// ---- stage (c)
// Per-row top-K (sv desc, id asc) over the row's partners: regime 1 rows read
// the dense count matrix, regime 2 rows their candidate list (off, list).
// queue: nrows + 2 words, its first nstatic entries the rows of the dense
// assemblies (written by the plan).
hipError_t launch_select(const uint64_t* list, const uint64_t* off, const uint16_t* mat,
                         const uint8_t* anynz, const AsmDesc* desc, const uint16_t* asm_of,
                         uint32_t nrows, const uint16_t* level, uint32_t m, uint32_t K,
                         uint32_t* topk_id, uint16_t* topk_cnt, uint32_t* topk_len,
                         uint32_t* rowmax, uint32_t* fence, uint32_t FW, uint32_t* queue,
                         uint32_t nstatic, hipStream_t st);
// Fence rows of the selected lists: FW words per row (16-id stride, padded
// with 0xFFFFFFFF); FW = 16 * ceil(ceil(K/16) / 16).
inline uint32_t fence_words(uint32_t K) { return 16u * (((K + 15u) / 16u + 15u) / 16u); }
// Edge emission: pass 0 marks kept top-K entries (keep: nrows x ceil(K/64)
// u64) and counts them per row; pass 1 writes them at the rows' offsets.
// Rows are visited in `order` (a permutation of the rows; identity when null).
hipError_t launch_emit_mark(const uint32_t* topk_id, const uint16_t* topk_cnt,
                            const uint32_t* topk_len, const uint32_t* fence, uint32_t FW,
                            uint32_t nrows, uint32_t K, const BuildScalars* bs,
                            const AsmDesc* desc, const uint16_t* asm_of, const uint32_t* order,
                            uint64_t* keep, uint32_t* row_edges, hipStream_t st);
// Pass 1 also writes each edge's sComp from its count (the row's assembly's svf).
hipError_t launch_emit_write(const uint32_t* topk_id, const uint16_t* topk_cnt,
                             const uint32_t* topk_len, uint32_t nrows, uint32_t K,
                             const uint64_t* keep, const uint64_t* row_off, const uint16_t* asm_of,
                             uint32_t m, const float* svf, uint32_t* efrom, uint32_t* eto,
                             uint16_t* ecnt, float* escomp, hipStream_t st);

// ---- stage (d)
// Edge weights of the min(*ne, cap) edges of a batch: unit/raw rows padded to
// 16*NC floats, each edge with its assembly's S, w_comp and colnorm.  The
// correlation is summed in the reference's order; edges whose FP64 weight or
// correlation lies within near_tol of a cut are appended (edge index, FP64
// weight) to near_edge / near_w64 at *near_cnt (entries past near_cap are
// counted, not written).
hipError_t launch_edge_weights_batch(const uint32_t* efrom, const uint32_t* eto,
                                     const float* scomp, uint64_t cap, const uint64_t* ne,
                                     const float* unit, const float* raw, uint32_t NC,
                                     const double* colnorm, const AsmDesc* desc,
                                     const uint16_t* asm_of, double min_edge, double neg_thr,
                                     double near_tol, float* score, unsigned long long* near_cnt,
                                     uint64_t near_cap, uint64_t* near_edge, double* near_w64,
                                     hipStream_t st);
// Edge weights of an explicit edge list (unpadded n x S rows, host count).
hipError_t launch_edge_weights(const uint32_t* efrom, const uint32_t* eto, const float* scomp,
                               uint64_t ne, const float* unit, const float* raw,
                               const double* colnorm, uint32_t S, double w_comp, double min_edge,
                               double neg_thr, float* score, double* score64, hipStream_t st);

// ---- utilities
hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t n, void* tmp,
                                     size_t& tmp_bytes, hipStream_t st);
// Fills count 32-bit words with value (a kernel, so it can sit in a captured graph).
hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t st);
// Per assembly a: out[4a..4a+3] = row_off at its first and one-past-last row,
// and the same for off (0 when off is null).
hipError_t launch_gather_bounds(const uint64_t* row_off, const uint64_t* off, const AsmDesc* desc,
                                uint32_t nasm, uint64_t* out, hipStream_t st);

}  // namespace rbg
