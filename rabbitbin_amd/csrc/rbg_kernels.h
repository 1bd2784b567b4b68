// Kernel launchers (internal to librbg).  All launch on `st` and return the
// launch status; shapes are checked by the callers in rbg_api.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rbg {

// Scalars of one graph build, kept in device memory and written by the
// kernels that compute them (b0 and the sv tables, then the calibration).
struct BuildScalars {
  double b0;
  double sim_cutoff;   // per-mille, as the reference's simCutoff
  uint32_t cmin;       // first count whose float similarity exceeds simCutoff / 1000
  uint32_t nlevels;    // distinct float similarities
  uint32_t rand_used;  // rand() values consumed by calib_converge
  uint32_t pad;
};

// b0 from the sampled pairs' counts, then svd/svf/level for counts 0..m.
hipError_t launch_baseline_tables(const uint32_t* pcnt, uint32_t np, uint32_t m, BuildScalars* bs,
                                  double* svd, float* svf, uint16_t* level, hipStream_t st);
// simCutoff and cmin from the per-row maxima (regime 1: all n rows; regime 2:
// the round members), consuming rnd[] in order.
hipError_t launch_calibrate(const uint32_t* rowmax, uint32_t n, uint32_t m, const double* svd,
                            const float* svf, const int32_t* rnd, uint32_t nrnd,
                            const uint32_t* members, const uint32_t* round_off, int regime,
                            double coverage, BuildScalars* bs, hipStream_t st);

hipError_t launch_sketch(const uint8_t* seqs, const uint64_t* offs, uint32_t n, uint32_t m,
                         uint32_t W, const double* sT, const uint8_t* sX, const uint8_t* cls,
                         const uint8_t* codes,
                         uint8_t* wx, uint32_t mpad, uint32_t* planes, hipStream_t st);
hipError_t launch_list_counts(const uint32_t* planes, uint32_t W, const uint32_t* pi,
                              const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st);
hipError_t launch_win16_counts(const uint16_t* win16, uint32_t m, const uint32_t* pi,
                               const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st);
hipError_t launch_dense_counts(const uint32_t* planes, uint32_t W, uint32_t n, uint16_t* out,
                               hipStream_t st);
// Work items of the fused pair pass: (bi << 16 | bj0, tile count) per item.
void gate_work_items(uint32_t n, std::vector<uint32_t>& items);
// q holds the int8 rows in pair-pass order; perm maps that order back to
// contig ids (candidates are written as (min id, max id)).
hipError_t launch_gate_mfma(const int8_t* q, uint32_t kpad, uint32_t n, int32_t qcut,
                            const uint32_t* perm, const uint32_t* items, uint32_t nitems,
                            uint64_t* cand, unsigned long long* ncand, uint64_t cap,
                            hipStream_t st);
// Exact gate + register compare of the candidates, in place (rejected ->
// all ones, survivors -> count in the low 16 bits; deg counts survivors per row).
hipError_t launch_gate_compare(uint64_t* cand, uint64_t nc, const float* unit, uint32_t NC,
                               const uint8_t* anynz, float thr, uint32_t acc1_end,
                               const uint32_t* planes, uint32_t W, uint32_t* deg, hipStream_t st);
hipError_t launch_scatter(const uint64_t* surv, uint64_t ns, const uint64_t* off, uint32_t* cursor,
                          uint64_t* list, hipStream_t st);
hipError_t launch_select_dense(const uint16_t* mat, const uint8_t* anynz, uint32_t n,
                               const uint16_t* level, uint32_t nlevels, uint32_t K,
                               uint32_t* topk_id, uint16_t* topk_cnt, uint32_t* topk_len,
                               uint32_t* rowmax, uint32_t* fence, uint32_t FW, hipStream_t st);
hipError_t launch_select_list(const uint64_t* list, const uint64_t* off, uint32_t n,
                              const uint16_t* level, uint32_t nlevels, uint32_t K,
                              uint32_t* topk_id, uint16_t* topk_cnt, uint32_t* topk_len,
                              uint32_t* rowmax, uint32_t* fence, uint32_t FW, hipStream_t st);
// Fence rows of the selected lists: FW words per row (16-id stride, padded
// with 0xFFFFFFFF); FW = 16 * ceil(ceil(K/16) / 16).
inline uint32_t fence_words(uint32_t K) { return 16u * (((K + 15u) / 16u + 15u) / 16u); }
// Edge emission: pass 0 marks kept top-K entries (keep: n x ceil(K/64) u64)
// and counts them per row; pass 1 writes them at the rows' offsets.
hipError_t launch_emit_mark(const uint32_t* topk_id, const uint16_t* topk_cnt,
                            const uint32_t* topk_len, const uint32_t* fence, uint32_t FW,
                            uint32_t n, uint32_t K, const BuildScalars* bs, int mutual,
                            uint64_t* keep, uint32_t* row_edges, hipStream_t st);
hipError_t launch_emit_write(const uint32_t* topk_id, const uint16_t* topk_cnt,
                             const uint32_t* topk_len, uint32_t n, uint32_t K, const uint64_t* keep,
                             const uint64_t* row_off, uint32_t* efrom, uint32_t* eto,
                             uint16_t* ecnt, hipStream_t st);
hipError_t launch_edge_weights(const uint32_t* efrom, const uint32_t* eto, const float* scomp,
                               uint64_t ne, const float* unit, const float* raw,
                               const double* colnorm, uint32_t S, double w_comp, double min_edge,
                               double neg_thr, float* score, double* score64, hipStream_t st,
                               bool padded = false, const uint64_t* ne_dev = nullptr);
// padded: unit and raw rows are zero-padded to 16*ceil(S/16) floats (S <= 64).
// ne_dev (optional): device-resident element count; ne is then the capacity.
hipError_t launch_cnt_to_sv(const uint16_t* c, uint64_t ne, const float* svf, float* out,
                            hipStream_t st, const uint64_t* ne_dev = nullptr);
hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t n, void* tmp,
                                     size_t& tmp_bytes, hipStream_t st);

}  // namespace rbg
