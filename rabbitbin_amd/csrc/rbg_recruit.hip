// Frozen-core recruitment: the contig x core scores on the device.
//
// One thread per candidate contig holds its unit row (S <= 64 floats) in
// registers; the cores' centroids (FP64) and thresholds pass through LDS in
// tiles of kCoreTile cores, every thread of the workgroup reading the same
// core at the same time (an LDS broadcast).  The score of a (contig, core)
// pair is the reference's sum over samples in sample order, one fused
// multiply-add per sample in FP64 on the unit value widened to double
// (src/impl/rb_split.cpp:739-762: `score += (double)unit[k] *
// core.centroid[k]`, which the reference build contracts).  A contig whose
// unit row is all zero matches nothing.  The thread counts the cores whose
// threshold its score reaches and keeps the last one's id: with exactly one,
// that id is the contig's core (the reference stops at a second match and
// reports it ambiguous; the count gives the same answer).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbg_kernels.h"

namespace rbg {
namespace {

constexpr uint32_t kRecruitMaxS = 64;
constexpr uint32_t kCoreTile = 32;

__global__ __launch_bounds__(256) void recruit_match_kernel(const float* __restrict__ unit, uint32_t n,
                                                            uint32_t S, const double* __restrict__ cent,
                                                            const double* __restrict__ thr,
                                                            const int32_t* __restrict__ core_id,
                                                            uint32_t ncores, int32_t* __restrict__ out_id,
                                                            uint8_t* __restrict__ out_amb) {
  __shared__ double tc[kCoreTile][kRecruitMaxS];
  __shared__ double tt[kCoreTile];
  __shared__ int32_t ti[kCoreTile];
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  const bool live = r < n;
  float u[kRecruitMaxS];
  bool nz = false;
#pragma unroll
  for (uint32_t k = 0; k < kRecruitMaxS; ++k) {
    u[k] = (live && k < S) ? unit[(size_t)r * S + k] : 0.f;
    nz = nz || u[k] != 0.f;
  }
  uint32_t hits = 0;
  int32_t id = -1;
  for (uint32_t c0 = 0; c0 < ncores; c0 += kCoreTile) {
    const uint32_t nt = min(kCoreTile, ncores - c0);
    __syncthreads();  // the previous tile is read
    for (uint32_t v = threadIdx.x; v < nt * S; v += 256) tc[v / S][v % S] = cent[(size_t)(c0 + v / S) * S + v % S];
    if (threadIdx.x < nt) {
      tt[threadIdx.x] = thr[c0 + threadIdx.x];
      ti[threadIdx.x] = core_id[c0 + threadIdx.x];
    }
    __syncthreads();
    for (uint32_t c = 0; c < nt; ++c) {
      double score = 0.0;
#pragma unroll
      for (uint32_t k = 0; k < kRecruitMaxS; ++k)
        if (k < S) score = __builtin_fma((double)u[k], tc[c][k], score);
      if (score >= tt[c]) {
        ++hits;
        id = ti[c];
      }
    }
  }
  if (live) {
    out_id[r] = nz && hits == 1 ? id : -1;
    out_amb[r] = nz && hits >= 2 ? 1 : 0;
  }
}

}  // namespace

hipError_t launch_recruit_match(const float* unit, uint32_t n, uint32_t S, const double* cent, const double* thr,
                                const int32_t* core_id, uint32_t ncores, int32_t* out_id, uint8_t* out_amb,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (S == 0 || S > kRecruitMaxS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(recruit_match_kernel, dim3((n + 255) / 256), dim3(256), 0, st, unit, n, S, cent, thr, core_id,
                     ncores, out_id, out_amb);
  return hipGetLastError();
}

}  // namespace rbg
