// librbg: rbg_edge_weights_noalloc, the edge-weight entry point for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It computes what
// rbg_edge_weights computes, with the same kernel, but stages the edge list,
// the depth rows and the results in `scratch`, a file-scope device array of
// kScratchBytes that each GPU holds until the process ends, instead of
// buffers allocated per call.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// array's device address is taken once per device with hipGetSymbolAddress
// (at rbg_open, with that device current) and only that address is used.
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: calls that overlap on one GPU race, also when they come from
// different contexts, so callers serialise them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "rbg_build.h"

namespace rbg {

constexpr size_t kScratchBytes = 64u << 20;  // 64 MiB per GPU

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Takes the device address of `scratch` on the current device (once per
// device).  Called by rbg_open with `device` current.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

}  // namespace rbg

using namespace rbg;

extern "C" {

rbg_status rbg_edge_weights_noalloc(rbg_ctx* ctx, const uint32_t* from, const uint32_t* to,
                                    const float* scomp, uint64_t ne, const float* unit, const float* raw,
                                    const double* colnorm, uint32_t n, uint32_t S, double w_comp,
                                    double min_edge_weight, double neg_depth_thr, float* score,
                                    double* score64) {
  if (!ctx || (ne && (!from || !to || !scomp || !score))) return RBG_ERR_ARG;
  if (S > 1 && (!unit || !raw)) return RBG_ERR_ARG;
  if (ne == 0) return RBG_OK;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return ctx->fail(RBG_ERR_ARG, "device ordinal");
  uint8_t* base;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = (uint8_t*)g_scratch[ctx->device];
  }
  if (!base) return ctx->fail(RBG_ERR_HIP, "scratch of this device not set up (rbg_open failed?)");
  for (uint64_t e = 0; e < ne; ++e)
    if (from[e] >= n || to[e] >= n) return ctx->fail(RBG_ERR_ARG, "edge endpoint out of range");
  const bool rows = S > 1;
  auto too_big = [&](size_t need_mib) {
    return ctx->fail(RBG_ERR_UNSUPPORTED,
                     "rbg_edge_weights_noalloc: " + std::to_string(ne) + " edges and " + std::to_string(n) +
                         " x " + std::to_string(S) + " depth rows need " + std::to_string(need_mib) +
                         " MiB of scratch or more; the limit is " + std::to_string(kScratchBytes >> 20) +
                         " MiB (use rbg_edge_weights)");
  };
  if (ne > kScratchBytes / 16 || (rows && (uint64_t)n * S > kScratchBytes / 8))
    return too_big(kScratchBytes >> 20);
  // layout in scratch: from, to, scomp, score, score64, unit, raw, colnorm
  size_t at = 0;
  const size_t o_from = at;
  at = align256(at + ne * 4);
  const size_t o_to = at;
  at = align256(at + ne * 4);
  const size_t o_sc = at;
  at = align256(at + ne * 4);
  const size_t o_w = at;
  at = align256(at + ne * 4);
  const size_t o_w64 = at;
  at = align256(at + (score64 ? ne * 8 : 0));
  const size_t o_unit = at;
  at = align256(at + (rows ? (size_t)n * S * 4 : 0));
  const size_t o_raw = at;
  at = align256(at + (rows ? (size_t)n * S * 4 : 0));
  const size_t o_cn = at;
  at = align256(at + (rows && colnorm ? (size_t)S * 8 : 0));
  if (at > kScratchBytes) return too_big((at + (1u << 20) - 1) >> 20);
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_from, from, ne * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_to, to, ne * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_sc, scomp, ne * 4, hipMemcpyHostToDevice, st));
  if (rows) {
    RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_unit, unit, (size_t)n * S * 4, hipMemcpyHostToDevice, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_raw, raw, (size_t)n * S * 4, hipMemcpyHostToDevice, st));
    if (colnorm) RBG_TRY_HIP(ctx, hipMemcpyAsync(base + o_cn, colnorm, (size_t)S * 8, hipMemcpyHostToDevice, st));
  }
  // Every output word is written by the kernel, so nothing is cleared.
  RBG_TRY_HIP(ctx, launch_edge_weights((const uint32_t*)(base + o_from), (const uint32_t*)(base + o_to),
                                       (const float*)(base + o_sc), ne,
                                       rows ? (const float*)(base + o_unit) : nullptr,
                                       rows ? (const float*)(base + o_raw) : nullptr,
                                       rows && colnorm ? (const double*)(base + o_cn) : nullptr, S, w_comp,
                                       min_edge_weight, neg_depth_thr, (float*)(base + o_w),
                                       score64 ? (double*)(base + o_w64) : nullptr, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(score, base + o_w, ne * 4, hipMemcpyDeviceToHost, st));
  if (score64) RBG_TRY_HIP(ctx, hipMemcpyAsync(score64, base + o_w64, ne * 8, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  return RBG_OK;
}

}  // extern "C"
