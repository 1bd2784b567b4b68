// librbg: the graph-build pipeline over a batch of device-resident assemblies
// (upload, plan, launch sequence, results), shared by the C ABI and the
// request server.
#include "rbg_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include <sys/mman.h>

using namespace rbg;

namespace rbg {

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

namespace {

constexpr size_t kStageHalf = 32u << 20;   // bytes per staging half
constexpr size_t kStageMin = 4u << 20;     // smaller copies go direct

// Workers kept for the life of the process (kCopyThreads - 1 of them, the
// calling thread makes up the rest): a parallel section's parts
// run on them and on the calling thread (creating threads per copy cost more
// than small copies).  One section at a time.
constexpr size_t kCopyThreads = 16;

class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool;  // never destroyed: workers sleep until exit
    return *pool;
  }
  void run(size_t n, const std::function<void(size_t)>& f) {
    if (n <= 1) {
      if (n) f(0);
      return;
    }
    std::lock_guard<std::mutex> section(call_mu_);
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &f;
      n_ = n;
      next_ = 0;
      left_ = n;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return left_ == 0; });
    fn_ = nullptr;
  }

 private:
  CopyPool() {
    for (size_t i = 0; i + 1 < kCopyThreads; ++i) std::thread([this] { work(); }).detach();
  }
  // Runs parts until none is left to take.
  void drain() {
    for (;;) {
      size_t t;
      const std::function<void(size_t)>* fn;
      {
        std::lock_guard<std::mutex> g(mu_);
        if (next_ >= n_) return;
        t = next_++;
        fn = fn_;
      }
      (*fn)(t);
      std::lock_guard<std::mutex> g(mu_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void work() {
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return next_ < n_; });
      }
      drain();
    }
  }
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t n_ = 0, next_ = 0, left_ = 0;
};

// dst = src (sub == 0) or dst[i] = src[i] - sub over 32-bit words, on up to
// kCopyThreads threads (one per 2 MiB).
void parallel_copy(void* dst, const void* src, size_t bytes, uint32_t sub) {
  const size_t nt = std::min<size_t>(kCopyThreads, std::max<size_t>(1, bytes >> 21));
  auto part = [&](size_t t) {
    const size_t words = bytes / 4;
    const size_t w0 = words * t / nt, w1 = words * (t + 1) / nt;
    if (sub) {
      const uint32_t* s = (const uint32_t*)src;
      uint32_t* d = (uint32_t*)dst;
      for (size_t w = w0; w < w1; ++w) d[w] = s[w] - sub;
    } else {
      std::memcpy((char*)dst + 4 * w0, (const char*)src + 4 * w0, 4 * (w1 - w0));
    }
    if (t + 1 == nt && (bytes & 3))  // a tail of bytes (sub is only used on whole words)
      std::memcpy((char*)dst + (bytes & ~(size_t)3), (const char*)src + (bytes & ~(size_t)3), bytes & 3);
  };
  CopyPool::get().run(nt, part);
}

rbg_status ensure_staging(rbg_ctx* ctx) {
  if (ctx->pin.p) return RBG_OK;
  RBG_TRY_HIP(ctx, hipHostMalloc((void**)&ctx->pin.p, 2 * kStageHalf, hipHostMallocDefault));
  ctx->pin.half = kStageHalf;
  for (auto& e : ctx->pin.done) RBG_TRY_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return RBG_OK;
}

// Device -> host through the staging halves: chunk c lands in half c & 1
// while the host drains chunk c - 1.
rbg_status d2h_chunks(rbg_ctx* ctx, void* dst, const void* src, size_t bytes, uint32_t sub) {
  hipStream_t st = ctx->stream;
  if (bytes < kStageMin) {
    RBG_TRY_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (sub)
      for (size_t w = 0; w < bytes / 4; ++w) ((uint32_t*)dst)[w] -= sub;
    return RBG_OK;
  }
  RBG_TRY(ensure_staging(ctx));
  const size_t H = ctx->pin.half, nch = (bytes + H - 1) / H;
  for (size_t c = 0; c <= nch; ++c) {
    if (c < nch) {
      const size_t off = c * H, len = std::min(H, bytes - off);
      RBG_TRY_HIP(ctx, hipMemcpyAsync(ctx->pin.p + (c & 1) * H, (const char*)src + off, len,
                                      hipMemcpyDeviceToHost, st));
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->pin.done[c & 1], st));
    }
    if (c > 0) {
      const size_t pc = c - 1, off = pc * H, len = std::min(H, bytes - off);
      RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[pc & 1]));
      parallel_copy((char*)dst + off, ctx->pin.p + (pc & 1) * H, len, sub);
    }
  }
  return RBG_OK;
}

// Several device -> host copies through the staging halves at once: the
// parts are laid end to end, a half is filled with whole parts or pieces of
// them, and the host drains half h - 1 while half h is in flight.  Each part's
// length is a multiple of 4 bytes.
struct D2HPart {
  void* dst;
  const void* src;
  size_t bytes;
  uint32_t sub;  // subtracted from every 32-bit word when nonzero
};
rbg_status d2h_parts(rbg_ctx* ctx, const D2HPart* parts, size_t np) {
  hipStream_t st = ctx->stream;
  RBG_TRY(ensure_staging(ctx));
  const size_t H = ctx->pin.half;
  struct Piece {
    size_t part, off, len, at;  // bytes [off, off + len) of the part, at `at` in the half
  };
  std::vector<std::vector<Piece>> halves(1);
  size_t used = 0;
  for (size_t k = 0; k < np; ++k)
    for (size_t off = 0; off < parts[k].bytes;) {
      if (used == H) {
        halves.emplace_back();
        used = 0;
      }
      const size_t len = std::min(parts[k].bytes - off, H - used);
      halves.back().push_back(Piece{k, off, len, used});
      used += len;
      off += len;
    }
  for (size_t h = 0; h <= halves.size(); ++h) {
    if (h < halves.size()) {
      for (const Piece& pc : halves[h])
        RBG_TRY_HIP(ctx, hipMemcpyAsync(ctx->pin.p + (h & 1) * H + pc.at, (const char*)parts[pc.part].src + pc.off,
                                        pc.len, hipMemcpyDeviceToHost, st));
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->pin.done[h & 1], st));
    }
    if (h > 0) {
      RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[(h - 1) & 1]));
      for (const Piece& pc : halves[h - 1])
        parallel_copy((char*)parts[pc.part].dst + pc.off, ctx->pin.p + ((h - 1) & 1) * H + pc.at, pc.len,
                      parts[pc.part].sub);
    }
  }
  return RBG_OK;
}

}  // namespace

// Several host -> device copies through the staging halves at once (parts
// laid end to end as in d2h_parts): the host fills half h while half h - 1 is
// in flight.  Returns once the staging is free again.
rbg_status h2d_parts(rbg_ctx* ctx, const H2DPart* parts, size_t np) {
  hipStream_t st = ctx->stream;
  RBG_TRY(ensure_staging(ctx));
  const size_t H = ctx->pin.half;
  size_t used = 0, h = 0;
  bool any = false;
  auto flush_half = [&]() -> rbg_status {  // the half's pieces are queued; mark it
    RBG_TRY_HIP(ctx, hipEventRecord(ctx->pin.done[h & 1], st));
    ++h;
    used = 0;
    return RBG_OK;
  };
  for (size_t k = 0; k < np; ++k)
    for (size_t off = 0; off < parts[k].bytes;) {
      if (used == 0 && h >= 2) RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[h & 1]));  // its last DMA
      const size_t len = std::min(parts[k].bytes - off, H - used);
      char* stage = ctx->pin.p + (h & 1) * H + used;
      parallel_copy(stage, (const char*)parts[k].src + off, len, 0);
      RBG_TRY_HIP(ctx, hipMemcpyAsync((char*)parts[k].dst + off, stage, len, hipMemcpyHostToDevice, st));
      any = true;
      used += len;
      off += len;
      if (used == H) RBG_TRY(flush_half());
    }
  if (used) RBG_TRY(flush_half());
  if (any) {
    RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[0]));
    RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[1]));
  }
  return RBG_OK;
}


void* alloc_host_result(size_t bytes) {
  // Large arrays: 2 MiB aligned with transparent huge pages requested, so the
  // first write faults in few pages.  Released with free().
  if (bytes < (8u << 20)) return std::malloc(bytes);
  void* p = nullptr;
  const size_t huge = 2u << 20;
  if (posix_memalign(&p, huge, (bytes + huge - 1) / huge * huge) != 0) return nullptr;
  (void)madvise(p, (bytes + huge - 1) / huge * huge, MADV_HUGEPAGE);
  return p;
}

rbg_status h2d_staged(rbg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t st = ctx->stream;
  if (bytes < kStageMin) {
    RBG_TRY_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return RBG_OK;
  }
  RBG_TRY(ensure_staging(ctx));
  const size_t H = ctx->pin.half;
  for (size_t off = 0, c = 0; off < bytes; off += H, ++c) {
    const size_t len = std::min(H, bytes - off);
    if (c >= 2) RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[c & 1]));  // the half's last DMA
    parallel_copy(ctx->pin.p + (c & 1) * H, (const char*)src + off, len, 0);
    RBG_TRY_HIP(ctx, hipMemcpyAsync((char*)dst + off, ctx->pin.p + (c & 1) * H, len,
                                    hipMemcpyHostToDevice, st));
    RBG_TRY_HIP(ctx, hipEventRecord(ctx->pin.done[c & 1], st));
  }
  // the staging halves may still be read by the DMA: the next staged copy
  // waits on their events, and callers synchronise before returning results
  RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[0]));
  RBG_TRY_HIP(ctx, hipEventSynchronize(ctx->pin.done[1]));
  return RBG_OK;
}

rbg_status d2h_staged(rbg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return d2h_chunks(ctx, dst, src, bytes, 0);
}

rbg_status d2h_u32_staged(rbg_ctx* ctx, uint32_t* dst, const uint32_t* src, size_t count, uint32_t sub) {
  return d2h_chunks(ctx, dst, src, count * 4, sub);
}

void release_staging(rbg_ctx* ctx) {
  if (ctx->pin.p) (void)hipHostFree(ctx->pin.p);
  for (auto& e : ctx->pin.done)
    if (e) (void)hipEventDestroy(e);
  ctx->pin = rbg_ctx::Staging{};
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

}  // namespace rbg

size_t rbg_asm::device_bytes() const {
  const DevBuf* all[] = {&seqs, &offs, &sk_rows, &asm_of, &unit_pad, &raw_pad, &colnorm, &anynz,
                         &q8, &perm, &pi, &pj, &pcnt, &merge, &desc, &bscal, &svd, &svf, &level,
                         &cal_rnd, &cal_members, &cal_roff, &cal_hist, &dense_items, &gate_items, &wx, &planes,
                         &mat, &cand, &counters, &deg, &slot, &off, &list, &topk_id, &topk_cnt, &selq,
                         &topk_len, &rowmax, &fence, &keep, &row_edges, &row_off, &efrom, &eto,
                         &ecnt, &escomp, &escore, &scan_tmp, &bounds, &near_e, &near_w, &cls_in};
  size_t s = 0;
  for (const DevBuf* d : all) s += d->bytes;
  return s;
}

void rbg_asm::release_buffers() {
  drop_graph();
  DevBuf* all[] = {&seqs, &offs, &sk_rows, &asm_of, &unit_pad, &raw_pad, &colnorm, &anynz,
                   &q8, &perm, &pi, &pj, &pcnt, &merge, &desc, &bscal, &svd, &svf, &level,
                   &cal_rnd, &cal_members, &cal_roff, &cal_hist, &dense_items, &gate_items, &wx, &planes,
                   &mat, &cand, &counters, &deg, &slot, &off, &list, &topk_id, &topk_cnt, &selq,
                   &topk_len, &rowmax, &fence, &keep, &row_edges, &row_off, &efrom, &eto,
                   &ecnt, &escomp, &escore, &scan_tmp, &bounds, &near_e, &near_w, &cls_in};
  for (DevBuf* d : all) d->release();
  planned = false;
  built = false;
}

namespace rbg {

namespace {

// Candidate capacity of a fused assembly whose pass covers `pairs` pairs.
uint64_t cand_capacity(uint64_t pairs, uint32_t n) {
  return std::min<uint64_t>(pairs, std::max<uint64_t>(1ull << 26, 2048ull * n));
}

// Pairs of an assembly's fused pass: all pairs, or the pairs with a new row.
uint64_t pass_pairs(uint32_t n, uint32_t n_old, bool incr) {
  const uint64_t all = (uint64_t)n * (n - 1) / 2;
  if (!incr) return all;
  return all - (uint64_t)n_old * (n_old - 1) / 2;
}

}  // namespace

rbg_status batch_upload(rbg_ctx* ctx, rbg_asm* b, const AsmInput* in, uint32_t count, double min_cv) {
  hipStream_t st = ctx->stream;
  b->planned = false;
  b->built = false;
  b->drop_graph();
  b->asms.assign(count, AsmHost{});
  uint64_t rows = 0, bytes = 0, pairs = 0, ncol = 0, nsk = 0;
  uint32_t nc_max = 1, kc_max = 1;
  b->has_src = false;
  b->n_incr = 0;
  b->list_extra = 0;
  b->cls_rows = 0;
  for (uint32_t a = 0; a < count; ++a) {
    const rbg_assembly_in& x = in[a].in;
    const Resident* src = in[a].src;
    const SketchInput* sk = in[a].sk;
    const uint32_t n_old = src ? src->n : 0;
    if (sk && (count != 1 || src || !sk->classes || (x.S > 1 && (!sk->unit || !sk->raw))))
      return ctx->fail(RBG_ERR_ARG, "sketch input: one assembly with class bytes (and unit/raw rows when S > 1)");
    if (!sk && ((x.n > 0 && (!x.seqs || !x.offsets)) || (x.S > 0 && x.n > 0 && !x.depth)))
      return ctx->fail(RBG_ERR_ARG, "assembly " + std::to_string(a) + ": missing input");
    if (n_old + (uint64_t)x.n == 0)
      return ctx->fail(RBG_ERR_ARG, "assembly " + std::to_string(a) + ": no contigs");
    if (src && src->S != x.S)
      return ctx->fail(RBG_ERR_ARG, "assembly " + std::to_string(a) + ": added contigs need the session's "
                                    "sample count");
    if (x.S > 64)
      return ctx->fail(RBG_ERR_UNSUPPORTED, "graph builds support at most 64 depth samples");
    AsmHost& h = b->asms[a];
    h.n = n_old + x.n;
    h.n_old = n_old;
    h.src = src;
    h.S = x.S;
    h.row0 = (uint32_t)rows;
    h.colnorm0 = (uint32_t)ncol;
    if (sk) h.b0_in = sk->b0;
    rows += h.n;
    ncol += x.S;
    if (!sk) {
      // the sketch addresses a workgroup's four contigs with 32-bit offsets
      for (uint32_t r = 0; r < x.n; ++r)
        if (x.offsets[r + 1] - x.offsets[r] >= (1ull << 29))
          return ctx->fail(RBG_ERR_UNSUPPORTED, "assembly " + std::to_string(a) + ": contig " +
                                                    std::to_string(r) + " is 512 MiB or longer (or offsets decrease)");
      nsk += x.n;
      if (x.n) bytes += x.offsets[x.n] - x.offsets[0];
    }
    if (x.S > 1) {
      nc_max = std::max(nc_max, nc_for(x.S));
      kc_max = std::max(kc_max, kpad_for(x.S) / 32);
    }
    if (src) b->has_src = true;
  }
  if (rows >= (1u << 24)) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^24 contigs in a batch");
  b->nrows = (uint32_t)rows;
  b->n_sketch = (uint32_t)nsk;
  b->NC = nc_max;
  b->kpad = 32 * kc_max;
  const uint32_t NC = nc_max, kpad = b->kpad;
  // The sketch kernel reads whole 16-byte chunks: keep 64 readable bytes past
  // the last contig (their contents are never counted).
  RBG_TRY_HIP(ctx, b->seqs.ensure(bytes + 64));
  RBG_TRY_HIP(ctx, b->offs.ensure((nsk + 1) * 8));
  RBG_TRY_HIP(ctx, b->sk_rows.ensure(std::max<uint64_t>(nsk, 1) * 4));
  RBG_TRY_HIP(ctx, b->asm_of.ensure(rows * 2));
  RBG_TRY_HIP(ctx, b->anynz.ensure(rows));
  RBG_TRY_HIP(ctx, b->unit_pad.ensure(rows * 16 * NC * 4));
  RBG_TRY_HIP(ctx, b->raw_pad.ensure(rows * 16 * NC * 4));
  RBG_TRY_HIP(ctx, b->colnorm.ensure(std::max<uint64_t>(ncol, 1) * 8));
  RBG_TRY_HIP(ctx, b->q8.ensure(rows * kpad));
  RBG_TRY_HIP(ctx, b->perm.ensure(rows * 4));
  std::vector<uint64_t> offs(nsk + 1);
  std::vector<uint32_t> sk_rows(nsk);
  std::vector<uint16_t> asm_of(rows);
  std::vector<uint8_t> anynz(rows, 1);  // no depth: is_nz always holds
  std::vector<double> colnorm(std::max<uint64_t>(ncol, 1), 0.0);
  std::vector<uint32_t> perm(rows), pi, pj;
  // Padded rows start as zeros; rows from the host are staged per assembly and
  // copied as one range, rows from a session are copied on the device.
  RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->unit_pad.p, rows * 16 * NC, 0u, st));
  RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->raw_pad.p, rows * 16 * NC, 0u, st));
  RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->q8.p, rows * kpad / 4, 0u, st));
  // Host preparation of each assembly's new rows (depth features, column
  // norms, pair-pass order, int8 rows, b0 pair draws), assemblies in parallel;
  // then the copies, in order.
  // Host rows: 2 MiB aligned with transparent huge pages requested
  // (alloc_host_result), first touched by the threads that fill them.
  struct HostRows {
    float* p = nullptr;
    size_t count = 0;
    HostRows() = default;
    HostRows(const HostRows&) = delete;
    HostRows& operator=(const HostRows&) = delete;
    ~HostRows() { std::free(p); }
    void alloc(size_t c) {
      p = (float*)alloc_host_result(c * sizeof(float));
      if (!p && c) throw std::bad_alloc();
      count = c;
    }
  };
  struct Prep {
    HostRows upad, rpad;            // new rows, padded to 16 * NC
    std::vector<int8_t> q8;         // new rows in pair-pass order, kpad bytes each
    std::vector<uint32_t> ai, aj;   // b0 pair draws (local rows)
  };
  std::vector<Prep> prep(count);
  auto prepare = [&](uint32_t a, bool parallel) {
    const rbg_assembly_in& x = in[a].in;
    AsmHost& h = b->asms[a];
    Prep& pr = prep[a];
    const Resident* src = h.src;
    const SketchInput* sk = in[a].sk;
    const uint32_t n_old = h.n_old, n_new = x.n, S = x.S;
    const size_t P = 16 * NC;  // padded row
    // depth features of the new rows (row-local) written straight into the
    // padded rows, colnorm from running column sums (the same additions in
    // the same order as over all rows at once)
    h.anynz_local.assign(h.n, 1);
    if (src) std::copy(src->anynz.begin(), src->anynz.end(), h.anynz_local.begin());
    const bool rows = S > 1 && n_new;
    if (rows) {
      pr.upad.alloc((size_t)n_new * P);
      pr.rpad.alloc((size_t)n_new * P);
    }
    if (sk) {  // the caller's rows (g_depth_unit, g_depth_raw, g_depth_colnorm, g_anynz)
      if (sk->anynz) std::copy(sk->anynz, sk->anynz + n_new, h.anynz_local.begin());
      if (S > 1) {
#pragma omp parallel for schedule(static) if (parallel)
        for (uint32_t r = 0; r < n_new; ++r) {
          float* u = pr.upad.p + (size_t)r * P;
          float* w = pr.rpad.p + (size_t)r * P;
          std::copy(sk->unit + (size_t)r * S, sk->unit + (size_t)(r + 1) * S, u);
          std::copy(sk->raw + (size_t)r * S, sk->raw + (size_t)(r + 1) * S, w);
          std::fill(u + S, u + P, 0.0f);
          std::fill(w + S, w + P, 0.0f);
        }
        for (uint32_t c = 0; c < S; ++c) colnorm[h.colnorm0 + c] = sk->colnorm ? sk->colnorm[c] : 1.0;
      }
    } else if (S > 0) {
      depth_features(x.depth, n_new, S, min_cv, nullptr, rows ? pr.upad.p : nullptr, rows ? pr.rpad.p : nullptr,
                     nullptr, h.anynz_local.data() + n_old, parallel, P);
      h.colsum.assign(S, 0.0);
      if (src) h.colsum = src->colsum;
      column_sums(x.depth, n_new, S, h.colsum.data());
      for (uint32_t c = 0; c < S; ++c) {
        const double mean = h.colsum[c] / (double)std::max<uint32_t>(h.n, 1);
        colnorm[h.colnorm0 + c] = S > 1 ? ((mean > 1e-12) ? 1.0 / mean : 0.0) : 0.0;
      }
    }
    // pair-pass order and int8 rows.  The pass visits the assembly's rows in
    // abundance order (pair_pass_order): a session keeps its earlier rows'
    // order and appends the new rows in their own order.
    h.perm_local.resize(h.n);
    if (src) std::copy(src->perm.begin(), src->perm.end(), h.perm_local.begin());
    if (rows) {
      std::vector<uint32_t> lp;
      pair_pass_order(pr.upad.p, n_new, S, lp, parallel, P);
      for (uint32_t r = 0; r < n_new; ++r) h.perm_local[n_old + r] = n_old + lp[r];
      if (src) {
        h.gq = src->gq;
        gate_quantize_more(pr.upad.p, n_new, S, h.gq, parallel, lp.data(), P);
      } else {
        gate_quantize(pr.upad.p, n_new, S, h.gq, parallel, lp.data(), P);
      }
      if (h.gq.kpad == kpad) {
        pr.q8.swap(h.gq.q);
      } else {
        pr.q8.assign((size_t)n_new * kpad, 0);
#pragma omp parallel for schedule(static) if (parallel)
        for (uint32_t r = 0; r < n_new; ++r)
          std::memcpy(&pr.q8[(size_t)r * kpad], &h.gq.q[(size_t)r * h.gq.kpad], h.gq.kpad);
      }
      std::vector<int8_t>().swap(h.gq.q);
    } else {
      for (uint32_t r = 0; r < n_new; ++r) h.perm_local[n_old + r] = n_old + r;
      if (src) h.gq = src->gq;
    }
    if (!(sk && sk->b0 == sk->b0)) baseline_pairs(h.n, pr.ai, pr.aj);  // b0 pair draws (they depend on n only)
  };
  // Every assembly's new contigs in one pass through the staging, on a
  // thread of its own while this one prepares the rows (the preparation makes
  // no HIP calls).
  rbg_status seq_status = RBG_OK;
  std::thread seq_upload([&] {
    std::vector<H2DPart> seq_parts;
    uint64_t at = 0;
    for (uint32_t a = 0; a < count; ++a) {
      const rbg_assembly_in& x = in[a].in;
      const uint64_t nb = (x.n && !in[a].sk) ? x.offsets[x.n] - x.offsets[0] : 0;
      if (nb && in[a].seqs_on_device) {
        const hipError_t e = hipMemcpyAsync((char*)b->seqs.p + at, x.seqs + x.offsets[0], nb,
                                            hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) {
          seq_status = ctx->fail(RBG_ERR_HIP, std::string("contig copy: ") + hipGetErrorString(e));
          return;
        }
      } else if (nb) {
        seq_parts.push_back(H2DPart{(char*)b->seqs.p + at, x.seqs + x.offsets[0], nb});
      }
      at += nb;
    }
    if (!seq_parts.empty()) seq_status = h2d_parts(ctx, seq_parts.data(), seq_parts.size());
  });
  try {
    if (count == 1) {
      prepare(0, true);
    } else {
      std::atomic<uint32_t> next{0};
      auto work = [&] {
        for (uint32_t a; (a = next.fetch_add(1)) < count;) prepare(a, false);
      };
      const uint32_t nt = std::min<uint32_t>(count, std::max(1u, std::thread::hardware_concurrency()));
      std::vector<std::thread> pool;
      for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(work);
      work();
      for (auto& th : pool) th.join();
    }
  } catch (...) {
    seq_upload.join();  // never leave the thread running
    throw;
  }
  seq_upload.join();
  RBG_TRY(seq_status);
  uint64_t seq0 = 0, sk = 0;
  for (uint32_t a = 0; a < count; ++a) {
    const rbg_assembly_in& x = in[a].in;
    AsmHost& h = b->asms[a];
    Prep& pr = prep[a];
    const Resident* src = h.src;
    const uint32_t n_old = h.n_old, S = x.S;
    const uint32_t n_new = in[a].sk ? 0u : x.n;  // contigs to sketch
    const uint64_t nb = n_new ? x.offsets[n_new] - x.offsets[0] : 0;
    for (uint32_t r = 0; r < n_new; ++r) {
      offs[sk + r] = seq0 + (x.offsets[r] - x.offsets[0]);
      sk_rows[sk + r] = h.row0 + n_old + r;
    }
    if (in[a].sk) {  // class bytes of the caller's sketches, bit-sliced at build time
      const SketchInput& si = *in[a].sk;
      RBG_TRY_HIP(ctx, b->cls_in.ensure((size_t)x.n * si.mpad));
      RBG_TRY_HIP(ctx, hipMemcpyAsync(b->cls_in.p, si.classes, (size_t)x.n * si.mpad, hipMemcpyHostToDevice, st));
      b->cls_rows = x.n;
      b->cls_mpad = si.mpad;
    }
    sk += n_new;
    seq0 += nb;
    offs[sk] = seq0;
    for (uint32_t r = 0; r < h.n; ++r) asm_of[h.row0 + r] = (uint16_t)a;
    std::copy(h.anynz_local.begin(), h.anynz_local.end(), anynz.begin() + h.row0);
    h.n_zero = (uint32_t)std::count(h.anynz_local.begin(), h.anynz_local.end(), (uint8_t)0);
    if (pr.upad.count) {
      const size_t o = (size_t)(h.row0 + n_old) * 16 * NC * 4;
      RBG_TRY_HIP(ctx, hipMemcpyAsync((char*)b->unit_pad.p + o, pr.upad.p, pr.upad.count * 4,
                                      hipMemcpyHostToDevice, st));
      RBG_TRY_HIP(ctx, hipMemcpyAsync((char*)b->raw_pad.p + o, pr.rpad.p, pr.rpad.count * 4,
                                      hipMemcpyHostToDevice, st));
    }
    if (!pr.q8.empty())
      RBG_TRY_HIP(ctx, hipMemcpyAsync((int8_t*)b->q8.p + (size_t)(h.row0 + n_old) * kpad, pr.q8.data(),
                                      pr.q8.size(), hipMemcpyHostToDevice, st));
    if (src && S > 1 && n_old) {  // the session's rows, copied on the device
      const size_t sp = 16 * nc_for(S) * 4, dp = 16 * NC * 4, kp = kpad_for(S);
      RBG_TRY_HIP(ctx, hipMemcpy2DAsync((char*)b->unit_pad.p + (size_t)h.row0 * dp, dp, src->unit_pad.p, sp,
                                        sp, n_old, hipMemcpyDeviceToDevice, st));
      RBG_TRY_HIP(ctx, hipMemcpy2DAsync((char*)b->raw_pad.p + (size_t)h.row0 * dp, dp, src->raw_pad.p, sp,
                                        sp, n_old, hipMemcpyDeviceToDevice, st));
      RBG_TRY_HIP(ctx, hipMemcpy2DAsync((int8_t*)b->q8.p + (size_t)h.row0 * kpad, kpad, src->q8.p, kp, kp,
                                        n_old, hipMemcpyDeviceToDevice, st));
    }
    for (uint32_t r = 0; r < h.n; ++r) perm[h.row0 + r] = h.row0 + h.perm_local[r];
    h.pair0 = (uint32_t)pi.size();
    h.np = (uint32_t)pr.ai.size();
    for (size_t t = 0; t < pr.ai.size(); ++t) {
      pi.push_back(h.row0 + pr.ai[t]);
      pj.push_back(h.row0 + pr.aj[t]);
    }
    pairs += pr.ai.size();
  }
  b->np_total = (uint32_t)pairs;
  RBG_TRY_HIP(ctx, b->pi.ensure(std::max<size_t>(pi.size(), 1) * 4));
  RBG_TRY_HIP(ctx, b->pj.ensure(std::max<size_t>(pj.size(), 1) * 4));
  RBG_TRY_HIP(ctx, b->pcnt.ensure(std::max<size_t>(pi.size(), 1) * 4));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->offs.p, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, st));
  if (nsk)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->sk_rows.p, sk_rows.data(), nsk * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->asm_of.p, asm_of.data(), asm_of.size() * 2, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->anynz.p, anynz.data(), anynz.size(), hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->colnorm.p, colnorm.data(), colnorm.size() * 8, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(b->perm.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, st));
  if (!pi.empty()) {
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->pi.p, pi.data(), pi.size() * 4, hipMemcpyHostToDevice, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->pj.p, pj.data(), pj.size() * 4, hipMemcpyHostToDevice, st));
  }
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));  // the host staging vectors go out of scope
  return RBG_OK;
}

rbg_status batch_plan(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P) {
  const uint32_t m = P->m, K = P->max_edges;
  if (m < 2 || m > 512) return ctx->fail(RBG_ERR_UNSUPPORTED, "sketch-m must be in [2, 512]");
  if (K == 0 || K > 1024) return ctx->fail(RBG_ERR_UNSUPPORTED, "max-edges must be in [1, 1024]");
  if (P->semantics) {
    char names[512];
    (void)rbg_env_semantics(names, sizeof names);
    return ctx->fail(RBG_ERR_UNSUPPORTED,
                     "non-default reference semantics requested (semantics=0x" + [&] {
                       char h[16];
                       std::snprintf(h, sizeof h, "%x", P->semantics);
                       return std::string(h);
                     }() + (names[0] ? std::string(", environment: ") + names : std::string()) +
                         "): this library implements the defaults only; use the reference's CPU path");
  }
  if (!(P->near_tol >= 0.0 && P->near_tol <= 1.0))
    return ctx->fail(RBG_ERR_ARG, "near_tol must be in [0, 1]");
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
  key.near_tol = P->near_tol;
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
  std::vector<MergeDesc> md;
  uint32_t max_old = 0;
  b->list_extra = 0;
  for (uint32_t a = 0; a < A; ++a) {
    AsmHost& h = b->asms[a];
    AsmDesc& d = D[a];
    const int regime = regime_for(P->regime, h.n);
    if (regime != 1 && regime != 2) return ctx->fail(RBG_ERR_ARG, "regime must be 0, 1 or 2");
    if (regime == 2 && h.n < 2500)
      return ctx->fail(RBG_ERR_ARG, "fused regime needs n >= 2500 (assembly " + std::to_string(a) + ")");
    const bool dense = dense_counts(regime, h.S, P->min_edge_weight);
    if (dense && (uint64_t)h.n * h.n * 2 > kDenseMaxBytes)
      return ctx->fail(RBG_ERR_UNSUPPORTED, "assembly " + std::to_string(a) + ": " + std::to_string(h.n) +
                                                " contigs without the abundance gate (S <= 2) need an n x n "
                                                "count matrix above 64 GiB");
    // A session whose earlier build kept its gated pairs (or its counts)
    // evaluates only the pairs with a new row.
    h.incr = !dense && h.src && h.src->regime == 2 && !h.src->dense && h.n_old > 0;
    h.dense_incr = dense && h.src && h.src->dense && h.n_old > 0 && h.src->mat.p;
    d.dense = dense ? 1u : 0u;
    d.row0 = h.row0;
    d.n = h.n;
    d.S = h.S;
    d.regime = (uint32_t)regime;
    // composition weight: auto_w_comp(S) (src/rabbitbin.cpp:470-473)
    d.w_comp = w_comp_for(h.S);
    // abundance-first gate (src/impl/rb_graph.cpp:693-706)
    const bool gate = gate_on(regime, h.S, P->min_edge_weight);
    d.thr = gate ? (float)((P->min_edge_weight - d.w_comp) / (1.0 - d.w_comp) - 1e-4)
                 : -__builtin_huge_valf();
    // prefilter cut: with the gate off every pair is a candidate
    d.qcut = gate ? gate_qcut(h.gq, d.thr) : INT32_MIN;
    d.qcut_hi = gate ? std::max(gate_qcut_hi(h.gq, d.thr), d.qcut) : INT32_MAX;
    d.acc1_end = gate ? 2 * (h.S / 32) : 0;
    d.pair0 = h.pair0;
    d.np = h.np;
    d.colnorm0 = h.colnorm0;
    d.b0_in = h.b0_in;
    if (dense) {
      d.mat0 = mat_total;
      d.jmin = h.dense_incr ? h.n_old : 0;
      mat_total += (uint64_t)h.n * h.n;
    } else {
      const uint32_t jmin = h.incr ? h.n_old : 0;
      gate_work_items(a, h.n, jmin, b->sched.gate_tiles_per_item, gfull, gpart);
      cap += cand_capacity(pass_pairs(h.n, h.n_old, h.incr), h.n);
      if (h.incr) {
        md.push_back(MergeDesc{(const uint64_t*)h.src->loff.p, (const uint64_t*)h.src->list.p, h.row0,
                               h.n_old});
        max_old = std::max(max_old, h.n_old);
        b->list_extra += h.src->nlist;
      }
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
  b->n_incr = (uint32_t)md.size();
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
  RBG_TRY_HIP(ctx, b->cal_hist.ensure((size_t)A * kCalibHistWords * 4));
  b->cal_max_items = 0;
  for (uint32_t a = 0; a < A; ++a)
    b->cal_max_items = std::max(b->cal_max_items, D[a].regime == 1 ? D[a].n : roff[(size_t)a * 11 + 10] - roff[(size_t)a * 11]);
  RBG_TRY_HIP(ctx, b->merge.ensure(std::max<size_t>(md.size(), 1) * sizeof(MergeDesc)));
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
  if (!md.empty())
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->merge.p, md.data(), md.size() * sizeof(MergeDesc),
                                    hipMemcpyHostToDevice, st));
  b->merge_max_old = max_old;
  const uint32_t FW = fence_words(K), KW = (K + 63) / 64;
  RBG_TRY_HIP(ctx, b->planes.ensure((size_t)N * 8 * W * 4));
  RBG_TRY_HIP(ctx, b->bscal.ensure(A * sizeof(BuildScalars)));
  RBG_TRY_HIP(ctx, b->svd.ensure((size_t)A * (m + 1) * 8));
  RBG_TRY_HIP(ctx, b->svf.ensure((size_t)A * (m + 1) * 4));
  RBG_TRY_HIP(ctx, b->level.ensure((size_t)A * (m + 1) * 2));
  // +16 bytes: the selection reads dense rows in whole 16-byte chunks
  RBG_TRY_HIP(ctx, b->mat.ensure(std::max<uint64_t>(mat_total, 1) * 2 + 16));
  RBG_TRY_HIP(ctx, b->cand.ensure(std::max<uint64_t>(cap, 1) * 8));
  RBG_TRY_HIP(ctx, b->counters.ensure(32));  // candidates, near-cut edges (u64 each)
  RBG_TRY_HIP(ctx, b->deg.ensure((size_t)N * 4));
  RBG_TRY_HIP(ctx, b->slot.ensure(std::max<uint64_t>(cap, 1) * 8));  // 2 list slots per candidate
  RBG_TRY_HIP(ctx, b->off.ensure((size_t)(N + 1) * 8));
  RBG_TRY_HIP(ctx, b->list.ensure(std::max<uint64_t>(2 * cap + b->list_extra, 1) * 8));
  // +16 ids: the mutual test reads whole 16-id segments of a row's list
  RBG_TRY_HIP(ctx, b->topk_id.ensure(((size_t)N * K + 16) * 4));
  RBG_TRY_HIP(ctx, b->topk_cnt.ensure((size_t)N * K * 2));
  RBG_TRY_HIP(ctx, b->topk_len.ensure((size_t)N * 4));
  // rows for the workgroup selection: the dense assemblies' rows, then the
  // long candidate lists the wave selection queues; two counters
  RBG_TRY_HIP(ctx, b->selq.ensure((size_t)(N + 2) * 4));
  {
    std::vector<uint32_t> srows;
    for (uint32_t a = 0; a < A; ++a)
      if (D[a].dense)
        for (uint32_t r = 0; r < D[a].n; ++r) srows.push_back(D[a].row0 + r);
    b->n_sel_static = (uint32_t)srows.size();
    if (!srows.empty())
      RBG_TRY_HIP(ctx, hipMemcpyAsync(b->selq.p, srows.data(), srows.size() * 4, hipMemcpyHostToDevice, st));
  }
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
  b->near_cap = std::min<uint64_t>(ce, 1u << 20);
  RBG_TRY_HIP(ctx, b->near_e.ensure(b->near_cap * 8));
  RBG_TRY_HIP(ctx, b->near_w.ensure(b->near_cap * 8));
  size_t tmp = 0;
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(nullptr, nullptr, N, nullptr, tmp, st));
  RBG_TRY_HIP(ctx, b->scan_tmp.ensure(tmp));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  b->key = key;
  b->planned = true;
  return RBG_OK;
}

size_t estimate_build_bytes(const AsmInput* in, uint32_t count, const rbg_graph_params* P) {
  const uint32_t m = P->m, K = P->max_edges, W = words_for(m);
  uint64_t N = 0, bytes = 0, mat = 0, cap = 0, extra = 0, nsk = 0;
  uint32_t NC = 1, kpad = 32;
  for (uint32_t a = 0; a < count; ++a) {
    const rbg_assembly_in& x = in[a].in;
    const Resident* src = in[a].src;
    const uint32_t n_old = src ? src->n : 0, n = n_old + x.n;
    N += n;
    nsk += x.n;
    if (x.n) bytes += x.offsets[x.n] - x.offsets[0];
    NC = std::max(NC, nc_for(x.S));
    kpad = std::max(kpad, kpad_for(x.S));
    const int regime = regime_for(P->regime, n);
    if (dense_counts(regime, x.S, P->min_edge_weight)) {
      mat += (uint64_t)n * n;
    } else {
      const bool incr = src && src->regime == 2 && !src->dense && n_old > 0;
      cap += cand_capacity(pass_pairs(n, n_old, incr), n);
      if (incr) extra += src->nlist;
    }
  }
  const uint64_t FW = fence_words(K), KW = (K + 63) / 64;
  uint64_t s = bytes + 64 + (nsk + 1) * 12;                 // seqs, offs, sk_rows
  s += N * (2 + 1 + 2 * 16 * NC * 4 + kpad + 4);            // asm_of, anynz, unit/raw, q8, perm
  s += N * 8 * W * 4;                                       // planes
  s += mat * 2 + cap * 16 + (2 * cap + extra) * 8;          // mat, cand + slots, list
  s += N * (4 + 8 + 4 + (uint64_t)K * 6 + 4 + 4 + FW * 4 + KW * 8 + 4 + 8);  // per-row arrays
  s += N * (uint64_t)K * (4 + 4 + 2 + 4 + 4);               // edges
  s += (uint64_t)count * (64 * 4 + (m + 1) * 14 + kCalibRandDev * 4 + kCalibHistWords * 4 + 30000 * 12);
  return (size_t)s + (64u << 20);                           // small tables and slack
}

// Every launch of one build, on ctx->stream, without a host round trip (so it
// can be captured).  `timed` records the stage events of the eager path.
static rbg_status enqueue_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P, bool want_wx,
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
  // events around single launches (eager builds): kbeg(name) ... kend()
  ctx->nk = 0;
  auto kbeg = [&](const char* name) -> rbg_status {
    if (timed && ctx->nk < RBG_MAX_KERNEL_TIMES) {
      ctx->kname[ctx->nk] = name;
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->kev[2 * ctx->nk], st));
    }
    return RBG_OK;
  };
  auto kend = [&]() -> rbg_status {
    if (timed && ctx->nk < RBG_MAX_KERNEL_TIMES) {
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->kev[2 * ctx->nk + 1], st));
      ++ctx->nk;
    }
    return RBG_OK;
  };
  // the same around a launch on the side stream
  auto kside = [&](const char* name, auto&& launch) -> rbg_status {
    const bool t = timed && ctx->nk < RBG_MAX_KERNEL_TIMES;
    if (t) {
      ctx->kname[ctx->nk] = name;
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->kev[2 * ctx->nk], ctx->side));
    }
    RBG_TRY_HIP(ctx, launch());
    if (t) {
      RBG_TRY_HIP(ctx, hipEventRecord(ctx->kev[2 * ctx->nk + 1], ctx->side));
      ++ctx->nk;
    }
    return RBG_OK;
  };
  RBG_TRY(mark(0));
  RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->counters.p, 8, 0u, st));
  unsigned long long* ncand = (unsigned long long*)b->counters.p;
  // The int8 prefilter of the pair pass reads only the int8 rows, the
  // pair-pass order and the gate items, which the plan has uploaded.  In a
  // build of several assemblies it runs on the side stream beside the sketch,
  // b0 and the small assemblies' dense counts, and joins before the exact
  // gate (44.0 against 46.5 ms for 64 assemblies).  A single assembly has no
  // dense work beside it, and the sketch and the prefilter compete for the
  // same workgroup slots (together 4.5 ms, one after the other 4.6): it runs
  // on the main stream, so every kernel's time is its own.
  auto prefilter = [&](hipStream_t s) {
    return launch_gate_mfma((const int8_t*)b->q8.p, b->kpad, desc, (const uint32_t*)b->perm.p,
                            (const uint32_t*)b->gate_items.p, b->n_gate_items, (uint64_t*)b->cand.p, ncand,
                            b->cand_cap, s);
  };
  const bool side = b->has_gate_items && A > 1;
  if (side) {
    RBG_TRY_HIP(ctx, hipEventRecord(ctx->fork_ev, st));
    RBG_TRY_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
    RBG_TRY(kside("gate_mfma_kernel", [&] { return prefilter(ctx->side); }));
    RBG_TRY_HIP(ctx, hipEventRecord(ctx->join_ev, ctx->side));
  }
  const uint32_t mpad = 32 * W;
  if (want_wx) RBG_TRY_HIP(ctx, b->wx.ensure((size_t)N * mpad));
  // sketches: rows of resident sessions are copied, the others computed
  if (b->has_src)
    for (const AsmHost& h : b->asms)
      if (h.src && h.n_old)
        RBG_TRY_HIP(ctx, hipMemcpyAsync((uint32_t*)b->planes.p + (size_t)h.row0 * 8 * W, h.src->planes.p,
                                        (size_t)h.n_old * 8 * W * 4, hipMemcpyDeviceToDevice, st));
  RBG_TRY(kbeg("sketch_kernel"));
  RBG_TRY_HIP(ctx, launch_sketch((const uint8_t*)b->seqs.p, (const uint64_t*)b->offs.p, b->n_sketch, m, W,
                                 (const double*)t->sT.p, (const uint8_t*)t->sX.p,
                                 (const uint8_t*)t->cls.p, (const uint8_t*)t->codes.p,
                                 want_wx ? (uint8_t*)b->wx.p : nullptr, mpad, (uint32_t*)b->planes.p,
                                 b->has_src ? (const uint32_t*)b->sk_rows.p : nullptr, st));
  RBG_TRY(kend());
  if (b->cls_rows)
    RBG_TRY_HIP(ctx, launch_planes_from_classes((const uint8_t*)b->cls_in.p, b->cls_rows, m, b->cls_mpad, W, 0,
                                                (uint32_t*)b->planes.p, st));
  RBG_TRY(mark(1));
  // b0 and the sv tables (the pair draws depend on n only)
  if (b->np_total) {
    RBG_TRY(kbeg("list_counts_kernel"));
    RBG_TRY_HIP(ctx, launch_list_counts((const uint32_t*)b->planes.p, W, (const uint32_t*)b->pi.p,
                                        (const uint32_t*)b->pj.p, b->np_total,
                                        (uint32_t*)b->pcnt.p, st));
    RBG_TRY(kend());
  }
  RBG_TRY_HIP(ctx, launch_baseline_tables((const uint32_t*)b->pcnt.p, desc, A, m, bs,
                                          (double*)b->svd.p, (float*)b->svf.p,
                                          (uint16_t*)b->level.p, st));
  RBG_TRY(mark(2));
  if (b->has_src)  // served sessions' earlier counts, into this build's row pitch
    for (uint32_t a = 0; a < A; ++a) {
      const AsmHost& h = b->asms[a];
      if (h.dense_incr)
        RBG_TRY_HIP(ctx, hipMemcpy2DAsync((uint16_t*)b->mat.p + b->desc_h[a].mat0, (size_t)h.n * 2,
                                          h.src->mat.p, (size_t)h.n_old * 2, (size_t)h.n_old * 2,
                                          h.n_old, hipMemcpyDeviceToDevice, st));
    }
  if (b->n_dense_items) {
    RBG_TRY(kbeg("dense_counts_kernel"));
    RBG_TRY_HIP(ctx, launch_dense_counts((const uint32_t*)b->planes.p, W, desc,
                                         (const uint32_t*)b->dense_items.p, b->n_dense_items,
                                         (uint16_t*)b->mat.p, st));
    RBG_TRY(kend());
  }
  if (side) {
    RBG_TRY_HIP(ctx, hipStreamWaitEvent(st, ctx->join_ev, 0));  // the candidates are in
  } else if (b->has_gate_items) {
    RBG_TRY(kbeg("gate_mfma_kernel"));
    RBG_TRY_HIP(ctx, prefilter(st));
    RBG_TRY(kend());
  }
  RBG_TRY(mark(6));
  RBG_TRY(mark(7));
  if (b->has_gate_items) {
    // Exact FP32 gate and register compare of the candidates, in place; the
    // survivors claim their list slots.  A served session's old rows start
    // at their old list lengths, so the new entries follow the old ones.
    const MergeDesc* md = (const MergeDesc*)b->merge.p;
    RBG_TRY_HIP(ctx, launch_fill_u32((uint32_t*)b->deg.p, N, 0u, st));
    RBG_TRY_HIP(ctx, launch_merge_deg(md, b->n_incr, b->merge_max_old, (uint32_t*)b->deg.p, st));
    RBG_TRY(kbeg("gate_compare_kernel"));
    RBG_TRY_HIP(ctx, launch_gate_compare((uint64_t*)b->cand.p, b->cand_cap, ncand,
                                         (const float*)b->unit_pad.p, b->NC,
                                         (const uint8_t*)b->anynz.p, desc, asm_of, A,
                                         (const uint32_t*)b->planes.p, W, (uint32_t*)b->deg.p,
                                         (uint32_t*)b->slot.p, st));
    RBG_TRY(kend());
  }
  RBG_TRY(mark(8));
  if (b->has_gate_items) {
    const MergeDesc* md = (const MergeDesc*)b->merge.p;
    size_t tmp = b->scan_tmp.bytes;
    RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)b->deg.p, (uint64_t*)b->off.p, N,
                                               b->scan_tmp.p, tmp, st));
    RBG_TRY_HIP(ctx, launch_merge_lists(md, b->n_incr, b->merge_max_old, (const uint64_t*)b->off.p,
                                        (uint64_t*)b->list.p, st));
    RBG_TRY(kbeg("scatter_kernel"));
    RBG_TRY_HIP(ctx, launch_scatter((const uint64_t*)b->cand.p, b->cand_cap, ncand,
                                    (const uint32_t*)b->slot.p, (const uint64_t*)b->off.p,
                                    (uint64_t*)b->list.p, st));
    RBG_TRY(kend());
  }
  RBG_TRY(mark(3));
  RBG_TRY(mark(9));
  RBG_TRY(kbeg("select_wave_kernel"));  // with the workgroup-per-row pass over the rows it queues
  RBG_TRY_HIP(ctx, launch_select((const uint64_t*)b->list.p, (const uint64_t*)b->off.p,
                                 (const uint16_t*)b->mat.p, (const uint8_t*)b->anynz.p, desc, asm_of,
                                 N, (const uint16_t*)b->level.p, m, K, (uint32_t*)b->topk_id.p,
                                 (uint16_t*)b->topk_cnt.p, (uint32_t*)b->topk_len.p,
                                 (uint32_t*)b->rowmax.p, (uint32_t*)b->fence.p, FW,
                                 (uint32_t*)b->selq.p, b->n_sel_static, st));
  RBG_TRY(kend());
  RBG_TRY(mark(10));
  // calibration (the glibc rand() draws are prepared per (n, seed))
  RBG_TRY(kbeg("calibrate_kernel"));
  RBG_TRY_HIP(ctx, launch_calibrate((const uint32_t*)b->rowmax.p, desc, A, m,
                                    (const double*)b->svd.p, (const float*)b->svf.p,
                                    (const int32_t*)b->cal_rnd.p, (const uint32_t*)b->cal_members.p,
                                    (const uint32_t*)b->cal_roff.p, b->cal_max_items,
                                    (uint32_t*)b->cal_hist.p, P->coverage, bs, st));
  RBG_TRY(kend());
  RBG_TRY(kbeg("emit_mark_kernel"));
  RBG_TRY_HIP(ctx, launch_emit_mark((const uint32_t*)b->topk_id.p, (const uint16_t*)b->topk_cnt.p,
                                    (const uint32_t*)b->topk_len.p, (const uint32_t*)b->fence.p, FW,
                                    N, K, bs, desc, asm_of, (const uint32_t*)b->perm.p,
                                    (uint64_t*)b->keep.p, (uint32_t*)b->row_edges.p, st));
  RBG_TRY(kend());
  {
    size_t tmp = b->scan_tmp.bytes;
    RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64((const uint32_t*)b->row_edges.p,
                                               (uint64_t*)b->row_off.p, N, b->scan_tmp.p, tmp, st));
  }
  const uint64_t* ne = (const uint64_t*)b->row_off.p + N;
  RBG_TRY(kbeg("emit_write_kernel"));
  RBG_TRY_HIP(ctx, launch_emit_write((const uint32_t*)b->topk_id.p, (const uint16_t*)b->topk_cnt.p,
                                     (const uint32_t*)b->topk_len.p, N, K, (const uint64_t*)b->keep.p,
                                     (const uint64_t*)b->row_off.p, asm_of, m,
                                     (const float*)b->svf.p, (uint32_t*)b->efrom.p,
                                     (uint32_t*)b->eto.p, (uint16_t*)b->ecnt.p,
                                     (float*)b->escomp.p, st));
  RBG_TRY(kend());
  RBG_TRY(mark(4));
  RBG_TRY(kbeg("edge_weight_pad_kernel"));
  RBG_TRY_HIP(ctx, launch_edge_weights_batch((const uint32_t*)b->efrom.p, (const uint32_t*)b->eto.p,
                                             (const float*)b->escomp.p, b->edge_cap, ne,
                                             (const float*)b->unit_pad.p, (const float*)b->raw_pad.p,
                                             b->NC, (const double*)b->colnorm.p, desc, asm_of,
                                             P->min_edge_weight, P->neg_depth_thr, P->near_tol,
                                             (float*)b->escore.p, ncand + 1, b->near_cap,
                                             (uint64_t*)b->near_e.p, (double*)b->near_w.p, st));
  RBG_TRY(kend());
  RBG_TRY_HIP(ctx, launch_gather_bounds((const uint64_t*)b->row_off.p,
                                        b->has_gate_items ? (const uint64_t*)b->off.p : nullptr, desc,
                                        A, (uint64_t*)b->bounds.p, st));
  RBG_TRY(mark(5));
  return RBG_OK;
}

rbg_status run_build(rbg_ctx* ctx, rbg_asm* b, const rbg_graph_params* P, bool want_wx,
                     bool capture, bool* replayed) {
  hipStream_t st = ctx->stream;
  const uint32_t A = (uint32_t)b->asms.size();
  if (replayed) *replayed = false;
  for (int attempt = 0; attempt < 2; ++attempt) {
    RBG_TRY(batch_plan(ctx, b, P));
    if (capture && !want_wx && !b->has_src) {
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
    unsigned long long cnt[2] = {0, 0};  // candidates, near-cut edges
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->bs_h.data(), b->bscal.p, A * sizeof(BuildScalars),
                                    hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(b->bounds_h.data(), b->bounds.p, (size_t)A * 32,
                                    hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(cnt, b->counters.p, 16, hipMemcpyDeviceToHost, st));
    RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
    const uint64_t nc = b->has_gate_items ? cnt[0] : 0;
    b->ncand_last = nc;
    if (nc <= b->cand_cap) {
      if (cnt[1] > b->near_cap)
        return ctx->fail(RBG_ERR_NOMEM, "more near-cut edges than the list holds (" +
                                            std::to_string(cnt[1]) + "): lower near_tol");
      b->near_e_h.resize(cnt[1]);
      b->near_w_h.resize(cnt[1]);
      if (cnt[1]) {
        RBG_TRY_HIP(ctx, hipMemcpyAsync(b->near_e_h.data(), b->near_e.p, cnt[1] * 8, hipMemcpyDeviceToHost, st));
        RBG_TRY_HIP(ctx, hipMemcpyAsync(b->near_w_h.data(), b->near_w.p, cnt[1] * 8, hipMemcpyDeviceToHost, st));
        RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
      }
      b->built = true;
      return RBG_OK;
    }
    b->cand_cap_min = nc;  // rebuild with room for every candidate
    b->built = false;
  }
  return ctx->fail(RBG_ERR_NOMEM, "candidate list overflow after a resize");
}

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
  if (b->asms.size() == 1) out->candidates = b->ncand_last;
  if (b->desc_h[a].dense && out->regime == 2)  // no gate: every pair passing is_nz was evaluated
    out->gated_pairs = (uint64_t)h.n * (h.n - 1) / 2 - (uint64_t)h.n_zero * (h.n_zero - (h.n_zero > 0)) / 2;
  const uint32_t K = P->max_edges, m = P->m, n = h.n, r0 = h.row0;
  {  // near-cut edges of this assembly, by edge index
    std::vector<std::pair<uint64_t, double>> nr;
    for (size_t t = 0; t < b->near_e_h.size(); ++t)
      if (b->near_e_h[t] >= e0 && b->near_e_h[t] < e1) nr.emplace_back(b->near_e_h[t] - e0, b->near_w_h[t]);
    std::sort(nr.begin(), nr.end());
    out->n_near = nr.size();
    if (!nr.empty()) {
      out->near_edge = (uint64_t*)std::malloc(nr.size() * 8);
      out->near_score64 = (double*)std::malloc(nr.size() * 8);
      if (!out->near_edge || !out->near_score64) return ctx->fail(RBG_ERR_NOMEM, "host allocation of the near list");
      for (size_t t = 0; t < nr.size(); ++t) {
        out->near_edge[t] = nr[t].first;
        out->near_score64[t] = nr[t].second;
      }
    }
  }
  if (!on_device) {
    const size_t nb = std::max<uint64_t>(ne, 1);
    out->from = (uint32_t*)alloc_host_result(nb * 4);
    out->to = (uint32_t*)alloc_host_result(nb * 4);
    out->scomp = (float*)alloc_host_result(nb * 4);
    out->score = h.S > 0 ? (float*)alloc_host_result(nb * 4) : nullptr;
    if (!out->from || !out->to || !out->scomp || (h.S > 0 && !out->score))
      return ctx->fail(RBG_ERR_NOMEM, "host allocation of the edge list");
    // global rows -> ids local to the assembly on the way
    const D2HPart parts[4] = {{out->from, (const uint32_t*)b->efrom.p + e0, ne * 4, r0},
                              {out->to, (const uint32_t*)b->eto.p + e0, ne * 4, r0},
                              {out->scomp, (const float*)b->escomp.p + e0, ne * 4, 0},
                              {out->score, (const float*)b->escore.p + e0, ne * 4, 0}};
    if (ne) RBG_TRY(d2h_parts(ctx, parts, h.S > 0 ? 4 : 3));
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

rbg_status save_resident(rbg_ctx* ctx, rbg_asm* b, uint32_t a, const rbg_graph_params* P,
                         Resident* r) {
  if (!b->built || a >= b->asms.size()) return ctx->fail(RBG_ERR_ARG, "no build to keep");
  hipStream_t st = ctx->stream;
  AsmHost& h = b->asms[a];
  const uint32_t n = h.n, S = h.S, W = words_for(P->m);
  // The build is complete, so r's buffers (this build's source when the
  // session was extended) may be overwritten: every row is copied from the
  // build.  Buffers grow with headroom so most extensions reuse them.
  auto fit = [&](DevBuf& d, size_t need) -> hipError_t {
    return need <= d.bytes && d.p ? hipSuccess : d.ensure(need + need / 4);
  };
  RBG_TRY_HIP(ctx, fit(r->planes, (size_t)n * 8 * W * 4));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(r->planes.p, (uint32_t*)b->planes.p + (size_t)h.row0 * 8 * W,
                                  (size_t)n * 8 * W * 4, hipMemcpyDeviceToDevice, st));
  if (S > 1) {
    const size_t sp = 16 * nc_for(S) * 4, bp = 16 * b->NC * 4;
    RBG_TRY_HIP(ctx, fit(r->unit_pad, (size_t)n * sp));
    RBG_TRY_HIP(ctx, fit(r->raw_pad, (size_t)n * sp));
    RBG_TRY_HIP(ctx, hipMemcpy2DAsync(r->unit_pad.p, sp, (char*)b->unit_pad.p + (size_t)h.row0 * bp, bp, sp,
                                      n, hipMemcpyDeviceToDevice, st));
    RBG_TRY_HIP(ctx, hipMemcpy2DAsync(r->raw_pad.p, sp, (char*)b->raw_pad.p + (size_t)h.row0 * bp, bp, sp,
                                      n, hipMemcpyDeviceToDevice, st));
    const size_t kp = kpad_for(S);
    RBG_TRY_HIP(ctx, fit(r->q8, (size_t)n * kp));
    RBG_TRY_HIP(ctx, hipMemcpy2DAsync(r->q8.p, kp, (int8_t*)b->q8.p + (size_t)h.row0 * b->kpad, b->kpad, kp,
                                      n, hipMemcpyDeviceToDevice, st));
  }
  const int regime = (int)b->desc_h[a].regime;
  const bool dense = b->desc_h[a].dense != 0;
  uint64_t nlist = 0;
  if (!dense) {
    nlist = b->bounds_h[4 * a + 3] - b->bounds_h[4 * a + 2];
    RBG_TRY_HIP(ctx, fit(r->loff, ((size_t)n + 1) * 8));
    RBG_TRY_HIP(ctx, fit(r->list, std::max<uint64_t>(nlist, 1) * 8));
    RBG_TRY_HIP(ctx, launch_extract_lists((const uint64_t*)b->off.p, (const uint64_t*)b->list.p, h.row0, n,
                                          (uint64_t*)r->loff.p, (uint64_t*)r->list.p, st));
    r->mat.release();
  } else {
    r->loff.release();
    r->list.release();
    const size_t mb = (size_t)n * n * 2;
    RBG_TRY_HIP(ctx, fit(r->mat, mb));
    RBG_TRY_HIP(ctx, hipMemcpyAsync(r->mat.p, (uint16_t*)b->mat.p + b->desc_h[a].mat0, mb,
                                    hipMemcpyDeviceToDevice, st));
  }
  // No wait: every later use of these buffers (the session's next build, or
  // their release, which synchronises) comes after the copies on the stream.
  RBG_TRY_HIP(ctx, hipGetLastError());
  r->nlist = nlist;
  r->n = n;
  r->S = S;
  r->regime = regime;
  r->dense = dense;
  r->perm = std::move(h.perm_local);
  r->anynz = std::move(h.anynz_local);
  r->colsum = std::move(h.colsum);
  r->gq = h.gq;
  return RBG_OK;
}

}  // namespace rbg
